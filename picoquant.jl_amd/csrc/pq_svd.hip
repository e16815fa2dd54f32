// decompose_tensor on the GPU: the SVD split of one tensor into two
// (src/backends/interactive.jl:130-152 -> src/layer1.jl:146-190), used by
// compress_bond! / contract_mps_tensor_network_circuit! (src/layer2.jl:
// 423-703).  The tensor is permuted to (left, right) and materialised as an
// m x n column-major matrix, factorised by rocSOLVER's gesdd (economy U and
// V^H), the bond dimension chi is taken from the normalised spectrum exactly
// as the reference does, and one kernel writes B = U[:, :chi] sqrt(S) and
// C = sqrt(S) V^H[:chi, :] as new context tensors.
#include "pq_rt.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <limits>

using namespace pqrt;

namespace {

// B[i + m j] = U[i + m j] sqrt(S[j]) for j < chi; C[j + chi l] = sqrt(S[j])
// V[j + k l] (V holds V^H with leading dimension k).
template <typename C, typename R>
__global__ void k_split_scale(const C* __restrict__ U, const C* __restrict__ V, const R* __restrict__ S,
                              C* __restrict__ B, C* __restrict__ Cout, int64_t m, int64_t n, int64_t k, int chi) {
  const int64_t nb = m * chi, total = nb + (int64_t)chi * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    if (e < nb) {
      const int64_t i = e % m, j = e / m;
      const R s = sqrt(S[j]);
      C u = U[i + m * j];
      u.x *= s;
      u.y *= s;
      B[e] = u;
    } else {
      const int64_t f = e - nb, j = f % chi, l = f / chi;
      const R s = sqrt(S[j]);
      C v = V[j + k * l];
      v.x *= s;
      v.y *= s;
      Cout[f] = v;
    }
  }
}

rocblas_handle blas_of(pq_ctx* c) {
  if (!c->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    if (rocblas_set_stream(h, c->stream) != rocblas_status_success) {
      rocblas_destroy_handle(h);
      return nullptr;
    }
    c->blas = h;
  }
  return static_cast<rocblas_handle>(c->blas);
}

template <typename C, typename R>
int decompose_typed(pq_ctx* c, const Tensor& A, int64_t m, int64_t n, double threshold, int max_rank, Buf*& Bb,
                    Buf*& Cb, int& chi) {
  const int64_t k = std::min(m, n);
  rocblas_handle h = blas_of(c);
  if (!h) return fail(c, PQ_ERR_HIP, "rocBLAS handle creation failed");
  Buf* U = new_buf(c, (size_t)(m * k) * sizeof(C));
  Buf* V = new_buf(c, (size_t)(k * n) * sizeof(C));
  Buf* S = new_buf(c, (size_t)k * sizeof(R));
  Buf* I = new_buf(c, sizeof(rocblas_int));
  auto release = [&]() {
    for (Buf* b : {U, V, S, I})
      if (b) unref(c, b);
  };
  if (!U || !V || !S || !I) {
    release();
    return fail(c, PQ_ERR_OOM, "SVD workspace allocation failed");
  }
  C* a = static_cast<C*>(A.buf->ptr) + A.off;
  rocblas_status st;
  // divide and conquer: 8-27x faster than gesvd's QR iteration for
  // 128..512-sided complex matrices on this device
  if constexpr (sizeof(R) == 8)
    st = rocsolver_zgesdd(h, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)m, (rocblas_int)n,
                          reinterpret_cast<rocblas_double_complex*>(a), (rocblas_int)m, static_cast<double*>(S->ptr),
                          static_cast<rocblas_double_complex*>(U->ptr), (rocblas_int)m,
                          static_cast<rocblas_double_complex*>(V->ptr), (rocblas_int)k,
                          static_cast<rocblas_int*>(I->ptr));
  else
    st = rocsolver_cgesdd(h, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)m, (rocblas_int)n,
                          reinterpret_cast<rocblas_float_complex*>(a), (rocblas_int)m, static_cast<float*>(S->ptr),
                          static_cast<rocblas_float_complex*>(U->ptr), (rocblas_int)m,
                          static_cast<rocblas_float_complex*>(V->ptr), (rocblas_int)k,
                          static_cast<rocblas_int*>(I->ptr));
  if (st != rocblas_status_success) {
    release();
    return fail(c, PQ_ERR_HIP, "rocsolver gesdd failed (status " + std::to_string((int)st) + ")");
  }
  std::vector<R> s((size_t)k);
  rocblas_int info = 0;
  if (hipMemcpyAsync(s.data(), S->ptr, (size_t)k * sizeof(R), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipMemcpyAsync(&info, I->ptr, sizeof(info), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    release();
    return fail(c, PQ_ERR_HIP, "SVD result download failed");
  }
  if (info != 0) {
    release();
    return fail(c, PQ_ERR_HIP, "gesdd did not converge (info " + std::to_string((int)info) + ")");
  }
  // src/layer1.jl:166-175: threshold at least sqrt(eps), relative to the
  // spectrum's 2-norm; max_rank caps the count
  const double thr = std::max(threshold, std::sqrt((double)std::numeric_limits<R>::epsilon()));
  double norm = 0.0;
  for (R v : s) norm += (double)v * (double)v;
  norm = std::sqrt(norm);
  chi = 0;
  for (R v : s)
    if ((double)v / norm > thr) ++chi;
  if (max_rank > 0) chi = std::min(chi, max_rank);
  if (chi < 1) {
    release();
    return fail(c, PQ_ERR_UNSUPPORTED, "decompose of a zero tensor (no singular value above the threshold)");
  }
  Bb = new_buf(c, (size_t)(m * chi) * sizeof(C));
  Cb = new_buf(c, (size_t)(chi * n) * sizeof(C));
  if (!Bb || !Cb) {
    if (Bb) unref(c, Bb);
    if (Cb) unref(c, Cb);
    release();
    return fail(c, PQ_ERR_OOM, "device allocation failed");
  }
  const int64_t total = (m + n) * chi;
  const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL((k_split_scale<C, R>), dim3(std::max(1u, g)), dim3(256), 0, c->stream,
                     static_cast<const C*>(U->ptr), static_cast<const C*>(V->ptr), static_cast<const R*>(S->ptr),
                     static_cast<C*>(Bb->ptr), static_cast<C*>(Cb->ptr), m, n, k, chi);
  hipError_t e = hipGetLastError();
  release();
  if (e != hipSuccess) {
    unref(c, Bb);
    unref(c, Cb);
    return hipfail(c, e, "split kernel");
  }
  return PQ_OK;
}

}  // namespace

namespace pqrt {
void release_blas(pq_ctx* c) {
  if (c && c->blas) {
    rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
    c->blas = nullptr;
  }
}
}  // namespace pqrt

extern "C" {

int pq_decompose(pq_ctx* c, pq_tensor h, int nleft, const int32_t* left, int nright, const int32_t* right,
                 double threshold, int max_rank, pq_tensor* out_b, pq_tensor* out_c, int* chi_out) {
  if (!c || !out_b || !out_c || nleft < 0 || nright < 0 || (nleft && !left) || (nright && !right) || max_rank < 0)
    return PQ_ERR_ARG;
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  if (nleft + nright != t->rank || nleft + 1 > PQ_MAX_RANK || nright + 1 > PQ_MAX_RANK)
    return fail(c, PQ_ERR_ARG, "left and right positions must cover the tensor's axes");
  // positions are 1-based (Julia); together a permutation of the axes
  std::vector<int> perm;
  std::vector<char> seen(t->rank, 0);
  for (int q = 0; q < nleft + nright; ++q) {
    const int p = (q < nleft ? left[q] : right[q - nleft]) - 1;
    if (p < 0 || p >= t->rank || seen[p]) return fail(c, PQ_ERR_ARG, "bad decompose positions");
    seen[p] = 1;
    perm.push_back(p);
  }
  Tensor v = *t;
  int64_t m = 1, n = 1;
  for (int q = 0; q < t->rank; ++q) {
    v.dims[q] = t->dims[perm[q]];
    v.strides[q] = t->strides[perm[q]];
    (q < nleft ? m : n) *= v.dims[q];
  }
  if (m < 1 || n < 1 || m > INT32_MAX || n > INT32_MAX) return fail(c, PQ_ERR_UNSUPPORTED, "matrix size out of range");
  Tensor A;
  int r = materialize_into(c, v, A);   // a private copy: gesdd overwrites it
  if (r) return r;
  Buf *Bb = nullptr, *Cb = nullptr;
  int chi = 0;
  r = c->dtype == PQ_C128 ? decompose_typed<c128, double>(c, A, m, n, threshold, max_rank, Bb, Cb, chi)
                          : decompose_typed<c64, float>(c, A, m, n, threshold, max_rank, Bb, Cb, chi);
  unref(c, A.buf);
  if (r) return r;
  Tensor B, Cn;
  B.buf = Bb;
  B.rank = nleft + 1;
  for (int q = 0; q < nleft; ++q) B.dims[q] = v.dims[q];
  B.dims[nleft] = chi;
  set_contiguous(B);
  Cn.buf = Cb;
  Cn.rank = nright + 1;
  Cn.dims[0] = chi;
  for (int q = 0; q < nright; ++q) Cn.dims[q + 1] = v.dims[nleft + q];
  set_contiguous(Cn);
  *out_b = add(c, B);
  *out_c = add(c, Cn);
  if (chi_out) *chi_out = chi;
  return PQ_OK;
}

}  // extern "C"
