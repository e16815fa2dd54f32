// pq_contract_noalloc: pq_contract into a caller-provided result tensor,
// with no device allocation at call time.  The partial products of split-K
// GEMMs and split reductions go to `scratch`, a device array the code
// object brings: every GPU the library runs on holds its own copy until the
// process ends, and every call on that GPU shares it (calls that overlap on
// one GPU race).  Both partial kernels write every partial they later read,
// so the array needs no clearing between calls.
#include "pq_rt.h"

#include <mutex>

using namespace pqrt;

namespace {

static __device__ __align__(256) unsigned char scratch[PQ_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
bool g_done[kMaxDevices] = {};
void* g_ptr[kMaxDevices] = {};

// Zero fill of n interleaved complex elements (graph-capture safe, unlike a
// memset node).
template <typename R>
__global__ void k_zero(R* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = R(0);
}

}  // namespace

namespace pqrt {

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_done[device]) {
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    void* ptr = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ptr = nullptr;
    }
    hipSetDevice(prev);
    g_ptr[device] = ptr;
    g_done[device] = true;
  }
  return g_ptr[device];
}

}  // namespace pqrt

extern "C" int pq_contract_noalloc(pq_ctx* c, pq_tensor ha, int ra, const int32_t* ia, pq_tensor hb, int rb,
                                   const int32_t* ib, pq_tensor hc) {
  if (!c) return PQ_ERR_ARG;
  Tensor* A = get(c, ha);
  Tensor* B = get(c, hb);
  Tensor* Co = get(c, hc);
  if (!A || !B || !Co) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  Tensor Ct;
  StepSpec s;
  int64_t kt = 0;
  int r = prepare_contract(c, *A, ra, ia, *B, rb, ib, Ct, s, kt);
  if (r) return r;
  if (Co->rank != Ct.rank) return fail(c, PQ_ERR_ARG, "result tensor has the wrong rank");
  for (int i = 0; i < Ct.rank; ++i)
    if (Co->dims[i] != Ct.dims[i]) return fail(c, PQ_ERR_ARG, "result tensor has the wrong dims");
  if (!is_contiguous(*Co)) return fail(c, PQ_ERR_ARG, "result tensor must be contiguous");
  if (Co->buf == A->buf || Co->buf == B->buf) return fail(c, PQ_ERR_ARG, "result tensor shares an operand's buffer");
  const int64_t nc = numel(Ct);
  unsigned char* pc = (unsigned char*)Co->buf->ptr + Co->off * c->esize;
  if (nc == 0) return PQ_OK;
  if (kt == 0) {
    const unsigned g = (unsigned)std::min<int64_t>((2 * nc + 255) / 256, 4096);
    if (c->dtype == PQ_C128)
      hipLaunchKernelGGL((k_zero<double>), dim3(g), dim3(256), 0, c->stream, (double*)pc, nc);
    else
      hipLaunchKernelGGL((k_zero<float>), dim3(g), dim3(256), 0, c->stream, (float*)pc, nc);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? PQ_OK : hipfail(c, e, "zero fill");
  }
  const size_t need = step_scratch_bytes(s, c->esize);
  void* sp = nullptr;
  if (need > 0) {
    if (need > (size_t)PQ_NOALLOC_SCRATCH_BYTES)
      return fail(c, PQ_ERR_UNSUPPORTED,
                  "contraction needs " + std::to_string(need) + " bytes of scratch, more than the " +
                      std::to_string((size_t)PQ_NOALLOC_SCRATCH_BYTES) + " pq_contract_noalloc holds");
    sp = c->noalloc_scratch;
    if (!sp) return fail(c, PQ_ERR_HIP, "the device scratch of pq_contract_noalloc is unavailable");
  }
  const unsigned char* pa = (const unsigned char*)A->buf->ptr + A->off * c->esize;
  const unsigned char* pb = (const unsigned char*)B->buf->ptr + B->off * c->esize;
  return launch_step(c, pa, pb, pc, s, sp);
}
