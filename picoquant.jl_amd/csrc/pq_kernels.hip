// Bandwidth-bound contraction kernels for gfx950 (64-wide wavefronts).
//
//  k_stream  : big x small pairwise contraction (gate application, absorbing
//              a rank-1 output tensor, outer products with a small factor).
//              One lane owns one X-open index; the small operand Y sits in LDS
//              as a dense [K][N] table, so the big operand is read once and C
//              written once -- the HBM roofline for bond-dimension-2 steps.
//  k_generic : any shape; one lane per output element, loop over K.
//  k_reduce  : tiny output (M*N <= 16) with a long summed extent: split-K
//              partial sums per workgroup, then a deterministic second pass.
//  k_copy    : strided gather into contiguous column-major storage
//              (materialises permuted / sliced views).
#include "pq_internal.h"

#include <type_traits>

namespace {

template <typename C>
__device__ __forceinline__ C czero() { C r; r.x = 0; r.y = 0; return r; }

template <typename C>
__device__ __forceinline__ C ldg(const C* p) { return *p; }

// ---------------------------------------------------------------- k_stream
// The lane's first X-open index has its K values loaded before the small
// operand is staged in LDS: those loads depend on nothing the staging writes,
// so the kernel waits for global memory once before the barrier, not twice.
template <typename C, int KP>
__global__ __launch_bounds__(256) void k_stream(const C* __restrict__ X, const C* __restrict__ Y,
                                                C* __restrict__ Out, StreamArgs a) {
  extern __shared__ unsigned char smem_raw[];
  const uint32_t Kt = a.K.total, Nt = a.N.total, Mt = a.M.total;
  {
    int64_t ox, oy, oc;
    batch_offsets(a.L, blockIdx.z, ox, oy, oc);
    X += ox;
    Y += oy;
    Out += oc;
  }
  // K offsets in X (uniform)
  int32_t xk[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    int32_t ox = 0, oy = 0;
    if (k < (int)Kt) group_offsets(a.K, (uint32_t)k, ox, oy);
    xk[k] = ox;
  }
  uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  C xv[KP];
  int32_t oc = 0;
  if (m < Mt) {
    int32_t ox;
    group_offsets(a.M, m, ox, oc);
#pragma unroll
    for (int k = 0; k < KP; ++k) xv[k] = (k < (int)Kt) ? ldg(X + ox + xk[k]) : czero<C>();
  }
  C* ys = reinterpret_cast<C*>(smem_raw);                              // [KP][Nt]
  int32_t* cn = reinterpret_cast<int32_t*>(ys + (size_t)KP * Nt);      // [Nt]
  for (uint32_t t = threadIdx.x; t < KP * Nt; t += blockDim.x) {
    uint32_t k = t % KP, n = t / KP;
    C v = czero<C>();
    if (k < Kt) {
      int32_t ox, oyk, oyn, ocn;
      group_offsets(a.K, k, ox, oyk);
      group_offsets(a.N, n, oyn, ocn);
      v = Y[oyk + oyn];
    }
    ys[(size_t)n * KP + k] = v;
  }
  for (uint32_t t = threadIdx.x; t < Nt; t += blockDim.x) {
    int32_t oy, ocn;
    group_offsets(a.N, t, oy, ocn);
    cn[t] = ocn;
  }
  __syncthreads();
  for (bool first = true; m < Mt; m += gridDim.x * blockDim.x, first = false) {
    if (!first) {
      int32_t ox;
      group_offsets(a.M, m, ox, oc);
#pragma unroll
      for (int k = 0; k < KP; ++k) xv[k] = (k < (int)Kt) ? ldg(X + ox + xk[k]) : czero<C>();
    }
    for (uint32_t n = 0; n < Nt; ++n) {
      C acc = czero<C>();
      const C* yc = ys + (size_t)n * KP;
#pragma unroll
      for (int k = 0; k < KP; ++k) cfma(acc, xv[k], yc[k]);
      Out[oc + cn[n]] = acc;
    }
  }
}

// --------------------------------------------------------------- k_generic
template <typename C>
__global__ __launch_bounds__(256) void k_generic(const C* __restrict__ A, const C* __restrict__ B,
                                                 C* __restrict__ Out, GenArgs a) {
  const uint32_t Mt = a.M.total, Nt = a.N.total, Kt = a.K.total;
  const uint64_t total = (uint64_t)Mt * Nt;
  {
    int64_t oa, ob, oc;
    batch_offsets(a.L, blockIdx.z, oa, ob, oc);
    A += oa;
    B += ob;
    Out += oc;
  }
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t m = (uint32_t)(t % Mt), n = (uint32_t)(t / Mt);
    int32_t oa, oc1, ob, oc2;
    group_offsets(a.M, m, oa, oc1);
    group_offsets(a.N, n, ob, oc2);
    C acc = czero<C>();
    for (uint32_t k = 0; k < Kt; ++k) {
      int32_t ka, kb;
      group_offsets(a.K, k, ka, kb);
      cfma(acc, A[oa + ka], B[ob + kb]);
    }
    Out[oc1 + oc2] = acc;
  }
}

// ---------------------------------------------------------------- k_reduce
// Phase 1: workgroup b sums its K chunk for every output; partials[b*MN + o].
template <typename C>
__global__ __launch_bounds__(256) void k_reduce_partial(const C* __restrict__ A, const C* __restrict__ B,
                                                        C* __restrict__ partials, GenArgs a) {
  __shared__ C red[256];
  const uint32_t Mt = a.M.total, Nt = a.N.total, Kt = a.K.total;
  const uint32_t MN = Mt * Nt;
  {
    int64_t oa, ob, oc;
    batch_offsets(a.L, blockIdx.z, oa, ob, oc);
    A += oa;
    B += ob;
  }
  partials += (size_t)blockIdx.z * gridDim.x * MN;
  const uint32_t chunk = (Kt + gridDim.x - 1) / gridDim.x;
  const uint32_t k0 = blockIdx.x * chunk;
  const uint32_t k1 = min(Kt, k0 + chunk);
  for (uint32_t o = 0; o < MN; ++o) {
    uint32_t m = o % Mt, n = o / Mt;
    int32_t oa, oc1, ob, oc2;
    group_offsets(a.M, m, oa, oc1);
    group_offsets(a.N, n, ob, oc2);
    C acc = czero<C>();
    for (uint32_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
      int32_t ka, kb;
      group_offsets(a.K, k, ka, kb);
      cfma(acc, A[oa + ka], B[ob + kb]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        red[threadIdx.x].x += red[threadIdx.x + w].x;
        red[threadIdx.x].y += red[threadIdx.x + w].y;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[(size_t)blockIdx.x * MN + o] = red[0];
    __syncthreads();
  }
}

// Phase 2: one workgroup per output sums the partials in a fixed order.
template <typename C>
__global__ __launch_bounds__(256) void k_reduce_final(const C* __restrict__ partials, C* __restrict__ Out,
                                                      GenArgs a, int nblocks) {
  __shared__ C red[256];
  const uint32_t Mt = a.M.total, Nt = a.N.total;
  const uint32_t MN = Mt * Nt;
  const uint32_t o = blockIdx.x;
  {
    int64_t oa, ob, oc;
    batch_offsets(a.L, blockIdx.z, oa, ob, oc);
    Out += oc;
  }
  partials += (size_t)blockIdx.z * nblocks * MN;
  C acc = czero<C>();
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    C p = partials[(size_t)b * MN + o];
    acc.x += p.x; acc.y += p.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x].x += red[threadIdx.x + w].x;
      red[threadIdx.x].y += red[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint32_t m = o % Mt, n = o / Mt;
    int32_t oa, oc1, ob, oc2;
    group_offsets(a.M, m, oa, oc1);
    group_offsets(a.N, n, ob, oc2);
    Out[oc1 + oc2] = red[0];
  }
}

// ------------------------------------------------------------------ k_copy
template <typename C>
__global__ __launch_bounds__(256) void k_copy(const C* __restrict__ In, C* __restrict__ Out, CopyArgs a) {
  const uint32_t T = a.G.total;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
    int32_t oi, oo;
    group_offsets(a.G, t, oi, oo);
    Out[oo] = In[oi];
  }
}

// ----------------------------------------------------------------- k_multi
// Grouped small contractions: problem blockIdx.y, batch entry blockIdx.z,
// one lane per output element (grid-stride over blockIdx.x), loop over K;
// a problem may instead be a small element-type conversion.
template <typename C>
__global__ __launch_bounds__(256) void k_multi(const MultiDesc* __restrict__ descs, int64_t group) {
  const MultiDesc& d = descs[blockIdx.y];
  if (blockIdx.z >= d.L.total) return;
  int64_t oa, ob, oc;
  batch_offsets(d.L, blockIdx.z, oa, ob, oc);
  for (int j = 0; j < d.nhigh; ++j)
    if ((group >> j) & 1) {
      oa += d.hA[j];
      ob += d.hB[j];
      oc += d.hC[j];
    }
  if (d.conv) {
    // element-type conversion (complex128 <-> complex64), same layout
    typedef typename std::conditional<sizeof(C) == 16, c64, c128>::type D;
    const C* A = reinterpret_cast<const C*>(d.A) + oa;
    D* Out = reinterpret_cast<D*>(d.C) + oc;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < d.M.total; t += gridDim.x * blockDim.x) {
      const C v = A[t];
      D w;
      w.x = (decltype(w.x))v.x;
      w.y = (decltype(w.y))v.y;
      Out[t] = w;
    }
    return;
  }
  const C* A = reinterpret_cast<const C*>(d.A) + oa;
  const C* B = reinterpret_cast<const C*>(d.B) + ob;
  C* Out = reinterpret_cast<C*>(d.C) + oc;
  const uint32_t Mt = d.M.total, Nt = d.N.total, Kt = d.K.total;
  const uint32_t total = Mt * Nt;
  // the problem's K offsets in LDS (grouped problems have K <= PQ_MULTI_KMAX),
  // so a lane issues the loads of eight K values at once instead of waiting
  // for each before decoding the next
  __shared__ int2 kof[PQ_MULTI_KMAX];
  const bool tab = Kt <= PQ_MULTI_KMAX;
  if (tab) {
    for (uint32_t k = threadIdx.x; k < Kt; k += blockDim.x) {
      int32_t ka, kb;
      group_offsets(d.K, k, ka, kb);
      kof[k] = make_int2(ka, kb);
    }
    __syncthreads();
  }
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    uint32_t m = t % Mt, n = t / Mt;
    int32_t pa, pc1, pb, pc2;
    group_offsets(d.M, m, pa, pc1);
    group_offsets(d.N, n, pb, pc2);
    C acc = czero<C>();
    uint32_t k = 0;
    if (tab) {
      for (; k + 8 <= Kt; k += 8) {
        C a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int2 o = kof[k + u];
          a[u] = A[pa + o.x];
          b[u] = B[pb + o.y];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) cfma(acc, a[u], b[u]);
      }
      for (; k < Kt; ++k) {
        const int2 o = kof[k];
        cfma(acc, A[pa + o.x], B[pb + o.y]);
      }
    } else {
      for (; k < Kt; ++k) {
        int32_t ka, kb;
        group_offsets(d.K, k, ka, kb);
        cfma(acc, A[pa + ka], B[pb + kb]);
      }
    }
    Out[pc1 + pc2] = acc;
  }
}

// ------------------------------------------------------------------ k_forest
// One item: an output element of a small unbatched step (one lane, the K
// offsets from the step's table), or one element of a conversion.
template <typename C>
__device__ __forceinline__ void forest_item(const ForestStep& f, const int4& it, const int2* ktab, char* dat) {
  // an operand in the tree's LDS data area holds its byte offset there
  auto at = [&](const void* q, int bit) -> char* {
    return ((f.lds >> bit) & 1) ? dat + (size_t)q : reinterpret_cast<char*>(const_cast<void*>(q));
  };
  if (f.conv) {
    typedef typename std::conditional<sizeof(C) == 16, c64, c128>::type D;
    const C v = reinterpret_cast<const C*>(at(f.A, 0))[it.y];
    D w;
    w.x = (decltype(w.x))v.x;
    w.y = (decltype(w.y))v.y;
    reinterpret_cast<D*>(at(f.C, 2))[it.w] = w;
    return;
  }
  const C* A = reinterpret_cast<const C*>(at(f.A, 0)) + it.y;
  const C* B = reinterpret_cast<const C*>(at(f.B, 1)) + it.z;
  C acc = czero<C>();
  // the loads of several K values in flight at once (a tree's level waits
  // for its slowest lane)
#pragma unroll 4
  for (int32_t k = 0; k < f.nk; ++k) {
    const int2 o = ktab[k];
    cfma(acc, A[o.x], B[o.y]);
  }
  reinterpret_cast<C*>(at(f.C, 2))[it.w] = acc;
}

// Workgroup b runs tree b level by level: a level's items (the outputs of
// all its steps, which read only earlier levels) are spread over the lanes,
// and the barrier after each level makes its stores visible to the
// workgroup's later levels (one workgroup runs on one CU).  Tree header at
// lv[lv[b]]: levels, first step record, step records, first K entry, K
// entries, first leaf copy, leaf copies, then the levels' item bounds.
// LDS: the tree's step records and K offsets are copied into LDS once, so
// a level waits on global memory only for its operands (the items of the
// next level are read ahead, as they depend on nothing the level computes);
// the results that the tree itself reads later, and the small leaves it
// reads, live in the LDS data area after them (leaves copied in at the
// start), so only the other leaves and the tree's root result go through
// global memory.
template <bool LDS>
__global__ __launch_bounds__(256) void k_forest(const ForestStep* __restrict__ fs, const int4* __restrict__ items,
                                                const int2* __restrict__ ktab,
                                                const ForestCopy* __restrict__ copies,
                                                const int32_t* __restrict__ lv) {
  extern __shared__ __align__(16) int32_t fsm[];
  const int32_t q = lv[blockIdx.x];
  const int32_t nlev = lv[q], fs0 = lv[q + 1], nfs = lv[q + 2], k0 = lv[q + 3], nkt = lv[q + 4];
  const int32_t c0 = lv[q + 5], nc = lv[q + 6];
  const int32_t* bnd = lv + q + 7;
  const ForestStep* F = fs + fs0;
  const int2* KT = ktab + k0;
  char* dat = nullptr;
  if constexpr (LDS) {
    constexpr int W = (int)(sizeof(ForestStep) / 4);
    const int tid = (int)threadIdx.x;
    // leaf copies first: their loads are in flight while the records move
    float2 cv[PQ_FOREST_COPY_PER_LANE];
    int32_t cd[PQ_FOREST_COPY_PER_LANE];
#pragma unroll
    for (int j = 0; j < PQ_FOREST_COPY_PER_LANE; ++j) {
      const int32_t e = tid + 256 * j;
      if (e < nc) {
        const ForestCopy cp = copies[c0 + e];
        cv[j] = *reinterpret_cast<const float2*>(cp.src);
        cd[j] = cp.dst;
      }
    }
    const int32_t* src = reinterpret_cast<const int32_t*>(fs + fs0);
    for (int32_t e = tid; e < nfs * W; e += (int32_t)blockDim.x) fsm[e] = src[e];
    const int32_t kw = nfs * W + (nfs * W & 1);
    int2* kd = reinterpret_cast<int2*>(fsm + kw);
    for (int32_t e = tid; e < nkt; e += (int32_t)blockDim.x) kd[e] = ktab[k0 + e];
    dat = reinterpret_cast<char*>(fsm) + (((size_t)kw * 4 + (size_t)nkt * 8 + 15) & ~(size_t)15);
#pragma unroll
    for (int j = 0; j < PQ_FOREST_COPY_PER_LANE; ++j)
      if (tid + 256 * j < nc) *reinterpret_cast<float2*>(dat + cd[j]) = cv[j];
    __syncthreads();
    F = reinterpret_cast<const ForestStep*>(fsm);
    KT = kd;
  }
  int4 nx = make_int4(0, 0, 0, 0);
  if (nlev > 0 && bnd[0] + (int32_t)threadIdx.x < bnd[1]) nx = items[bnd[0] + threadIdx.x];
  for (int32_t l = 0; l < nlev; ++l) {
    const int32_t i0 = bnd[l] + (int32_t)threadIdx.x, i1 = bnd[l + 1];
    const int4 cur = nx;
    if (l + 1 < nlev && i1 + (int32_t)threadIdx.x < bnd[l + 2]) nx = items[i1 + threadIdx.x];
    // the lane's next items of this level, loaded together (a level of a
    // few thousand outputs gives a lane several)
    constexpr int PRE = 3;
    int4 pre[PRE];
#pragma unroll
    for (int j = 0; j < PRE; ++j) {
      const int32_t i = i0 + (j + 1) * (int32_t)blockDim.x;
      if (i < i1) pre[j] = items[i];
    }
    int32_t c = 0;
    for (int32_t i = i0; i < i1; i += (int32_t)blockDim.x, ++c) {
      int4 it;
      if (c == 0) {
        it = cur;
      } else if (c <= PRE) {
        it = pre[0];
        if (c == 2) it = pre[1];
        if (c == 3) it = pre[2];
      } else {
        it = items[i];
      }
      const ForestStep& f = F[it.x - fs0];
      if (f.c128)
        forest_item<c128>(f, it, KT + (f.kt - k0), dat);
      else
        forest_item<c64>(f, it, KT + (f.kt - k0), dat);
    }
    __syncthreads();
  }
}

inline unsigned grid_for(uint64_t work, unsigned block, unsigned cap = 8192) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

template <typename C>
hipError_t launch_stream(const C* X, const C* Y, C* Out, const StreamArgs& a, hipStream_t s) {
  const uint32_t Kt = a.K.total, Nt = a.N.total;
  int KP = 1;
  while ((uint32_t)KP < Kt) KP <<= 1;
  size_t smem = (size_t)KP * Nt * sizeof(C) + (size_t)Nt * 4;
  unsigned grid = grid_for(a.M.total, 256, a.L.total > 1 ? 1024 : 4096);
  switch (KP) {
#define PQ_CASE(P) \
  case P: hipLaunchKernelGGL((k_stream<C, P>), dim3(grid, 1, a.L.total), dim3(256), smem, s, X, Y, Out, a); break;
    PQ_CASE(1) PQ_CASE(2) PQ_CASE(4) PQ_CASE(8) PQ_CASE(16) PQ_CASE(32)
#undef PQ_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename C>
hipError_t launch_generic(const C* A, const C* B, C* Out, const GenArgs& a, hipStream_t s) {
  uint64_t work = (uint64_t)a.M.total * a.N.total;
  hipLaunchKernelGGL((k_generic<C>), dim3(grid_for(work, 256, a.L.total > 1 ? 1024 : 8192), 1, a.L.total),
                     dim3(256), 0, s, A, B, Out, a);
  return hipGetLastError();
}

int reduce_blocks_for(uint32_t ktotal) {
  uint32_t b = ktotal / 4096;
  if (b < 1) b = 1;
  if (b > 2048) b = 2048;
  return (int)b;
}

template <typename C>
hipError_t launch_reduce(const C* A, const C* B, C* Out, const GenArgs& a, C* partials, int nblocks,
                         hipStream_t s) {
  hipLaunchKernelGGL((k_reduce_partial<C>), dim3(nblocks, 1, a.L.total), dim3(256), 0, s, A, B, partials, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_reduce_final<C>), dim3(a.M.total * a.N.total, 1, a.L.total), dim3(256), 0, s, partials,
                     Out, a, nblocks);
  return hipGetLastError();
}

template <typename C>
hipError_t launch_copy(const C* In, C* Out, const CopyArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_copy<C>), dim3(grid_for(a.G.total, 256)), dim3(256), 0, s, In, Out, a);
  return hipGetLastError();
}

template <typename C>
hipError_t launch_multi(const MultiDesc* descs, int n, unsigned blocks_x, unsigned batch, int64_t group,
                        hipStream_t s) {
  if (n < 1 || n > 65535 || batch < 1 || batch > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_multi<C>), dim3(blocks_x, n, batch), dim3(256), 0, s, descs, group);
  return hipGetLastError();
}

hipError_t launch_forest(const ForestStep* steps, const int4* items, const int2* ktab, const ForestCopy* copies,
                         const int32_t* lv, int ntrees, size_t lds_bytes, hipStream_t s) {
  if (ntrees < 1 || ntrees > (1 << 24) || lds_bytes > PQ_FOREST_LDS_BYTES) return hipErrorInvalidValue;
  if (lds_bytes > 0)
    hipLaunchKernelGGL(k_forest<true>, dim3(ntrees), dim3(256), lds_bytes, s, steps, items, ktab, copies, lv);
  else
    hipLaunchKernelGGL(k_forest<false>, dim3(ntrees), dim3(256), 0, s, steps, items, ktab, copies, lv);
  return hipGetLastError();
}

#define PQ_INST(C)                                                                                   \
  template hipError_t launch_multi<C>(const MultiDesc*, int, unsigned, unsigned, int64_t, hipStream_t); \
  template hipError_t launch_stream<C>(const C*, const C*, C*, const StreamArgs&, hipStream_t);      \
  template hipError_t launch_generic<C>(const C*, const C*, C*, const GenArgs&, hipStream_t);        \
  template hipError_t launch_reduce<C>(const C*, const C*, C*, const GenArgs&, C*, int, hipStream_t); \
  template hipError_t launch_copy<C>(const C*, C*, const CopyArgs&, hipStream_t);
PQ_INST(c128)
PQ_INST(c64)
