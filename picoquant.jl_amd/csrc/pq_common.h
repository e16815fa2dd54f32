// Shared device/host definitions for libpq_hip: complex element types, fast
// division by run-time invariant extents, and the index-group descriptors a
// pairwise contraction is lowered to.
//
// A pairwise contraction C = sum_K A * B (src/layer1.jl:85-92) is described
// by three index groups over column-major storage:
//   M : A's open indices   (strides in A and in C)
//   N : B's open indices   (strides in B and in C)
//   K : contracted indices (strides in A and in B)
// plus an optional batch group L present in A, B and C (used by the planned
// executor for slice batches; the reference has no batch indices).
// Every element offset is 32-bit: a tensor holds fewer than 2^31 elements.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PQ_MAX_MODES 24   // modes per group after merging contiguous runs

struct c128 { double x, y; };
struct c64  { float x, y; };

template <typename R> struct cplx_of;
template <> struct cplx_of<double> { typedef c128 type; };
template <> struct cplx_of<float>  { typedef c64 type; };

template <typename C> struct real_of;
template <> struct real_of<c128> { typedef double type; };
template <> struct real_of<c64>  { typedef float type; };

template <typename C>
__host__ __device__ __forceinline__ C cmake(typename real_of<C>::type x, typename real_of<C>::type y) {
  C r; r.x = x; r.y = y; return r;
}

// acc += a * b (complex multiply-add, 4 real FMAs)
template <typename C>
__device__ __forceinline__ void cfma(C& acc, const C a, const C b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// Unsigned division of n < 2^31 by a run-time invariant d >= 1 via
// multiplication (Granlund & Montgomery, "Division by Invariant Integers using
// Multiplication", PLDI 1994): with l = ceil(log2 d) and
// magic = ceil(2^(32+l) / d), q = (n * magic) >> (32 + l) is exact for
// n < 2^31.  Powers of two reduce to a shift (magic = 2^32).
struct FDiv {
  uint64_t magic;
  uint32_t d;
  uint32_t shift;  // 32 + l
};

static inline FDiv make_fdiv(uint32_t d) {
  FDiv f;
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  f.d = d;
  f.shift = 32 + l;
  // ceil(2^(32+l) / d)
  unsigned __int128 num = ((unsigned __int128)1) << (32 + l);
  f.magic = (uint64_t)((num + d - 1) / d);
  return f;
}

__device__ __forceinline__ uint32_t fdiv_q(uint32_t n, const FDiv& f) {
  return (uint32_t)(((uint64_t)n * f.magic) >> f.shift);
}

// One index group: modes are listed fastest-first in the group's own linear
// order; each mode has an extent and up to three strides (one per tensor the
// group touches; unused strides are zero).
struct Group {
  int32_t nmodes;
  uint32_t total;                 // product of extents
  FDiv ext[PQ_MAX_MODES];
  int32_t s0[PQ_MAX_MODES];       // stride in the first tensor
  int32_t s1[PQ_MAX_MODES];       // stride in the second tensor
};

// Decode a linear index of the group into the two tensor offsets.
__device__ __forceinline__ void group_offsets(const Group& g, uint32_t i, int32_t& o0, int32_t& o1) {
  o0 = 0; o1 = 0;
  for (int j = 0; j < g.nmodes; ++j) {
    uint32_t q = fdiv_q(i, g.ext[j]);
    int32_t r = (int32_t)(i - q * g.ext[j].d);
    o0 += r * g.s0[j];
    o1 += r * g.s1[j];
    i = q;
  }
}

// Batch group: entries of a batched contraction (e.g. the slices of one slice
// group).  Entry b reads A at A + offA(b), B at B + offB(b) and writes C at
// C + offC(b); each mode has its own stride per tensor (0 = shared operand).
#define PQ_MAX_BMODES 12
struct Batch {
  int32_t nmodes;
  uint32_t total;
  FDiv ext[PQ_MAX_BMODES];
  int64_t sA[PQ_MAX_BMODES];
  int64_t sB[PQ_MAX_BMODES];
  int64_t sC[PQ_MAX_BMODES];
};

__device__ __forceinline__ void batch_offsets(const Batch& L, uint32_t b, int64_t& oa, int64_t& ob,
                                              int64_t& oc) {
  oa = 0; ob = 0; oc = 0;
  for (int j = 0; j < L.nmodes; ++j) {
    uint32_t q = fdiv_q(b, L.ext[j]);
    int64_t r = (int64_t)(b - q * L.ext[j].d);
    oa += r * L.sA[j];
    ob += r * L.sB[j];
    oc += r * L.sC[j];
    b = q;
  }
}
