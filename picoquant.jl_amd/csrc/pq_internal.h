// Internal interface between the runtime (context, tensors, allocator) and
// the kernel launchers.
#pragma once
#include "pq_common.h"
#include "../../include/pq_hip.h"
#include <string>

// Big x small pairwise contraction: X is the large operand, Y the small one
// (at most PQ_STREAM_KMAX contracted elements, PQ_STREAM_YMAX elements).
//   M: X-open  (s0 = stride in X, s1 = stride in C), fastest C stride first
//   K: summed  (s0 = stride in X, s1 = stride in Y)
//   N: Y-open  (s0 = stride in Y, s1 = stride in C)
#define PQ_STREAM_KMAX 32
#define PQ_STREAM_YMAX 2048

struct StreamArgs {
  Group M, K, N;
  Batch L;   // sA = stride in X, sB = stride in Y, sC = stride in C
};

// General contraction (any shape):
//   M: A-open (s0 = A, s1 = C); N: B-open (s0 = B, s1 = C); K: (s0 = A, s1 = B)
struct GenArgs {
  Group M, N, K;
  Batch L;
};

// Strided copy: group over output linear order, s0 = input stride, s1 = output stride.
struct CopyArgs {
  Group G;
};

// Tiled GEMM on MFMA with gathered operand loads and scattered C stores.
// Offsets are split as (row group, col group) so a tile loads through two
// small tables instead of per-element decoding.
struct GemmArgs {
  Group M, N, K;      // as GenArgs
  Batch L;            // batch entries on grid z
  int32_t a_kfast;    // A's unit-stride mode is in K: lanes walk k when loading
  int32_t b_kfast;    // B's unit-stride mode is in K
  int32_t tile;       // see pq_set_gemm_tile
  int32_t contig;     // first K mode unit-stride in A and B, 16-byte aligned runs
};

// One problem of a grouped launch of small contractions (PQ_KC_BATCHED):
// many independent steps of one dependency level in a single kernel launch,
// one problem per grid row.  Pointers are for slice group 0; hA/hB/hC add
// the leaf offset of each slice-group bit (group index passed at launch).
#define PQ_MAX_HIGH_BITS 24
struct MultiDesc {
  const void* A;
  const void* B;
  void* C;
  Group M, N, K;      // as GenArgs
  Batch L;            // batch entries of one slice group (grid z)
  int32_t conv;       // 1: element-type conversion of M.total elements A -> C (the other type)
  int32_t c128;       // forest launches: the step computes in (conversions: converts from) complex128
  int32_t nhigh;
  int64_t hA[PQ_MAX_HIGH_BITS], hB[PQ_MAX_HIGH_BITS], hC[PQ_MAX_HIGH_BITS];
};

// K offsets of a grouped problem are staged in LDS up to this K (larger K
// decodes each offset in the loop)
#define PQ_MULTI_KMAX 64
template <typename C>
hipError_t launch_multi(const MultiDesc* descs, int n, unsigned blocks_x, unsigned batch, int64_t group,
                        hipStream_t s);

// Independent trees of small unbatched steps in one launch (k_forest), one
// workgroup per tree, run level by level with a barrier between levels.
// Every output element of a step is one item (step, A offset, B offset,
// C offset), precomputed on the host; a step's K offsets (A, B) are
// ktab[kt .. kt + nk).  Conversions: item = (step, source offset, -, target
// offset).  lv[b] is the offset in lv of tree b's entry: nlevels, its first
// step record and number of step records, its first K table entry and
// number of K entries (a tree's records and K entries are contiguous), its
// first leaf copy and number of leaf copies, then the nlevels + 1 item
// boundaries of its levels.
struct ForestStep {
  const void* A;
  const void* B;
  void* C;
  int32_t nk, kt;     // K extent, first K table entry
  int32_t conv;       // 1: element-type conversion A -> C (the other type)
  int32_t c128;       // computes in (converts from) complex128
  int32_t lds;        // bit 0 / 1 / 2: A / B / C is in the tree's LDS data area,
                      // the pointer holding its byte offset there
};
// Leaf operands copied into a tree's LDS data area when its workgroup
// starts, 8 bytes per entry: source address and byte offset in the area.
struct ForestCopy {
  const void* src;
  int32_t dst, pad;
};
// A tree's LDS: its step records (32-bit words, padded to 8 bytes), its K
// offsets, then (16-byte aligned) its data area.  Per-lane leaf copies:
#define PQ_FOREST_COPY_PER_LANE 8
// lds_bytes: LDS of the largest tree (0: the kernel reads records and K
// offsets from global memory and keeps every operand there), at most
// PQ_FOREST_LDS_BYTES
#define PQ_FOREST_LDS_BYTES (48 << 10)
hipError_t launch_forest(const ForestStep* steps, const int4* items, const int2* ktab, const ForestCopy* copies,
                         const int32_t* lv, int ntrees, size_t lds_bytes, hipStream_t s);

template <typename C>
hipError_t launch_stream(const C* X, const C* Y, C* Out, const StreamArgs& a, hipStream_t s);
template <typename C>
hipError_t launch_generic(const C* A, const C* B, C* Out, const GenArgs& a, hipStream_t s);
template <typename C>
hipError_t launch_reduce(const C* A, const C* B, C* Out, const GenArgs& a, C* partials,
                         int nblocks, hipStream_t s);
template <typename C>
hipError_t launch_copy(const C* In, C* Out, const CopyArgs& a, hipStream_t s);
template <typename C>
hipError_t launch_gemm(const C* A, const C* B, C* Out, const GemmArgs& a, hipStream_t s);
// Split-K second pass: Out[C offset of (m, n)] = sum over z < parts of
// P[z * mn + m + mt * n], summed in FP64; gm / gn give each output's C
// offset (s1) for its row m < mt and column n (s0 unused).
template <typename C>
hipError_t launch_sum_partials(const C* P, C* Out, const Group& gm, const Group& gn, uint32_t mt, uint64_t mn,
                               int parts, hipStream_t s);
int gemm_k_limit();
// complex64 GEMM on FP16 hi/lo matrix cores (contiguous K, multiple of 32).
bool gemm16_applicable(const GemmArgs& g);
hipError_t launch_gemm16(const c64* A, const c64* B, c64* Out, const GemmArgs& a, hipStream_t s);

// Number of partial blocks the split-K reduction uses for a given K extent.
int reduce_blocks_for(uint32_t ktotal);

// Fused chain of small complex64 operators on one large tensor (plan steps
// of kind 2; the program format is documented with PQ_CHAIN_* in pq_hip.h).
// The host expands each program into int32 address tables (logical tile
// addresses and element offsets) in one device buffer; a bit list of n bits
// is split into a low part (up to CH_LO_BITS: gather / scatter, 4 bits: R) taken
// from the lane and a high part taken from the loop counter, each with its
// own table, so the kernel only adds table entries.
#define CH_TB 13   // PQ_CHAIN_MAX_TILE_BITS
#define CH_OB 5    // PQ_CHAIN_MAX_OP_BITS
#define CH_ST 8    // PQ_CHAIN_MAX_STAGES
#define CH_THREADS 512   // 8 waves per workgroup, two persistent workgroups per CU
#define CH_LO_BITS 9     // gather / scatter bits taken from the lane (2^9 = CH_THREADS)
#define CH_NIT (1 << (CH_TB - 1 - CH_LO_BITS))   // gather / scatter iterations (element pairs) per lane and tile
struct ChainStage {
  const c64* S;
  int32_t nk, nm;          // 2^nk operator columns, 2^nm rows
  int32_t rlo, rhi;        // R bits from the lane (<= 4; 6 for K, M <= 4) and from the block counter
  int32_t kp, mp, sk, sm;  // table offsets: K / M tile addresses, K / M operator offsets
  int32_t rl, rh;          // table offsets: R tile addresses (low, high part)
  int32_t sd, ko, mo, ro, rho;  // LDS staging: dense operator [m][k] (c64 units), K / M /
                                // low-R / high-R tile addresses (int units)
};
struct ChainArgs {
  const c64* X;
  c64* Y;
  const int32_t* tab;
  int32_t tb, nst;
  int32_t ilo, ihi, olo, ohi;              // gather / scatter bits from the lane, from the counter
  int32_t ilp, ilg, ihp, ihg;              // gather tables: low / high tile address, element offset
  int32_t olp, olg, ohp, ohg;              // scatter tables
  int32_t sdn, idn;                        // staged c64 operator elements, staged ints
  // Each lane moves element pairs: the pair bit's tile address (XOR) and
  // tensor stride on the gather and scatter side, and whether the pair is
  // one aligned 16-byte access (pair stride 1, every other offset even)
  int32_t gpa, gps, spa, sps;
  int32_t gvec_ok, svec_ok;                // strides allow it (host); the launch also checks alignment
  int32_t gvec, svec;
  int32_t tof;  // table offset of the per-tile (X, Y) spectator offsets, -1: decode spec per tile
  Group spec;   // spectator modes: s0 = stride in X, s1 = stride in Y
  ChainStage st[CH_ST];
};
#define CH_STAGE_BYTES PQ_CHAIN_STAGE_BYTES   // LDS after the tile for the staged operators and offsets
#define CH_TOF_TILES (1 << 21)   // chains with at most this many tiles get a per-tile offset table
hipError_t launch_chain(const ChainArgs& a, hipStream_t s);
