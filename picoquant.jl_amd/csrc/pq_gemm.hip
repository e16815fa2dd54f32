// Complex GEMM on CDNA4 matrix cores with the index permutation fused into
// operand loads and result stores (TTGT without the transposes).
//
//   C[m, n] = sum_k A[m, k] * B[k, n]      (per batch entry z)
//
// m, n, k are linear indices of the M / N / K index groups; element offsets
// come from per-workgroup offset tables in LDS (row tables for the tile,
// K tables in windows of 2048), so any stride pattern is read and written
// in place.
//
// Complex product as four real MFMAs on split planes:
//   Cr += Ar*Br + (-Ai)*Bi ;  Ci += Ar*Bi + Ai*Br
// complex64  -> v_mfma_f32_16x16x4_f32 (fp32 accumulate)
// complex128 -> v_mfma_f64_16x16x4_f64 (fp64 accumulate)
//
// Tile: 256 lanes = 4 waves arranged WGM x WGN (2x2, 4x1 or 1x4); workgroup
// tile BM x BN, each wave (BM/WGM) x (BN/WGN) as a grid of 16x16 MFMA tiles;
// BK = 16 per LDS stage.  Extents below the tile are zero-padded.  Lane layout of a 16x16x4 MFMA (64 lanes): A operand lane l holds
// A[l%16][l/16], B operand B[l/16][l%16]; accumulator register r holds
// D[4*(l/16)+r][l%16] for f32 and D[l/16+4r][l%16] for f64.
#include "pq_internal.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename R> struct mfma_t;
template <> struct mfma_t<float> {
  typedef floatx4 acc;
  // accumulator register r of lane-group g (= lane/16) holds row 4*g + r
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
  static __device__ __forceinline__ acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <> struct mfma_t<double> {
  typedef doublex4 acc;
  // the f64 variant interleaves: register r of lane-group g holds row g + 4*r
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
  static __device__ __forceinline__ acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

constexpr int BK = 16;
constexpr int KTAB_MAX = 2048;   // K offsets held in LDS at a time
constexpr uint32_t GEMM_K_MAX = 1u << 24;

// Workgroups are dealt to the 8 XCDs round-robin in launch order, and each
// XCD has its own L2.  Renumber the (x, y) tiles of a batch entry so that
// each XCD gets one contiguous run of tiles: the tiles that share rows of
// the large operand (consecutive x for one y) then meet in the same L2.
__device__ __forceinline__ void xcd_tile(uint32_t& tx, uint32_t& ty) {
  const uint32_t gx = gridDim.x, n = gridDim.x * gridDim.y;
  const uint32_t b = blockIdx.x + gx * blockIdx.y;
  uint32_t l = b;
  if (n % 8 == 0 && n >= 16) l = (b % 8) * (n / 8) + b / 8;
  tx = l % gx;
  ty = l / gx;
}

// FP32 accumulation in two levels: each K tile's products are summed in a
// fresh accumulator (at most 16 MFMA roundings), which is then added to the
// running sum.  A single chain over K = 256 rounds 128 times into a growing
// sum; two levels cut the rounding error of long sums about 2-3x for 2x the
// accumulator registers.  FP64 accumulates in one chain.
template <typename MF, int TM, int TN>
__device__ __forceinline__ void zero_acc(typename MF::acc (&xr)[TM][TN], typename MF::acc (&xi)[TM][TN]) {
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        xr[i][j][r] = 0;
        xi[i][j][r] = 0;
      }
}

template <typename MF, int TM, int TN>
__device__ __forceinline__ void add_chunk(typename MF::acc (&accr)[TM][TN], typename MF::acc (&acci)[TM][TN],
                                          const typename MF::acc (&cr)[TM][TN], const typename MF::acc (&ci)[TM][TN]) {
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        accr[i][j][r] += cr[i][j][r];
        acci[i][j][r] += ci[i][j][r];
      }
}

template <typename C, int BM, int BN, int WGM, bool AK, bool BK_FAST>
__global__ __launch_bounds__(256) void k_gemm(const C* __restrict__ A, const C* __restrict__ B,
                                              C* __restrict__ Out, GemmArgs g) {
  typedef typename real_of<C>::type R;
  typedef mfma_t<R> MF;
  constexpr int WGN = 4 / WGM;
  constexpr int WM = BM / WGM, WN = BN / WGN;    // wave tile
  static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be a multiple of 16");
  constexpr int TM = WM / 16, TN = WN / 16;      // 16x16 tiles per wave
  constexpr bool CHUNKED = sizeof(R) == 4;

  __shared__ R As_r[BK][BM + 4], As_i[BK][BM + 4];
  __shared__ R Bs_r[BK][BN + 4], Bs_i[BK][BN + 4];
  __shared__ int32_t offAm[BM], offCm[BM], offBn[BN], offCn[BN];
  __shared__ int32_t offAk[KTAB_MAX], offBk[KTAB_MAX];

  const uint32_t Mt = g.M.total, Nt = g.N.total, Kt = g.K.total;
  uint32_t tile_x, tile_y;
  xcd_tile(tile_x, tile_y);
  const uint32_t m0 = tile_x * BM, n0 = tile_y * BN;
  const int tid = threadIdx.x;
  {
    int64_t oa, ob, oc;
    batch_offsets(g.L, blockIdx.z, oa, ob, oc);
    A += oa;
    B += ob;
    Out += oc;
  }

  for (int t = tid; t < BM; t += 256) {
    int32_t oa = 0, oc = 0;
    if (m0 + t < Mt) group_offsets(g.M, m0 + t, oa, oc);
    offAm[t] = oa;
    offCm[t] = oc;
  }
  for (int t = tid; t < BN; t += 256) {
    int32_t ob = 0, oc = 0;
    if (n0 + t < Nt) group_offsets(g.N, n0 + t, ob, oc);
    offBn[t] = ob;
    offCn[t] = oc;
  }
  const int wave = tid / 64, lane = tid % 64;
  const int wm = (wave / WGN) * WM, wn = (wave % WGN) * WN;
  const int li = lane % 16, lk = lane / 16;

  typename MF::acc accr[TM][TN], acci[TM][TN];
  zero_acc<MF, TM, TN>(accr, acci);

  // K runs in windows of KTAB_MAX indices, each with its own offset table
  // (every K tile of a window ends with a barrier, so the next window's
  // table may overwrite the last one's)
  for (uint32_t kb = 0; kb < Kt; kb += KTAB_MAX) {
    const uint32_t ke = Kt - kb < (uint32_t)KTAB_MAX ? Kt : kb + KTAB_MAX;
    for (uint32_t t = tid; t < ke - kb; t += 256) {
      int32_t oa, ob;
      group_offsets(g.K, kb + t, oa, ob);
      offAk[t] = oa;
      offBk[t] = ob;
    }
    __syncthreads();

    for (uint32_t k0 = kb; k0 < ke; k0 += BK) {
      typename MF::acc cr[TM][TN], ci[TM][TN];
      if constexpr (CHUNKED) zero_acc<MF, TM, TN>(cr, ci);
      // A tile BM x BK -> As[k][m]
      for (int t = tid; t < BM * BK; t += 256) {
        int m, k;
        if (AK) { k = t % BK; m = t / BK; } else { m = t % BM; k = t / BM; }
        C v;
        v.x = 0; v.y = 0;
        if (m0 + m < Mt && k0 + k < ke) v = A[offAm[m] + offAk[k0 - kb + k]];
        As_r[k][m] = v.x;
        As_i[k][m] = v.y;
      }
      // B tile BK x BN -> Bs[k][n]
      for (int t = tid; t < BK * BN; t += 256) {
        int n, k;
        if (BK_FAST) { k = t % BK; n = t / BK; } else { n = t % BN; k = t / BN; }
        C v;
        v.x = 0; v.y = 0;
        if (n0 + n < Nt && k0 + k < ke) v = B[offBk[k0 - kb + k] + offBn[n]];
        Bs_r[k][n] = v.x;
        Bs_i[k][n] = v.y;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        R ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          ar[i] = As_r[kk + lk][wm + 16 * i + li];
          ai[i] = As_i[kk + lk][wm + 16 * i + li];
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          br[j] = Bs_r[kk + lk][wn + 16 * j + li];
          bi[j] = Bs_i[kk + lk][wn + 16 * j + li];
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            typename MF::acc& xr = CHUNKED ? cr[i][j] : accr[i][j];
            typename MF::acc& xi = CHUNKED ? ci[i][j] : acci[i][j];
            xr = MF::op(ar[i], br[j], xr);
            xr = MF::op(-ai[i], bi[j], xr);
            xi = MF::op(ar[i], bi[j], xi);
            xi = MF::op(ai[i], br[j], xi);
          }
      }
      if constexpr (CHUNKED) add_chunk<MF, TM, TN>(accr, acci, cr, ci);
      __syncthreads();
    }
  }

  // Epilogue: accumulator register r of tile (i, j) holds D[MF::row(lk, r)][li].
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = wm + 16 * i + MF::row(lk, r);
        int n = wn + 16 * j + li;
        if (m0 + m < Mt && n0 + n < Nt) {
          C v;
          v.x = accr[i][j][r];
          v.y = acci[i][j][r];
          Out[offCm[m] + offCn[n]] = v;
        }
      }
}

// ------------------------------------------------------------------ k_gemm2
// Fast path when the summed indices run contiguously in both operands
// (first K mode has unit stride in A and in B and its extent is a multiple
// of BKT) and every offset keeps 16-byte alignment: each operand row (A) or
// column (B) of a K tile is one contiguous run, read with 16-byte vector
// loads.  The next tile is prefetched into registers while the matrix
// cores work on the current one (register double buffering).
//
// LDS holds the tiles as interleaved complex values, row-major [row][k]
// with an odd pitch of BKT + 1 elements: a lane's (re, im) operand pair is
// one 8-byte (16-byte for complex128) read, and the 16 lanes of an MFMA
// operand row fall on distinct banks.
template <typename C, int BM, int BN, int WGM, int BKT>
__global__ __launch_bounds__(256) void k_gemm2(const C* __restrict__ A, const C* __restrict__ B,
                                               C* __restrict__ Out, GemmArgs g) {
  typedef typename real_of<C>::type R;
  typedef mfma_t<R> MF;
  constexpr int WGN = 4 / WGM;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be a multiple of 16");
  constexpr bool CHUNKED = sizeof(R) == 4;
  constexpr int VEC = 16 / (int)sizeof(C);       // complex elements per 16-byte chunk
  constexpr int CPR = BKT / VEC;                  // chunks per row of a K tile
  constexpr int ACH = (BM * CPR + 255) / 256;     // chunks per lane, A tile
  constexpr int BCH = (BN * CPR + 255) / 256;     // chunks per lane, B tile
  constexpr int KT_MAX = 2048 / BKT;              // K tiles (K <= 2048)
  constexpr int LDK = BKT + 1;                    // LDS row pitch (complex elements)

  __shared__ C As[BM * LDK + VEC];   // + a spare chunk written by idle lanes
  __shared__ C Bs[BN * LDK + VEC];
  __shared__ int32_t offAm[BM], offCm[BM], offBn[BN], offCn[BN];
  __shared__ int32_t offAk[KT_MAX], offBk[KT_MAX];

  const uint32_t Mt = g.M.total, Nt = g.N.total, Kt = g.K.total;
  uint32_t tile_x, tile_y;
  xcd_tile(tile_x, tile_y);
  const uint32_t m0 = tile_x * BM, n0 = tile_y * BN;
  const int tid = threadIdx.x;
  {
    int64_t oa, ob, oc;
    batch_offsets(g.L, blockIdx.z, oa, ob, oc);
    A += oa;
    B += ob;
    Out += oc;
  }
  for (int t = tid; t < BM; t += 256) {
    int32_t oa = 0, oc = 0;
    if (m0 + t < Mt) group_offsets(g.M, m0 + t, oa, oc);
    offAm[t] = oa;
    offCm[t] = oc;
  }
  for (int t = tid; t < BN; t += 256) {
    int32_t ob = 0, oc = 0;
    if (n0 + t < Nt) group_offsets(g.N, n0 + t, ob, oc);
    offBn[t] = ob;
    offCn[t] = oc;
  }
  // K tile offsets: a table when the tiles fit it, else decoded per tile
  // (the same value in every lane)
  const uint32_t ntiles = Kt / BKT;
  const bool ktab = ntiles <= (uint32_t)KT_MAX;
  if (ktab) {
    for (uint32_t t = tid; t < ntiles; t += 256) {
      int32_t oa, ob;
      group_offsets(g.K, t * BKT, oa, ob);
      offAk[t] = oa;
      offBk[t] = ob;
    }
  }
  __syncthreads();

  // This lane's chunks of each tile: global offset (fixed over K) and LDS
  // position.  Loads and stores are unconditional: rows past the operand's
  // extent read its first row (offAm / offBn hold 0 there), and their
  // products only reach outputs that are never stored; lanes past the tile
  // (when 256 does not divide its chunk count) write a spare LDS chunk.
  int32_t ga[ACH], gb[BCH], la[ACH], lb[BCH];
#pragma unroll
  for (int q = 0; q < ACH; ++q) {
    const int id = tid + 256 * q;
    const int row = id / CPR, cc = id % CPR;
    ga[q] = (row < BM ? offAm[row] : 0) + cc * VEC;
    la[q] = row < BM ? row * LDK + cc * VEC : BM * LDK;
  }
#pragma unroll
  for (int q = 0; q < BCH; ++q) {
    const int id = tid + 256 * q;
    const int col = id / CPR, cc = id % CPR;
    gb[q] = (col < BN ? offBn[col] : 0) + cc * VEC;
    lb[q] = col < BN ? col * LDK + cc * VEC : BN * LDK;
  }

  const int wave = tid / 64, lane = tid % 64;
  const int wm = (wave / WGN) * WM, wn = (wave % WGN) * WN;
  const int li = lane % 16, lk = lane / 16;

  typename MF::acc accr[TM][TN], acci[TM][TN];
  zero_acc<MF, TM, TN>(accr, acci);

  // 16-byte staging registers: two complex64 or one complex128 per chunk
  typedef typename std::conditional<sizeof(C) == 8, float4, double2>::type V16;
  V16 ra[ACH], rb[BCH];
  auto load_tile = [&](uint32_t kt) {
    int32_t ka, kb;
    if (ktab) {
      ka = offAk[kt];
      kb = offBk[kt];
    } else {
      group_offsets(g.K, kt * BKT, ka, kb);
    }
#pragma unroll
    for (int q = 0; q < ACH; ++q) ra[q] = *reinterpret_cast<const V16*>(A + ga[q] + ka);
#pragma unroll
    for (int q = 0; q < BCH; ++q) rb[q] = *reinterpret_cast<const V16*>(B + gb[q] + kb);
  };
  auto put = [](C* dst, const V16& v) {
    if constexpr (VEC == 2) {
      C x, y;
      x.x = v.x;
      x.y = v.y;
      y.x = v.z;
      y.y = v.w;
      dst[0] = x;
      dst[1] = y;
    } else {
      C x;
      x.x = v.x;
      x.y = v.y;
      dst[0] = x;
    }
  };
  auto store_tile = [&]() {
#pragma unroll
    for (int q = 0; q < ACH; ++q) put(As + la[q], ra[q]);
#pragma unroll
    for (int q = 0; q < BCH; ++q) put(Bs + lb[q], rb[q]);
  };

  const C* Ap = As + (wm + li) * LDK + lk;
  const C* Bp = Bs + (wn + li) * LDK + lk;
  const typename MF::acc zero = {};
  load_tile(0);
  for (uint32_t kt = 0; kt < ntiles; ++kt) {
    typename MF::acc cr[TM][TN], ci[TM][TN];
    store_tile();
    __syncthreads();
    if (kt + 1 < ntiles) load_tile(kt + 1);
#pragma unroll
    for (int kk = 0; kk < BKT; kk += 4) {
      C a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = Ap[16 * i * LDK + kk];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bp[16 * j * LDK + kk];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          typename MF::acc& xr = CHUNKED ? cr[i][j] : accr[i][j];
          typename MF::acc& xi = CHUNKED ? ci[i][j] : acci[i][j];
          // a K tile's first products start the tile's fresh sums
          const bool fresh = CHUNKED && kk == 0;
          xr = MF::op(a[i].x, b[j].x, fresh ? zero : xr);
          xr = MF::op(-a[i].y, b[j].y, xr);
          xi = MF::op(a[i].x, b[j].y, fresh ? zero : xi);
          xi = MF::op(a[i].y, b[j].x, xi);
        }
    }
    if constexpr (CHUNKED) add_chunk<MF, TM, TN>(accr, acci, cr, ci);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = wm + 16 * i + MF::row(lk, r);
        int n = wn + 16 * j + li;
        if (m0 + m < Mt && n0 + n < Nt) {
          C v;
          v.x = accr[i][j][r];
          v.y = acci[i][j][r];
          Out[offCm[m] + offCn[n]] = v;
        }
      }
}

// K-tile depth of the fast path: 32 for complex64 tiles whose LDS allows it.
template <typename C, int BM, int BN>
constexpr int fast_bk() {
  // 32 when both tiles at pitch 33 plus the offset tables fit 80 KiB (two
  // workgroups per CU's 160 KiB); 64-deep tiles measured slower (their
  // staging registers halve the occupancy)
  constexpr size_t es = sizeof(C), tab = 16 * (BM + BN) + 2 * 4 * 2048 / 8;
  return (BM + BN + 2) * 33 * es + tab <= 80 * 1024 ? 32 : 16;
}

// Static LDS stays within 64 KiB: complex128 tiles up to 128 rows or
// columns and 4096 elements.
template <typename C, int BM, int BN>
constexpr bool tile_fits() {
  return sizeof(C) == 8 || (BM * BN <= 64 * 64 && BM <= 128 && BN <= 128);
}

template <typename C, int BM, int BN, int WGM>
hipError_t launch_tile(const C* A, const C* B, C* Out, const GemmArgs& g, hipStream_t s) {
  if constexpr (!tile_fits<C, BM, BN>()) {
    return launch_tile<C, (BM > 64 ? 64 : BM), (BN > 64 ? 64 : BN), WGM>(A, B, Out, g, s);
  } else {
    dim3 grid((g.M.total + BM - 1) / BM, (g.N.total + BN - 1) / BN, g.L.total);
    constexpr int BKF = fast_bk<C, BM, BN>();
    // K tile depth: the deepest of 32 / 16 / 8 (within the tile's LDS
    // budget) that divides the first (contiguous) K mode, so a tile never
    // straddles two runs.
    const uint32_t k0ext = g.K.nmodes > 0 ? g.K.ext[0].d : 0;
    if (g.contig && k0ext > 0) {
      if (k0ext % BKF == 0) {
        hipLaunchKernelGGL((k_gemm2<C, BM, BN, WGM, BKF>), grid, dim3(256), 0, s, A, B, Out, g);
        return hipGetLastError();
      }
      if (BKF > 16 && k0ext % 16 == 0) {
        hipLaunchKernelGGL((k_gemm2<C, BM, BN, WGM, 16>), grid, dim3(256), 0, s, A, B, Out, g);
        return hipGetLastError();
      }
      if (k0ext % 8 == 0) {
        hipLaunchKernelGGL((k_gemm2<C, BM, BN, WGM, 8>), grid, dim3(256), 0, s, A, B, Out, g);
        return hipGetLastError();
      }
    }
    if (g.a_kfast && g.b_kfast)
      hipLaunchKernelGGL((k_gemm<C, BM, BN, WGM, true, true>), grid, dim3(256), 0, s, A, B, Out, g);
    else if (g.a_kfast)
      hipLaunchKernelGGL((k_gemm<C, BM, BN, WGM, true, false>), grid, dim3(256), 0, s, A, B, Out, g);
    else if (g.b_kfast)
      hipLaunchKernelGGL((k_gemm<C, BM, BN, WGM, false, true>), grid, dim3(256), 0, s, A, B, Out, g);
    else
      hipLaunchKernelGGL((k_gemm<C, BM, BN, WGM, false, false>), grid, dim3(256), 0, s, A, B, Out, g);
    return hipGetLastError();
  }
}

// ----------------------------------------------------------------- k_gemm_h
// complex64 GEMM on FP16 matrix cores with a hi/lo split (the fast path's
// contiguous-K layout; K tiles of 32).  Per K tile every row of A (column of
// B) gets a power-of-two scale 2^s that brings its largest magnitude into
// [2^14, 2^15); each scaled FP32 value x is split into hi = fp16(x) and
// lo = fp16(x - hi), and the product is summed as hi*hi + hi*lo + lo*hi
// (lo*lo dropped, about 2^-22 relative) on v_mfma_f32_16x16x32_f16 with FP32
// accumulation: 12 FP16 MFMAs per 16x16 complex block and K tile.  The
// tile's sums are unscaled exactly (ldexp) and added to the running FP32
// sums, as in the FP32 path's two-level accumulation.
//
// MFMA 16x16x32 f16 lane layout (64 lanes): A lane l holds A[l%16][8(l/16)+i],
// B lane l holds B[8(l/16)+i][l%16], accumulator r holds D[4(l/16)+r][l%16].
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ floatx4 mfma16(const half8& a, const half8& b, const floatx4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

template <int BM, int BN, int WGM>
__global__ __launch_bounds__(256) void k_gemm_h(const c64* __restrict__ A, const c64* __restrict__ B,
                                                c64* __restrict__ Out, GemmArgs g) {
  constexpr int BKT = 32;                         // complex K per tile = one MFMA depth
  constexpr int WGN = 4 / WGM;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be a multiple of 16");
  constexpr int CPR = BKT / 2;                    // 16-byte chunks (2 complex) per row
  constexpr int ACH = (BM * CPR + 255) / 256;
  constexpr int BCH = (BN * CPR + 255) / 256;
  constexpr int KT_MAX = 2048 / BKT;
  constexpr int PK = BKT + 8;                     // plane pitch in halves (16-byte aligned rows)

  // planes: 0 re-hi, 1 re-lo, 2 im-hi, 3 im-lo; one spare row for idle lanes
  __shared__ __align__(16) _Float16 Ap[4][(BM + 1) * PK];
  __shared__ __align__(16) _Float16 Bp[4][(BN + 1) * PK];
  __shared__ __align__(16) int32_t sA[BM + 1], sB[BN + 1];
  __shared__ int32_t offAm[BM], offCm[BM], offBn[BN], offCn[BN];
  __shared__ int32_t offAk[KT_MAX], offBk[KT_MAX];

  const uint32_t Mt = g.M.total, Nt = g.N.total, Kt = g.K.total;
  uint32_t tile_x, tile_y;
  xcd_tile(tile_x, tile_y);
  const uint32_t m0 = tile_x * BM, n0 = tile_y * BN;
  const int tid = threadIdx.x;
  {
    int64_t oa, ob, oc;
    batch_offsets(g.L, blockIdx.z, oa, ob, oc);
    A += oa;
    B += ob;
    Out += oc;
  }
  for (int t = tid; t < BM; t += 256) {
    int32_t oa = 0, oc = 0;
    if (m0 + t < Mt) group_offsets(g.M, m0 + t, oa, oc);
    offAm[t] = oa;
    offCm[t] = oc;
  }
  for (int t = tid; t < BN; t += 256) {
    int32_t ob = 0, oc = 0;
    if (n0 + t < Nt) group_offsets(g.N, n0 + t, ob, oc);
    offBn[t] = ob;
    offCn[t] = oc;
  }
  const uint32_t ntiles = Kt / BKT;
  const bool ktab = ntiles <= (uint32_t)KT_MAX;
  if (ktab) {
    for (uint32_t t = tid; t < ntiles; t += 256) {
      int32_t oa, ob;
      group_offsets(g.K, t * BKT, oa, ob);
      offAk[t] = oa;
      offBk[t] = ob;
    }
  }
  __syncthreads();

  // Lane chunks: 16 consecutive lanes hold the 16 chunks of one row, so a
  // row's scale is a reduction over a 16-lane group.  Rows past the
  // operand read its first row; lanes past the tile use the spare row.
  int32_t ga[ACH], gb[BCH], la[ACH], lb[BCH], ra_row[ACH], rb_row[BCH];
#pragma unroll
  for (int q = 0; q < ACH; ++q) {
    const int id = tid + 256 * q;
    const int row = id / CPR, cc = id % CPR;
    ga[q] = (row < BM ? offAm[row] : 0) + cc * 2;
    ra_row[q] = row < BM ? row : BM;
    la[q] = ra_row[q] * PK + cc * 2;
  }
#pragma unroll
  for (int q = 0; q < BCH; ++q) {
    const int id = tid + 256 * q;
    const int col = id / CPR, cc = id % CPR;
    gb[q] = (col < BN ? offBn[col] : 0) + cc * 2;
    rb_row[q] = col < BN ? col : BN;
    lb[q] = rb_row[q] * PK + cc * 2;
  }
  const bool lead = (tid % CPR) == 0;

  const int wave = tid / 64, lane = tid % 64;
  const int wm = (wave / WGN) * WM, wn = (wave % WGN) * WN;
  const int li = lane % 16, lk = lane / 16;

  floatx4 accr[TM][TN], acci[TM][TN];
  zero_acc<mfma_t<float>, TM, TN>(accr, acci);

  float4 ra[ACH], rb[BCH];
  auto load_tile = [&](uint32_t kt) {
    int32_t ka, kb;
    if (ktab) {
      ka = offAk[kt];
      kb = offBk[kt];
    } else {
      group_offsets(g.K, kt * BKT, ka, kb);
    }
#pragma unroll
    for (int q = 0; q < ACH; ++q) ra[q] = *reinterpret_cast<const float4*>(A + ga[q] + ka);
#pragma unroll
    for (int q = 0; q < BCH; ++q) rb[q] = *reinterpret_cast<const float4*>(B + gb[q] + kb);
  };
  // Scale, split and store one chunk (2 complex) of a row into the planes.
  auto split_store = [&](_Float16* p0, _Float16* p1, _Float16* p2, _Float16* p3, int at, const float4& v,
                         int* srow, int row, bool lead_lane) {
    float m = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
    m = fmaxf(m, __shfl_xor(m, 8, 16));
    m = fmaxf(m, __shfl_xor(m, 4, 16));
    m = fmaxf(m, __shfl_xor(m, 2, 16));
    m = fmaxf(m, __shfl_xor(m, 1, 16));
    int e = 0;
    (void)frexpf(m, &e);
    const int sc = m > 0.f ? 15 - e : 0;         // max * 2^sc in [2^14, 2^15)
    const float xr0 = ldexpf(v.x, sc), xi0 = ldexpf(v.y, sc), xr1 = ldexpf(v.z, sc), xi1 = ldexpf(v.w, sc);
    half2v hr, lr, hi, li2;
    hr.x = (_Float16)xr0;
    hr.y = (_Float16)xr1;
    hi.x = (_Float16)xi0;
    hi.y = (_Float16)xi1;
    lr.x = (_Float16)(xr0 - (float)hr.x);
    lr.y = (_Float16)(xr1 - (float)hr.y);
    li2.x = (_Float16)(xi0 - (float)hi.x);
    li2.y = (_Float16)(xi1 - (float)hi.y);
    *reinterpret_cast<half2v*>(p0 + at) = hr;
    *reinterpret_cast<half2v*>(p1 + at) = lr;
    *reinterpret_cast<half2v*>(p2 + at) = hi;
    *reinterpret_cast<half2v*>(p3 + at) = li2;
    if (lead_lane) srow[row] = sc;
  };

  load_tile(0);
  for (uint32_t kt = 0; kt < ntiles; ++kt) {
#pragma unroll
    for (int q = 0; q < ACH; ++q) split_store(Ap[0], Ap[1], Ap[2], Ap[3], la[q], ra[q], sA, ra_row[q], lead);
#pragma unroll
    for (int q = 0; q < BCH; ++q) split_store(Bp[0], Bp[1], Bp[2], Bp[3], lb[q], rb[q], sB, rb_row[q], lead);
    __syncthreads();
    if (kt + 1 < ntiles) load_tile(kt + 1);
    half8 a[TM][4], an[TM][2], b[TN][4];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int at = (wm + 16 * i + li) * PK + 8 * lk;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) a[i][pl] = *reinterpret_cast<const half8*>(&Ap[pl][at]);
      an[i][0] = -a[i][2];
      an[i][1] = -a[i][3];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int at = (wn + 16 * j + li) * PK + 8 * lk;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) b[j][pl] = *reinterpret_cast<const half8*>(&Bp[pl][at]);
    }
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      // row scales of this lane's four output rows
      const int4 sa = *reinterpret_cast<const int4*>(&sA[wm + 16 * i + 4 * lk]);
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        floatx4 cr = mfma16(a[i][0], b[j][1], zero);      // re: hi*lo + lo*hi + hi*hi
        cr = mfma16(a[i][1], b[j][0], cr);
        cr = mfma16(an[i][0], b[j][3], cr);                // - im * im
        cr = mfma16(an[i][1], b[j][2], cr);
        cr = mfma16(a[i][0], b[j][0], cr);
        cr = mfma16(an[i][0], b[j][2], cr);
        floatx4 ci = mfma16(a[i][0], b[j][3], zero);       // im: re*im + im*re
        ci = mfma16(a[i][1], b[j][2], ci);
        ci = mfma16(a[i][2], b[j][1], ci);
        ci = mfma16(a[i][3], b[j][0], ci);
        ci = mfma16(a[i][0], b[j][2], ci);
        ci = mfma16(a[i][2], b[j][0], ci);
        const int sb = sB[wn + 16 * j + li];
        accr[i][j][0] += ldexpf(cr[0], -(sa.x + sb));
        acci[i][j][0] += ldexpf(ci[0], -(sa.x + sb));
        accr[i][j][1] += ldexpf(cr[1], -(sa.y + sb));
        acci[i][j][1] += ldexpf(ci[1], -(sa.y + sb));
        accr[i][j][2] += ldexpf(cr[2], -(sa.z + sb));
        acci[i][j][2] += ldexpf(ci[2], -(sa.z + sb));
        accr[i][j][3] += ldexpf(cr[3], -(sa.w + sb));
        acci[i][j][3] += ldexpf(ci[3], -(sa.w + sb));
      }
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int m = wm + 16 * i + 4 * lk + r;
        int n = wn + 16 * j + li;
        if (m0 + m < Mt && n0 + n < Nt) {
          c64 v;
          v.x = accr[i][j][r];
          v.y = acci[i][j][r];
          Out[offCm[m] + offCn[n]] = v;
        }
      }
}

template <int BM, int BN, int WGM>
hipError_t launch_tile_h(const c64* A, const c64* B, c64* Out, const GemmArgs& g, hipStream_t s) {
  dim3 grid((g.M.total + BM - 1) / BM, (g.N.total + BN - 1) / BN, g.L.total);
  hipLaunchKernelGGL((k_gemm_h<BM, BN, WGM>), grid, dim3(256), 0, s, A, B, Out, g);
  return hipGetLastError();
}

template <typename C>
__global__ __launch_bounds__(256) void k_sum_partials(const C* P, C* Out, const Group gm, const Group gn,
                                                      uint32_t mt, uint64_t mn, int parts) {
  const uint64_t o = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= mn) return;
  int32_t pm, cm, pn, cn;
  group_offsets(gm, (uint32_t)(o % mt), pm, cm);
  group_offsets(gn, (uint32_t)(o / mt), pn, cn);
  double re = 0.0, im = 0.0;
  for (int z = 0; z < parts; ++z) {
    const C v = P[(uint64_t)z * mn + o];
    re += (double)v.x;
    im += (double)v.y;
  }
  C r;
  r.x = (decltype(r.x))re;
  r.y = (decltype(r.y))im;
  Out[(int64_t)cm + cn] = r;
}

}  // namespace

int gemm_k_limit() { return (int)GEMM_K_MAX; }

template <typename C>
hipError_t launch_sum_partials(const C* P, C* Out, const Group& gm, const Group& gn, uint32_t mt, uint64_t mn,
                               int parts, hipStream_t s) {
  if (mt < 1 || mn < 1 || parts < 1 || mn % mt != 0 || (uint64_t)gm.total != mt || (uint64_t)gn.total != mn / mt)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_sum_partials<C>), dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, P, Out, gm, gn, mt,
                     mn, parts);
  return hipGetLastError();
}
template hipError_t launch_sum_partials<c128>(const c128*, c128*, const Group&, const Group&, uint32_t, uint64_t,
                                              int, hipStream_t);
template hipError_t launch_sum_partials<c64>(const c64*, c64*, const Group&, const Group&, uint32_t, uint64_t, int,
                                             hipStream_t);

template <typename C>
hipError_t launch_gemm(const C* A, const C* B, C* Out, const GemmArgs& g, hipStream_t s) {
  if (g.K.total > GEMM_K_MAX || g.L.total < 1 || g.L.total > 65535) return hipErrorInvalidValue;
  switch (g.tile) {
    case 0: return launch_tile<C, 32, 32, 2>(A, B, Out, g, s);
    case 1: return launch_tile<C, 64, 32, 2>(A, B, Out, g, s);
    case 2: return launch_tile<C, 32, 64, 2>(A, B, Out, g, s);
    case 3: return launch_tile<C, 64, 64, 2>(A, B, Out, g, s);
    case 4: return launch_tile<C, 128, 64, 2>(A, B, Out, g, s);
    case 5: return launch_tile<C, 64, 128, 2>(A, B, Out, g, s);
    case 6: return launch_tile<C, 64, 16, 4>(A, B, Out, g, s);
    case 7: return launch_tile<C, 128, 16, 4>(A, B, Out, g, s);
    case 8: return launch_tile<C, 256, 16, 4>(A, B, Out, g, s);
    case 9: return launch_tile<C, 16, 64, 1>(A, B, Out, g, s);
    case 10: return launch_tile<C, 128, 32, 4>(A, B, Out, g, s);
    case 11: return launch_tile<C, 128, 128, 2>(A, B, Out, g, s);
    case 12: return launch_tile<C, 16, 128, 1>(A, B, Out, g, s);
    case 13: return launch_tile<C, 16, 256, 1>(A, B, Out, g, s);
    case 14: return launch_tile<C, 32, 128, 1>(A, B, Out, g, s);
    default: return hipErrorInvalidValue;
  }
}

bool gemm16_applicable(const GemmArgs& g) {
  return g.contig && g.K.nmodes > 0 && g.K.ext[0].d % 32 == 0 && g.K.total <= GEMM_K_MAX &&
         g.L.total >= 1 && g.L.total <= 65535;
}

hipError_t launch_gemm16(const c64* A, const c64* B, c64* Out, const GemmArgs& g, hipStream_t s) {
  if (!gemm16_applicable(g)) return hipErrorInvalidValue;
  switch (g.tile) {
    case 0: return launch_tile_h<32, 32, 2>(A, B, Out, g, s);
    case 1: return launch_tile_h<64, 32, 2>(A, B, Out, g, s);
    case 2: return launch_tile_h<32, 64, 2>(A, B, Out, g, s);
    case 4: return launch_tile_h<128, 64, 2>(A, B, Out, g, s);
    case 5: return launch_tile_h<64, 128, 2>(A, B, Out, g, s);
    case 6: return launch_tile_h<64, 16, 4>(A, B, Out, g, s);
    case 7: return launch_tile_h<128, 16, 4>(A, B, Out, g, s);
    case 9: return launch_tile_h<16, 64, 1>(A, B, Out, g, s);
    case 10: return launch_tile_h<128, 32, 4>(A, B, Out, g, s);
    case 12: return launch_tile_h<16, 128, 1>(A, B, Out, g, s);
    case 14: return launch_tile_h<32, 128, 1>(A, B, Out, g, s);
    default: return launch_tile_h<64, 64, 2>(A, B, Out, g, s);     // 3, and 8 / 11 / 13 (LDS)
  }
}

template hipError_t launch_gemm<c128>(const c128*, const c128*, c128*, const GemmArgs&, hipStream_t);
template hipError_t launch_gemm<c64>(const c64*, const c64*, c64*, const GemmArgs&, hipStream_t);
