// Fused chain of small operators on one large complex64 tensor.
//
// A run of pairwise steps (src/layer2.jl:346-405 applied along a plan) in
// which each step contracts a small operand S_t (at most 2^5 x 2^5
// elements) with the previous step's result rewrites a few bits of one big
// tensor per step and is HBM-bound: every step reads and writes the whole
// tensor.  Here the tensor's modes are split into bits; the bits any
// operator touches form a tile of at most 2^13 elements.  One workgroup per
// configuration of the other ("spectator") modes:
//   1. gathers its tile from X into LDS (bits ordered by X stride, so
//      adjacent lanes read adjacent addresses),
//   2. applies the operators in order, in place:
//        T[r, m] = sum_k S_t[m, k] T[r, k]
//      as 16 x 16 x 4 complex blocks on v_mfma_f32_16x16x4_f32 (4 real
//      MFMAs per block: Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br),
//   3. scatters the tile to Y (bits ordered by Y stride).
// Tile positions of every bit, per stage, come from the host program
// (pq_hip.h, PQ_CHAIN_*); in place is safe because a wave reads all K
// values of its 16 r's before it writes their M values, and different
// waves own different r's (distinct R bits).
#include "pq_internal.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// LDS bank spreading: the tile is stored at swz(e), a bijection that is
// linear over GF(2) (the low 5 bits are XORed with bits 5-9 and 10-14), so
// swz(a | b) = swz(a) ^ swz(b) for addresses built from disjoint bit sets.
// 16 lanes whose addresses differ by multiples of any power of two up to
// 2^11 then fall on 16 distinct 8-byte bank slots.
__device__ __forceinline__ int swz(int e) { return e ^ ((e >> 5) & 31) ^ ((e >> 10) & 31); }

__device__ __forceinline__ floatx4 mfma(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Tile <-> tensor copy: element j = tid + 2^lo it (lo <= 9 bits for the
// CH_THREADS = 512 lanes); lane tid takes the low tables' entry tid,
// iteration it the high tables' entry it.
template <bool GATHER>
__device__ __forceinline__ void move_tile(c64* tile, const c64* X, c64* Y, int32_t base, const int32_t* tab,
                                          int lo, int hi, int lp, int lg, int hp, int hg, int tid) {
  if (tid >= (1 << lo)) return;
  const int p0 = swz(tab[lp + tid]);
  const int32_t g0 = base + tab[lg + tid];
  const int nit = 1 << hi;
  for (int it = 0; it < nit; ++it) {
    const int p = p0 ^ swz(tab[hp + it]);
    const int32_t g = g0 + tab[hg + it];
    if (GATHER)
      tile[p] = X[g];
    else
      Y[g] = tile[p];
  }
}

// One stage, T[r, m] = sum_k S[m, k] T[r, k], with NKC chunks of 4 K values
// and NMB blocks of 16 M values (K = 2^nk <= 4 NKC, M = 2^nm <= 16 NMB).
// Wave w of the CH_THREADS / 64 takes the blocks of 16 r's numbered w,
// w + CH_THREADS / 64, ...
template <int NKC, int NMB>
__device__ __forceinline__ void stage(c64* tile, const c64* sd, const int* si_, const ChainStage& g, const int32_t* tab,
                                      int wave, int li, int lg) {
  const int nk = 1 << g.nk, nm = 1 << g.nm;
  // This lane's operator values (MFMA A operand: row m = 16 mb + li,
  // column k = 4 kc + lg), zero outside the operator; its K and M tile
  // offsets (B operand row k = 4 kc + lg; output row m = 16 mb + 4 lg + i);
  // the tile offset of its column li within a block of 16 r's.
  float sr[NKC][NMB], si[NKC][NMB], sn[NKC][NMB];
  int kof[NKC], mof[NMB][4];
  // operator values and offsets come from the LDS staging area
#pragma unroll
  for (int kc = 0; kc < NKC; ++kc) {
    const int k = 4 * kc + lg;
    const bool kv = k < nk;
    kof[kc] = kv ? si_[g.ko + k] : 0;
#pragma unroll
    for (int mb = 0; mb < NMB; ++mb) {
      const int m = 16 * mb + li;
      c64 v;
      v.x = 0.f;
      v.y = 0.f;
      if (kv && m < nm) v = sd[g.sd + m * nk + k];
      sr[kc][mb] = v.x;
      si[kc][mb] = v.y;
      sn[kc][mb] = -v.y;
    }
  }
#pragma unroll
  for (int mb = 0; mb < NMB; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = 16 * mb + 4 * lg + i;
      mof[mb][i] = m < nm ? si_[g.mo + m] : -1;
    }
  const bool rv = li < (1 << g.rlo);
  const int rlo = si_[g.ro + (rv ? li : 0)];
  const bool kpart = NKC == 1 && nk < 4;   // fewer K values than lane groups
  // every lane stores all its outputs (no partial R or M block)
  const bool full = (1 << g.rlo) == 16 && nm >= 16 * NMB;
  const int nrb = 1 << g.rhi;
  constexpr int NW = CH_THREADS / 64;
  // the next block's high R offset is fetched one iteration ahead
  int rh = wave < nrb ? tab[g.rh + wave] : 0;
  for (int rb = wave; rb < nrb; rb += NW) {
    const int roff = rlo ^ swz(rh);
    if (rb + NW < nrb) rh = tab[g.rh + rb + NW];
    // every K value of these r's is read before any M value is written
    float br[NKC], bi[NKC];
#pragma unroll
    for (int kc = 0; kc < NKC; ++kc) {
      c64 v = tile[roff ^ kof[kc]];
      if (kpart && lg >= nk) {
        v.x = 0.f;
        v.y = 0.f;
      }
      br[kc] = v.x;
      bi[kc] = v.y;
    }
#pragma unroll
    for (int mb = 0; mb < NMB; ++mb) {
      floatx4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < NKC; ++kc) {
        cr = mfma(sr[kc][mb], br[kc], cr);
        ci = mfma(sr[kc][mb], bi[kc], ci);
        cr = mfma(sn[kc][mb], bi[kc], cr);
        ci = mfma(si[kc][mb], br[kc], ci);
      }
      // accumulator register i holds row 4 lg + i, column li
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (full || (rv && mof[mb][i] >= 0)) {
          c64 v;
          v.x = cr[i];
          v.y = ci[i];
          tile[roff ^ mof[mb][i]] = v;
        }
      }
    }
  }
}

__global__ __launch_bounds__(CH_THREADS) void k_chain(const ChainArgs a) {
  extern __shared__ c64 tile[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int32_t* tab = a.tab;
  int32_t ox, oy;
  group_offsets(a.spec, blockIdx.x, ox, oy);

  // stage every operator (dense [m][k]) and the swizzled K / M / low-R
  // offsets in LDS after the tile; the loads overlap the gather's
  c64* sd = tile + (1 << a.tb);
  int* sidx = reinterpret_cast<int*>(sd + a.sdn);
  for (int s = 0; s < a.nst; ++s) {
    const ChainStage& g = a.st[s];
    const int nk = 1 << g.nk, nkm = 1 << (g.nk + g.nm);
    for (int e = tid; e < nkm; e += CH_THREADS) {
      const int m = e >> g.nk, k = e & (nk - 1);
      sd[g.sd + e] = g.S[tab[g.sk + k] + tab[g.sm + m]];
    }
    for (int e = tid; e < nk; e += CH_THREADS) sidx[g.ko + e] = swz(tab[g.kp + e]);
    for (int e = tid; e < (1 << g.nm); e += CH_THREADS) sidx[g.mo + e] = swz(tab[g.mp + e]);
    for (int e = tid; e < (1 << g.rlo); e += CH_THREADS) sidx[g.ro + e] = swz(tab[g.rl + e]);
  }
  move_tile<true>(tile, a.X, nullptr, ox, tab, a.ilo, a.ihi, a.ilp, a.ilg, a.ihp, a.ihg, tid);
  __syncthreads();

  for (int s = 0; s < a.nst; ++s) {
    const ChainStage& g = a.st[s];
    const int nkc = ((1 << g.nk) + 3) / 4, nmb = ((1 << g.nm) + 15) / 16;
    switch (nkc * 4 + nmb) {
      case 1 * 4 + 1: stage<1, 1>(tile, sd, sidx, g, tab, wave, li, lg); break;
      case 2 * 4 + 1: stage<2, 1>(tile, sd, sidx, g, tab, wave, li, lg); break;
      case 4 * 4 + 1: stage<4, 1>(tile, sd, sidx, g, tab, wave, li, lg); break;
      case 8 * 4 + 1: stage<8, 1>(tile, sd, sidx, g, tab, wave, li, lg); break;
      case 1 * 4 + 2: stage<1, 2>(tile, sd, sidx, g, tab, wave, li, lg); break;
      case 2 * 4 + 2: stage<2, 2>(tile, sd, sidx, g, tab, wave, li, lg); break;
      case 4 * 4 + 2: stage<4, 2>(tile, sd, sidx, g, tab, wave, li, lg); break;
      default: stage<8, 2>(tile, sd, sidx, g, tab, wave, li, lg); break;
    }
    __syncthreads();
  }

  move_tile<false>(tile, nullptr, a.Y, oy, tab, a.olo, a.ohi, a.olp, a.olg, a.ohp, a.ohg, tid);
}

}  // namespace

hipError_t launch_chain(const ChainArgs& a, hipStream_t s) {
  if (a.tb < 0 || a.tb > CH_TB || a.nst < 1 || a.nst > CH_ST || a.spec.total < 1) return hipErrorInvalidValue;
  if ((int64_t)a.sdn * 8 + (int64_t)a.idn * 4 > CH_STAGE_BYTES) return hipErrorInvalidValue;
  const size_t lds = (sizeof(c64) << a.tb) + sizeof(c64) * a.sdn + sizeof(int) * a.idn;
  hipLaunchKernelGGL(k_chain, dim3(a.spec.total), dim3(CH_THREADS), lds, s, a);
  return hipGetLastError();
}
