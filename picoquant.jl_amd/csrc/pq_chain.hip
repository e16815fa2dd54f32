// Fused chain of small operators on one large complex64 tensor.
//
// A run of pairwise steps (src/layer2.jl:346-405 applied along a plan) in
// which each step contracts a small operand S_t (at most 2^5 x 2^5
// elements) with the previous step's result rewrites a few bits of one big
// tensor per step and is HBM-bound: every step reads and writes the whole
// tensor.  Here the tensor's modes are split into bits; the bits any
// operator touches form a tile of at most 2^13 elements, one tile per
// configuration of the other ("spectator") modes.  For each tile:
//   1. gather it from X into LDS (bits ordered by X stride, so adjacent
//      lanes read adjacent addresses),
//   2. apply the operators in order, in place:
//        T[r, m] = sum_k S_t[m, k] T[r, k]
//      as 16 x 16 x 4 complex blocks on v_mfma_f32_16x16x4_f32, with
//      three real products per complex one (Gauss): P1 = Ar Br, P2 = Ai Bi,
//      P3 = (Ar + Ai)(Br + Bi), each summed over K in an MFMA accumulator,
//      then Cr = P1 - P2, Ci = P3 - P1 - P2,
//   3. scatter it to Y (bits ordered by Y stride).
// Workgroups are persistent (two per CU, 8 waves each, one tile buffer
// each): while the stages run on one tile, the next tile's loads are in
// flight into registers, and the previous tile's stores drain, so HBM
// traffic overlaps the MFMA work instead of every workgroup on the chip
// gathering, computing and scattering in lockstep; a workgroup waiting at
// a stage barrier leaves the CU's matrix cores to the other.
// Tile positions of every bit, per stage, come from the host program
// (pq_hip.h, PQ_CHAIN_*); in place is safe because a wave reads all K
// values of its 16 r's before it writes their M values, and different
// waves own different r's (distinct R bits).
#include "pq_internal.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// A complex64 value in LDS: 8-byte aligned, so each lane moves it with one
// ds_read_b64 / ds_write_b64 (a 4-byte aligned pair becomes two 32-bit
// accesses, which put 32 lanes on only 16 of the banks).
struct __attribute__((aligned(8))) lc {
  float x, y;
};

// LDS addresses of tile elements come from the host's tables (pq_plan.hip,
// chain_layout): a linear map over GF(2) chosen per chain so that the lanes
// of every access below fall on distinct banks; addresses of disjoint bit
// sets combine by XOR.

// Tile addresses are byte offsets (the host's element addresses times 8,
// scaled when staged): an access is the XOR of table entries, with no shift.
__device__ __forceinline__ lc& at(lc* t, int b) { return *reinterpret_cast<lc*>(reinterpret_cast<char*>(t) + b); }
__device__ __forceinline__ const lc& at(const lc* t, int b) {
  return *reinterpret_cast<const lc*>(reinterpret_cast<const char*>(t) + b);
}

__device__ __forceinline__ floatx4 mfma(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// One stage, T[r, m] = sum_k S[m, k] T[r, k], with NKC chunks of 4 K values
// and NMB blocks of 16 M values (K = 2^nk <= 4 NKC, M = 2^nm <= 16 NMB).
// FULL: every lane writes all its outputs (16 low R values and M = 16 NMB),
// so the writes need no per-lane predicate.
// Wave w of the CH_THREADS / 64 takes the blocks of 16 r's numbered w,
// w + CH_THREADS / 64, ...  The lane's operator values (MFMA A operand:
// row m = 16 mb + li, column k = 4 kc + lg) stay in registers when there
// are at most 8 chunk-blocks of them; the 32 x 32 case rereads them from
// LDS per M block (register room for the next tile's prefetch).  With at
// most 4 chunk-blocks (16 x 16 and smaller, 8 -> 32), and with G8 (the
// one-M-block kernel form) the 32-column ones, a complex product takes
// three real MFMAs (Gauss), otherwise four: in the full kernel form the
// third accumulator and the sums of the larger operators cost registers,
// and measured slower.
template <int NKC, int NMB, bool FULL, bool G8 = false>
__device__ __forceinline__ void stage(lc* tile, const lc* sd, const int* si_, const ChainStage& g, int wave, int li,
                                      int lg) {
  constexpr bool HOLD = NKC * NMB <= 8;
  constexpr bool GAUSS = NKC * NMB <= 4 || G8;
  constexpr int HMB = HOLD ? NMB : 1;
  const int nk = 1 << g.nk, nm = 1 << g.nm;
  auto opval = [&](int kc, int mb) -> lc {
    const int k = 4 * kc + lg, m = 16 * mb + li;
    lc v;
    v.x = 0.f;
    v.y = 0.f;
    if (k < nk && m < nm) v = sd[g.sd + k * nm + m];
    return v;
  };
  float sr[NKC][HMB], si[NKC][HMB];
  int kof[NKC], mof[NMB][4];
#pragma unroll
  for (int kc = 0; kc < NKC; ++kc) {
    const int k = 4 * kc + lg;
    kof[kc] = k < nk ? si_[g.ko + k] : 0;
    if constexpr (HOLD) {
#pragma unroll
      for (int mb = 0; mb < NMB; ++mb) {
        const lc v = opval(kc, mb);
        sr[kc][mb] = v.x;
        si[kc][mb] = v.y;
      }
    }
  }
  // output row m = 16 mb + 4 lg + i of the lane's accumulator register i
#pragma unroll
  for (int mb = 0; mb < NMB; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = 16 * mb + 4 * lg + i;
      mof[mb][i] = m < nm ? si_[g.mo + m] : 0;
    }
  const bool rv = li < (1 << g.rlo);
  const int rlo = si_[g.ro + (rv ? li : 0)];
  const bool kpart = NKC == 1 && nk < 4;   // fewer K values than lane groups
  const int nrb = 1 << g.rhi;
  constexpr int NW = CH_THREADS / 64;
  // UNR blocks of 16 r's per iteration: the K values of all of them are
  // read first, so a wave pays the LDS latency once per iteration; then each
  // block's MFMAs run and its M values are written (blocks hold different
  // r's, so a block may be written before a later one is read).  A block
  // number past the end repeats the wave's first block and is not written.
  // 16 x 16 takes 2 blocks: with three accumulators per block, 4 spill.
  constexpr int UNR = (NKC <= 2 && NMB == 1) ? 4 : (NKC * NMB <= 4 ? 2 : 1);
  for (int rb0 = wave * UNR; rb0 < nrb; rb0 += NW * UNR) {
    int roff[UNR];
    float br[UNR][NKC], bi[UNR][NKC];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int rb = rb0 + u < nrb ? rb0 + u : rb0;
      roff[u] = rlo ^ si_[g.rho + rb];
#pragma unroll
      for (int kc = 0; kc < NKC; ++kc) {
        lc v = at(tile, roff[u] ^ kof[kc]);
        if (kpart && lg >= nk) {
          v.x = 0.f;
          v.y = 0.f;
        }
        br[u][kc] = v.x;
        bi[u][kc] = v.y;
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const bool wr = UNR == 1 || rb0 + u < nrb;
#pragma unroll
      for (int mb = 0; mb < NMB; ++mb) {
        if constexpr (!HOLD) {
#pragma unroll
          for (int kc = 0; kc < NKC; ++kc) {
            const lc v = opval(kc, mb);
            sr[kc][0] = v.x;
            si[kc][0] = v.y;
          }
        }
        constexpr int hm = 0;
        floatx4 cr, ci;
        if constexpr (GAUSS) {
          // three real products per complex one
          floatx4 p1 = {0.f, 0.f, 0.f, 0.f}, p2 = {0.f, 0.f, 0.f, 0.f}, p3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kc = 0; kc < NKC; ++kc) {
            p1 = mfma(sr[kc][mb], br[u][kc], p1);
            p2 = mfma(si[kc][mb], bi[u][kc], p2);
            p3 = mfma(sr[kc][mb] + si[kc][mb], br[u][kc] + bi[u][kc], p3);
          }
          cr = p1 - p2;
          ci = (p3 - p1) - p2;
        } else {
          // four real products per complex one
          cr = floatx4{0.f, 0.f, 0.f, 0.f};
          ci = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kc = 0; kc < NKC; ++kc) {
            const float ar = sr[kc][HOLD ? mb : hm], ai = si[kc][HOLD ? mb : hm];
            cr = mfma(ar, br[u][kc], cr);
            ci = mfma(ar, bi[u][kc], ci);
            cr = mfma(-ai, bi[u][kc], cr);
            ci = mfma(ai, br[u][kc], ci);
          }
        }
        // accumulator register i holds row 4 lg + i, column li
        if (wr) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (FULL || (rv && 16 * mb + 4 * lg + i < nm)) {
              // real and imaginary parts written as two 32-bit values (one
              // ds_write2_b32), no register pairing
              float* e = &at(tile, roff[u] ^ mof[mb][i]).x;
              e[0] = cr[i];
              e[1] = ci[i];
            }
          }
        }
      }
    }
  }
}

// Stages of at most 4 x 4 (K, M <= 4) with at least 64 R values (the host
// then takes 6 R bits from the lane): each lane owns one r (the lane's R
// table entry XOR its group's) and runs the whole product on
// v_mfma_f32_4x4x1f32, whose 16 blocks of 4 lanes take 4 r's each: lane
// 4b + j supplies T[r, k] as block b's B column j and S[j, k] as its A row
// j (the same operator in every block), and receives its r's M values in
// its 4 accumulator registers.  Three real products per complex one
// (Gauss), one instruction per K value: no padded MFMA rows, where a
// 16 x 16 x 4 block would leave 12 of its 16 rows unused.
__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void stage_rows(lc* tile, const lc* sd, const int* si_, const ChainStage& g, int wave,
                                           int lane) {
  const int nk = 1 << g.nk, nm = 1 << g.nm;
  const int m = lane & 3;
  float sr[4], si[4];
  int kof[4], mof[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lc v;
    v.x = 0.f;
    v.y = 0.f;
    if (k < nk && m < nm) v = sd[g.sd + k * nm + m];
    sr[k] = v.x;
    si[k] = v.y;
    kof[k] = k < nk ? si_[g.ko + k] : 0;
    mof[k] = k < nm ? si_[g.mo + k] : 0;
  }
  const int rlo = si_[g.ro + lane];
  const int ng = 1 << g.rhi;
  constexpr int NW = CH_THREADS / 64;
  for (int grp = wave; grp < ng; grp += NW) {
    const int roff = rlo ^ si_[g.rho + grp];
    float tr[4], ti[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      tr[k] = 0.f;
      ti[k] = 0.f;
      if (k < nk) {
        const lc v = at(tile, roff ^ kof[k]);
        tr[k] = v.x;
        ti[k] = v.y;
      }
    }
    floatx4 p1 = {0.f, 0.f, 0.f, 0.f}, p2 = {0.f, 0.f, 0.f, 0.f}, p3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k < nk) {
        p1 = mfma4(sr[k], tr[k], p1);
        p2 = mfma4(si[k], ti[k], p2);
        p3 = mfma4(sr[k] + si[k], tr[k] + ti[k], p3);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i < nm) {
        float* e = &at(tile, roff ^ mof[i]).x;
        e[0] = p1[i] - p2[i];
        e[1] = (p3[i] - p1[i]) - p2[i];
      }
    }
  }
}

// GV / SV: the gather / scatter moves each element pair as one 16-byte
// access (a.gvec / a.svec); separate instantiations keep the two address
// forms out of each other's code.  MODE 1: every stage's operator has at
// most 4 chunk-blocks (the three-product stages), so only those stage loops
// and the lane-row 4 x 4 stage are compiled in; MODE 2: every operator has
// one M block (up to 32 -> 16), and the 32-column stages take three
// products too; MODE 0: every stage form.  The register allocation of the
// whole kernel follows the stages it holds: the lane-row stage and the
// three-product 32 -> 8 stage fit only in the leaner forms without spills
// or slower code.  Four waves per SIMD (two workgroups per CU) is the
// design point: the register allocator is held to 128 VGPRs rather than
// trading the second workgroup for registers.
template <bool GV, bool SV, int MODE>
__global__ __launch_bounds__(CH_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_chain(const ChainArgs a) {
  extern __shared__ lc smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lg = lane >> 4;
  const int32_t* tab = a.tab;
  const uint32_t total = a.spec.total, G = gridDim.x;
  lc* tile = smem;

  // stage every operator (dense [k][m]: the 16 lanes of a lane group read
  // 16 consecutive m) and the K / M / R tile addresses in LDS after the
  // tile, once per workgroup
  lc* sd = smem + (1 << a.tb);
  int* sidx = reinterpret_cast<int*>(sd + a.sdn);
  for (int s = 0; s < a.nst; ++s) {
    const ChainStage& g = a.st[s];
    const int nk = 1 << g.nk, nkm = 1 << (g.nk + g.nm);
    for (int e = tid; e < nkm; e += CH_THREADS) {
      const int k = e >> g.nm, m = e & ((1 << g.nm) - 1);
      const c64 v = g.S[tab[g.sk + k] + tab[g.sm + m]];
      sd[g.sd + e] = lc{v.x, v.y};
    }
    for (int e = tid; e < nk; e += CH_THREADS) sidx[g.ko + e] = 8 * tab[g.kp + e];
    for (int e = tid; e < (1 << g.nm); e += CH_THREADS) sidx[g.mo + e] = 8 * tab[g.mp + e];
    for (int e = tid; e < (1 << g.rlo); e += CH_THREADS) sidx[g.ro + e] = 8 * tab[g.rl + e];
    for (int e = tid; e < (1 << g.rhi); e += CH_THREADS) sidx[g.rho + e] = 8 * tab[g.rh + e];
  }
  // Gather / scatter without branches: lane l, iteration it moves the
  // element pair (l + 2^lo it) along the side's pair bit; each lane does
  // CH_NIT iterations, and the staged iteration tables wrap around (entry
  // it holds iteration it mod 2^hi), as do the lanes beyond 2^lo.  A
  // repeated iteration rereads the same elements into the same tile
  // positions, or stores the same values to the same addresses.
  int* hid = sidx + a.idn;   // tile address / tensor offset per iteration
  if (tid < CH_NIT) {
    hid[tid] = 8 * tab[a.ihp + (tid & ((1 << a.ihi) - 1))];
    hid[CH_NIT + tid] = tab[a.ihg + (tid & ((1 << a.ihi) - 1))];
    hid[2 * CH_NIT + tid] = 8 * tab[a.ohp + (tid & ((1 << a.ohi) - 1))];
    hid[3 * CH_NIT + tid] = tab[a.ohg + (tid & ((1 << a.ohi) - 1))];
  }
  __syncthreads();
  const int gt = tid & ((1 << a.ilo) - 1), st = tid & ((1 << a.olo) - 1);
  const int gp0 = 8 * tab[a.ilp + gt], sp0 = 8 * tab[a.olp + st];
  const int gpa = 8 * a.gpa, spa = 8 * a.spa;
  const int32_t gg0 = tab[a.ilg + gt], sg0 = tab[a.olg + st];
  // wait for these four loads here (vmcnt(0), other counters untouched):
  // otherwise the compiler, merging the loop's entry and back edge, keeps
  // them pending and waits for every vector memory operation in flight
  // before their first use inside the loop
  __builtin_amdgcn_s_waitcnt(0x0F70);
  // spectator offsets (X, Y) of tile t: the host's per-tile table, or
  // decoded from the spectator modes (uniform across the workgroup)
  auto offsets = [&](uint32_t t, int32_t& ox, int32_t& oy) {
    if (a.tof >= 0) {
      // uniform address in the constant address space: a scalar load, which
      // waits on no vector memory operation
      const __attribute__((address_space(4))) int32_t* ct =
          (const __attribute__((address_space(4))) int32_t*)(tab + a.tof);
      ox = ct[2 * t];
      oy = ct[2 * t + 1];
    } else {
      group_offsets(a.spec, t, ox, oy);
    }
    ox = __builtin_amdgcn_readfirstlane(ox);
    oy = __builtin_amdgcn_readfirstlane(oy);
  };
  // Element accesses are a uniform tile base plus a 32-bit byte offset per
  // lane (chained tensors hold fewer than 2^29 elements, checked on the
  // host), so each access needs one address register.
  c64 P[2 * CH_NIT];   // the next tile, in flight (pairs)
  auto load = [&](int32_t ox) {
    const char* src = reinterpret_cast<const char*>(a.X + ox);
    uint32_t o[CH_NIT];
#pragma unroll
    for (int it = 0; it < CH_NIT; ++it) o[it] = (uint32_t)(gg0 + hid[CH_NIT + it]) * 8u;
    if constexpr (GV) {
#pragma unroll
      for (int it = 0; it < CH_NIT; ++it) {
        const float4 v = *reinterpret_cast<const float4*>(src + o[it]);
        P[2 * it].x = v.x;
        P[2 * it].y = v.y;
        P[2 * it + 1].x = v.z;
        P[2 * it + 1].y = v.w;
      }
    } else {
      const uint32_t ps = (uint32_t)a.gps * 8u;
#pragma unroll
      for (int it = 0; it < CH_NIT; ++it) {
        P[2 * it] = *reinterpret_cast<const c64*>(src + o[it]);
        P[2 * it + 1] = *reinterpret_cast<const c64*>(src + o[it] + ps);
      }
    }
  };
  auto put = [&](lc* buf) {
#pragma unroll
    for (int it = 0; it < CH_NIT; ++it) {
      const int q = gp0 ^ hid[it];
      at(buf, q) = lc{P[2 * it].x, P[2 * it].y};
      at(buf, q ^ gpa) = lc{P[2 * it + 1].x, P[2 * it + 1].y};
    }
  };
  // scatter in two halves: the finished tile is read from LDS into
  // registers (readq), and its stores are issued once the next tile is in
  // LDS (storeq)
  lc Q[2 * CH_NIT];
  auto readq = [&](const lc* buf) {
#pragma unroll
    for (int it = 0; it < CH_NIT; ++it) {
      const int q = sp0 ^ hid[2 * CH_NIT + it];
      Q[2 * it] = at(buf, q);
      Q[2 * it + 1] = at(buf, q ^ spa);
    }
  };
  auto storeq = [&](int32_t oy) {
    char* dst = reinterpret_cast<char*>(a.Y + oy);
    uint32_t o[CH_NIT];
#pragma unroll
    for (int it = 0; it < CH_NIT; ++it) o[it] = (uint32_t)(sg0 + hid[3 * CH_NIT + it]) * 8u;
    if constexpr (SV) {
#pragma unroll
      for (int it = 0; it < CH_NIT; ++it) {
        float4 w;
        w.x = Q[2 * it].x;
        w.y = Q[2 * it].y;
        w.z = Q[2 * it + 1].x;
        w.w = Q[2 * it + 1].y;
        *reinterpret_cast<float4*>(dst + o[it]) = w;
      }
    } else {
      const uint32_t ps = (uint32_t)a.sps * 8u;
#pragma unroll
      for (int it = 0; it < CH_NIT; ++it) {
        c64 w0, w1;
        w0.x = Q[2 * it].x;
        w0.y = Q[2 * it].y;
        w1.x = Q[2 * it + 1].x;
        w1.y = Q[2 * it + 1].y;
        *reinterpret_cast<c64*>(dst + o[it]) = w0;
        *reinterpret_cast<c64*>(dst + o[it] + ps) = w1;
      }
    }
  };

  uint32_t t = blockIdx.x;
  // The two workgroups on a CU do identical work and would stay in lockstep,
  // both in their memory phase at once; the first half of the grid (one
  // workgroup of each pair when the dispatcher fills every CU once, then
  // again) issues at a higher wave priority, so its stages run ahead and
  // its memory phase overlaps the other's stages (2 % per chain when the
  // loads waited for the previous tile's stores; neutral since).
  if (blockIdx.x < G / 2) __builtin_amdgcn_s_setprio(3);
  int32_t cx = 0, cy = 0, nx = 0, ny = 0;   // offsets of the current and the next tile
  if (t < total) {
    offsets(t, cx, cy);
    load(cx);
    put(tile);
    if (t + G < total) offsets(t + G, nx, ny);
  }
  __syncthreads();

  // every wave of a workgroup runs the same iterations, and the loop ends
  // when the workgroup's tiles run out.  Per tile: the next tile's loads are
  // issued (into registers; the memory pipeline is otherwise idle during the
  // stages), the stages run, the finished tile is read into registers, the
  // prefetched one is put into LDS, and the finished tile's stores are
  // issued; they drain during the next tile's stages.  Stores follow the
  // loads they could delay: loads and stores share the vector memory
  // counter, which retires in order, so waiting for a load never waits for a
  // store issued after it.
  for (; t < total; t += G) {
    const bool more = t + G < total;
    int32_t mx = 0, my = 0;
    if (more) {
      load(nx);
      if (t + 2 * G < total) offsets(t + 2 * G, mx, my);
    }
    for (int s = 0; s < a.nst; ++s) {
      const ChainStage& g = a.st[s];
      const int nkc = ((1 << g.nk) + 3) / 4, nmb = ((1 << g.nm) + 15) / 16;
      // FULL: 16 low R values (rlo = 4) and every M row of the blocks used
      const bool full = g.rlo == 4 && g.nm >= 4 + (nmb == 2);
      if constexpr (MODE == 2) {
        switch ((nkc * 4 + nmb) * 2 + (full ? 1 : 0)) {
          case (4 * 4 + 1) * 2 + 1: stage<4, 1, true>(tile, sd, sidx, g, wave, li, lg); break;
          case (1 * 4 + 1) * 2:
          case (1 * 4 + 1) * 2 + 1: stage<1, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (2 * 4 + 1) * 2:
          case (2 * 4 + 1) * 2 + 1: stage<2, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (4 * 4 + 1) * 2: stage<4, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          default: stage<8, 1, false, true>(tile, sd, sidx, g, wave, li, lg); break;
        }
      } else if constexpr (MODE == 1) {
        if (g.rlo == 6) {
          stage_rows(tile, sd, sidx, g, wave, lane);
          __syncthreads();
          continue;
        }
        switch ((nkc * 4 + nmb) * 2 + (full ? 1 : 0)) {
          case (4 * 4 + 1) * 2 + 1: stage<4, 1, true>(tile, sd, sidx, g, wave, li, lg); break;
          case (2 * 4 + 2) * 2 + 1: stage<2, 2, true>(tile, sd, sidx, g, wave, li, lg); break;
          case (1 * 4 + 1) * 2:
          case (1 * 4 + 1) * 2 + 1: stage<1, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (2 * 4 + 1) * 2:
          case (2 * 4 + 1) * 2 + 1: stage<2, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (4 * 4 + 1) * 2: stage<4, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (1 * 4 + 2) * 2:
          case (1 * 4 + 2) * 2 + 1: stage<1, 2, false>(tile, sd, sidx, g, wave, li, lg); break;
          default: stage<2, 2, false>(tile, sd, sidx, g, wave, li, lg); break;
        }
      } else {
        switch ((nkc * 4 + nmb) * 2 + (full ? 1 : 0)) {
          case (4 * 4 + 1) * 2 + 1: stage<4, 1, true>(tile, sd, sidx, g, wave, li, lg); break;
          case (2 * 4 + 2) * 2 + 1: stage<2, 2, true>(tile, sd, sidx, g, wave, li, lg); break;
          case (1 * 4 + 1) * 2:
          case (1 * 4 + 1) * 2 + 1: stage<1, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (2 * 4 + 1) * 2:
          case (2 * 4 + 1) * 2 + 1: stage<2, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (4 * 4 + 1) * 2: stage<4, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (8 * 4 + 1) * 2:
          case (8 * 4 + 1) * 2 + 1: stage<8, 1, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (1 * 4 + 2) * 2:
          case (1 * 4 + 2) * 2 + 1: stage<1, 2, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (2 * 4 + 2) * 2: stage<2, 2, false>(tile, sd, sidx, g, wave, li, lg); break;
          case (4 * 4 + 2) * 2:
          case (4 * 4 + 2) * 2 + 1: stage<4, 2, false>(tile, sd, sidx, g, wave, li, lg); break;
          default: stage<8, 2, false>(tile, sd, sidx, g, wave, li, lg); break;
        }
      }
      __syncthreads();
    }
    readq(tile);
    __syncthreads();
    if (more) put(tile);
    storeq(cy);
    cy = ny;
    nx = mx;
    ny = my;
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_chain(const ChainArgs& a, hipStream_t s) {
  if (a.tb < 0 || a.tb > CH_TB || a.nst < 1 || a.nst > CH_ST || a.spec.total < 1) return hipErrorInvalidValue;
  if (a.ilo > CH_LO_BITS || a.olo > CH_LO_BITS || a.ihi > CH_TB - 1 - CH_LO_BITS || a.ohi > CH_TB - 1 - CH_LO_BITS ||
      a.ihi < 0 || a.ohi < 0)
    return hipErrorInvalidValue;
  if ((a.gvec && ((uintptr_t)a.X % 16 != 0 || !a.gvec_ok)) || (a.svec && ((uintptr_t)a.Y % 16 != 0 || !a.svec_ok)))
    return hipErrorInvalidValue;
  if ((int64_t)a.sdn * 8 + (int64_t)a.idn * 4 > CH_STAGE_BYTES) return hipErrorInvalidValue;
  const size_t lds = (sizeof(lc) << a.tb) + sizeof(lc) * a.sdn + sizeof(int) * (a.idn + 4 * CH_NIT);
  // persistent workgroups: as many as fit on the device at once with this
  // LDS size (two per CU for a 2^13-element tile)
  static int ncu[64] = {0};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  if (ncu[dev] == 0) {
    e = hipDeviceGetAttribute(&ncu[dev], hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (ncu[dev] < 1) ncu[dev] = 1;
  }
  // the small-stage instantiation for chains whose every operator has at
  // most 4 chunk-blocks (lane-row stages, 6 R bits from the lane, exist
  // only in it), the one-M-block instantiation for chains whose every
  // operator has one M block, else the full one
  bool small = true, rows = false, col = true;
  for (int t = 0; t < a.nst; ++t) {
    const ChainStage& g = a.st[t];
    const int nkc = ((1 << g.nk) + 3) / 4, nmb = ((1 << g.nm) + 15) / 16;
    small = small && nkc * nmb <= 4;
    col = col && nmb == 1;
    if (g.rlo == 6) rows = true;
    if (g.rlo < 0 || g.rlo > 6 || (g.rlo == 6 && (g.nk > 2 || g.nm > 2)) || (g.rlo == 5)) return hipErrorInvalidValue;
  }
  if (rows && !small) return hipErrorInvalidValue;
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chain<true, true, 0>, CH_THREADS, lds);
  if (e != hipSuccess) return e;
  per_cu = std::max(per_cu, 1);
  const uint32_t grid = std::min<uint32_t>(a.spec.total, (uint32_t)(per_cu * ncu[dev]));
#define PQ_CHAIN_LAUNCH(G, S)                                                         \
  do {                                                                                \
    if (small)                                                                        \
      hipLaunchKernelGGL((k_chain<G, S, 1>), dim3(grid), dim3(CH_THREADS), lds, s, a); \
    else if (col)                                                                     \
      hipLaunchKernelGGL((k_chain<G, S, 2>), dim3(grid), dim3(CH_THREADS), lds, s, a); \
    else                                                                              \
      hipLaunchKernelGGL((k_chain<G, S, 0>), dim3(grid), dim3(CH_THREADS), lds, s, a); \
  } while (0)
  if (a.gvec && a.svec)
    PQ_CHAIN_LAUNCH(true, true);
  else if (a.gvec)
    PQ_CHAIN_LAUNCH(true, false);
  else if (a.svec)
    PQ_CHAIN_LAUNCH(false, true);
  else
    PQ_CHAIN_LAUNCH(false, false);
#undef PQ_CHAIN_LAUNCH
  return hipGetLastError();
}
