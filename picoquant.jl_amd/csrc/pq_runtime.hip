// libpq_hip runtime: context (one device, one stream), a stream-ordered
// caching allocator, the tensor table (refcounted buffers + strided views),
// the lowering of an ncon-labelled pairwise contraction to index groups, and
// the kernel dispatch.  The C ABI is declared in include/pq_hip.h.
#include "pq_rt.h"

#include <algorithm>
#include <cstring>

namespace pqrt {

int64_t numel(const Tensor& t) {
  int64_t n = 1;
  for (int i = 0; i < t.rank; ++i) n *= t.dims[i];
  return n;
}

bool is_contiguous(const Tensor& t) {
  int64_t s = 1;
  for (int i = 0; i < t.rank; ++i) {
    if (t.dims[i] != 1 && t.strides[i] != s) return false;
    s *= t.dims[i];
  }
  return true;
}

void set_contiguous(Tensor& t) {
  int64_t s = 1;
  for (int i = 0; i < t.rank; ++i) {
    t.strides[i] = s;
    s *= t.dims[i];
  }
}

// Drop unit extents and merge neighbours that are contiguous in both tensors.
std::vector<Mode> merge_modes(const std::vector<Mode>& in) {
  std::vector<Mode> out;
  for (const Mode& m : in) {
    if (m.ext == 1) continue;
    if (!out.empty()) {
      Mode& p = out.back();
      if (m.s0 == p.s0 * p.ext && m.s1 == p.s1 * p.ext) {
        p.ext *= m.ext;
        continue;
      }
    }
    out.push_back(m);
  }
  return out;
}

bool fill_group(Group& g, const std::vector<Mode>& modes) {
  if ((int)modes.size() > PQ_MAX_MODES) return false;
  memset(&g, 0, sizeof(g));
  g.nmodes = (int)modes.size();
  uint64_t total = 1;
  for (size_t j = 0; j < modes.size(); ++j) {
    if (modes[j].s0 >= (1ll << 31) || modes[j].s1 >= (1ll << 31)) return false;
    g.ext[j] = make_fdiv((uint32_t)modes[j].ext);
    g.s0[j] = (int32_t)modes[j].s0;
    g.s1[j] = (int32_t)modes[j].s1;
    total *= (uint64_t)modes[j].ext;
  }
  if (total >= (1ull << 31)) return false;
  g.total = (uint32_t)total;
  return true;
}

Batch single_batch() {
  Batch L;
  memset(&L, 0, sizeof(L));
  L.total = 1;
  return L;
}

int fail(pq_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hipfail(pq_ctx* c, hipError_t e, const char* where) {
  return fail(c, PQ_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

static size_t round_bytes(size_t b) {
  if (b < 256) return 256;
  if (b < (2u << 20)) {
    size_t p = 256;
    while (p < b) p <<= 1;
    return p;
  }
  const size_t g = 2u << 20;
  return (b + g - 1) / g * g;
}

void* dev_alloc(pq_ctx* c, size_t bytes, size_t* got) {
  size_t want = round_bytes(bytes);
  auto it = c->free_blocks.lower_bound(want);
  if (it != c->free_blocks.end() && it->first <= want + want / 4 + (2u << 20)) {
    void* p = it->second;
    *got = it->first;
    c->cached -= it->first;
    c->free_blocks.erase(it);
    return p;
  }
  void* p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    hipStreamSynchronize(c->stream);
    for (auto& kv : c->free_blocks) hipFree(kv.second);
    c->free_blocks.clear();
    c->cached = 0;
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  *got = want;
  return p;
}

void dev_release(pq_ctx* c, void* p, size_t bytes) {
  // Stream-ordered reuse: later work on the same stream runs after the
  // kernels that still read this block.
  c->free_blocks.emplace(bytes, p);
  c->cached += bytes;
}

int ensure_aux_streams(pq_ctx* c) {
  for (hipStream_t& a : c->aux)
    if (!a && hipStreamCreateWithFlags(&a, hipStreamNonBlocking) != hipSuccess) {
      a = nullptr;
      return fail(c, PQ_ERR_HIP, "stream creation failed");
    }
  return PQ_OK;
}

Buf* new_buf(pq_ctx* c, size_t bytes) {
  size_t got = 0;
  void* p = dev_alloc(c, bytes ? bytes : 1, &got);
  if (!p) return nullptr;
  Buf* b = new Buf;
  b->ptr = p;
  b->bytes = got;
  b->refs = 1;
  return b;
}

void unref(pq_ctx* c, Buf* b) {
  if (--b->refs == 0) {
    dev_release(c, b->ptr, b->bytes);
    delete b;
  }
}

Tensor* get(pq_ctx* c, pq_tensor h) {
  auto it = c->tensors.find(h);
  return it == c->tensors.end() ? nullptr : &it->second;
}

pq_tensor add(pq_ctx* c, const Tensor& t) {
  pq_tensor id = c->next_id++;
  c->tensors[id] = t;
  return id;
}

ProfScope::ProfScope(pq_ctx* c_, int kc_, double bytes, double flops) : c(c_), kc(kc_) {
  on = c->prof && !c->capturing;
  if (!on) return;
  hipEvent_t e0;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, c->stream);
  c->pe[kc].events.push_back({e0, e1});
  c->pe[kc].launches++;
  c->pe[kc].bytes += bytes;
  c->pe[kc].flops += flops;
}

ProfScope::~ProfScope() {
  if (on) hipEventRecord(e1, c->stream);
}

template <typename C>
static C* dptr(const Tensor& t) {
  return reinterpret_cast<C*>(t.buf->ptr) + t.off;
}

int materialize_into(pq_ctx* c, const Tensor& src, Tensor& dst) {
  dst = Tensor();
  dst.rank = src.rank;
  for (int i = 0; i < src.rank; ++i) dst.dims[i] = src.dims[i];
  set_contiguous(dst);
  int64_t n = numel(src);
  dst.buf = new_buf(c, (size_t)n * c->esize);
  if (!dst.buf) return fail(c, PQ_ERR_OOM, "device allocation failed");
  if (n == 0) return PQ_OK;
  std::vector<Mode> modes;
  for (int i = 0; i < src.rank; ++i) modes.push_back({src.dims[i], src.strides[i], dst.strides[i]});
  modes = merge_modes(modes);
  CopyArgs a;
  if (!fill_group(a.G, modes)) {
    unref(c, dst.buf);
    return fail(c, PQ_ERR_UNSUPPORTED, "tensor too large or too fragmented");
  }
  ProfScope ps(c, PQ_KC_COPY, 2.0 * n * c->esize, 0);
  hipError_t e = (c->dtype == PQ_C128)
                     ? launch_copy<c128>(dptr<c128>(src), dptr<c128>(dst), a, c->stream)
                     : launch_copy<c64>(dptr<c64>(src), dptr<c64>(dst), a, c->stream);
  if (e != hipSuccess) return hipfail(c, e, "copy kernel");
  return PQ_OK;
}

// Default workgroup tile for the MFMA GEMM: big tiles when both free extents
// fill them, narrow ones otherwise (complex128 is limited to 64x64).
static int pick_tile(uint64_t Mt, uint64_t Nt, size_t esize) {
  if (Nt <= 16) return Mt >= 4096 && esize == 8 ? 8 : (Mt >= 128 ? 7 : 6);
  if (Mt <= 16) return Nt >= 4096 && esize == 8 ? 13 : (Nt >= 128 ? 12 : 9);
  if (Mt <= 32 && Nt <= 32) return 0;
  if (Nt <= 32) return Mt >= 128 ? 10 : 1;
  if (Mt <= 32) return 2;
  if (esize == 8 && Mt >= 4096 && Nt >= 128) return 5;
  if (esize == 8 && Mt >= 4096 && Nt >= 64) return 4;
  return 3;
}

// Workgroup tile (rows, columns) of GEMM tile index t, as launch_gemm runs it.
static void gemm_tile_dims(int t, uint64_t& bm, uint64_t& bn) {
  static const int d[PQ_GEMM_TILES][2] = {{32, 32},  {64, 32},  {32, 64},  {64, 64},   {128, 64},
                                          {64, 128}, {64, 16},  {128, 16}, {256, 16},  {16, 64},
                                          {128, 32}, {128, 128}, {16, 128}, {16, 256}, {32, 128}};
  bm = (uint64_t)d[t][0];
  bn = (uint64_t)d[t][1];
}

// Split-K: a GEMM whose output tiles cannot fill the device (a long K with
// few rows and columns) runs as splitk partial products over consecutive
// parts of its slowest K mode, written to scratch and summed in FP64 by
// k_sum_partials.  Parts are halved until about GEMM_SPLIT_WGS workgroups
// run, each part keeps at least GEMM_SPLIT_MIN_K values of K, and at most
// GEMM_SPLIT_MAX parts are made.
#define GEMM_SPLIT_WGS 1024
#define GEMM_SPLIT_MIN_K 512
#define GEMM_SPLIT_MAX 16
static int gemm_split(const StepSpec& s, uint64_t m, uint64_t n) {
  if (s.L.total != 1 || s.K.empty() || s.L.nmodes >= PQ_MAX_BMODES) return 1;
  uint64_t bm, bn;
  gemm_tile_dims(s.tile, bm, bn);
  const uint64_t tiles = ((m + bm - 1) / bm) * ((n + bn - 1) / bn);
  const uint64_t last = (uint64_t)s.K.back().ext;
  int parts = 1;
  while (tiles * parts < GEMM_SPLIT_WGS && parts < GEMM_SPLIT_MAX && last % (2 * parts) == 0 &&
         s.Kt / (2 * parts) >= GEMM_SPLIT_MIN_K)
    parts *= 2;
  return parts;
}

static bool stream_fits(uint64_t Kt, uint64_t Ny, size_t esize) {
  uint64_t KP = 1;
  while (KP < Kt) KP <<= 1;
  size_t smem = KP * Ny * esize + KP * 4 + Ny * 4;
  return Kt <= PQ_STREAM_KMAX && Kt * Ny <= PQ_STREAM_YMAX && smem <= 48 * 1024;
}

int resolve_step(pq_ctx* c, StepSpec& s, size_t esize) {
  s.Mt = s.Nt = s.Kt = 1;
  for (auto& m : s.M) s.Mt *= m.ext;
  for (auto& m : s.N) s.Nt *= m.ext;
  for (auto& m : s.K) s.Kt *= m.ext;
  const uint64_t Mt = s.Mt, Nt = s.Nt, Kt = s.Kt;
  const bool ab = stream_fits(Kt, Nt, esize), ba = stream_fits(Kt, Mt, esize);
  // MFMA tiles zero-pad short extents: worth it once one free side fills a
  // 16-wide tile and there is enough work per output.
  const bool gemm_ok = Kt <= (uint64_t)gemm_k_limit() && Kt >= 4 && std::max(Mt, Nt) >= 16 &&
                       Mt * Nt >= 128;
  const bool step_kc = s.kc >= 0;   // a step's own class is taken as given
  int kc = step_kc ? s.kc : c->override_kc;
  if (kc < 0) {
    if ((ab && Kt * Nt <= 64) || (ba && Kt * Mt <= 64)) kc = PQ_KC_STREAM;
    else if (Mt * Nt <= 16 && Kt > 64) kc = PQ_KC_REDUCE;
    else if (gemm_ok) kc = PQ_KC_GEMM;
    else if (ab || ba) kc = PQ_KC_STREAM;
    else kc = PQ_KC_GENERIC;
  }
  if (kc == PQ_KC_STREAM) {
    if (!ab && !ba) return fail(c, PQ_ERR_UNSUPPORTED, "stream kernel not applicable");
    s.stream_ab = ab && (Mt >= Nt || !ba);
  } else if (kc == PQ_KC_GEMM || kc == PQ_KC_GEMM16) {
    if (!gemm_ok) return fail(c, PQ_ERR_UNSUPPORTED, "gemm kernel not applicable");
    // FP16 hi/lo arithmetic: complex64 steps whose first K mode runs
    // contiguously in both operands with an extent that is a multiple of 32.
    const int arith = s.arith >= 0 ? s.arith : c->arith;
    const bool k32 = !s.K.empty() && s.K[0].s0 == 1 && s.K[0].s1 == 1 && s.K[0].ext % 32 == 0;
    const bool c64 = (s.dtype >= 0 ? s.dtype : c->dtype) == PQ_C64;
    if (kc == PQ_KC_GEMM16 && !(c64 && k32))
      return fail(c, PQ_ERR_UNSUPPORTED, "FP16 split GEMM needs complex64 and contiguous K of 32s");
    if (kc == PQ_KC_GEMM && !step_kc && arith == 1 && c64 && k32) kc = PQ_KC_GEMM16;
    // The MFMA result layout puts 16 consecutive lanes on 16 consecutive
    // N-side outputs; when C's unit stride lies on the M side the product
    // is computed transposed so that those lanes write adjacent addresses.
    bool c_unit_m = false, c_unit_n = false;
    for (auto& m : s.M) c_unit_m = c_unit_m || m.s1 == 1;
    for (auto& m : s.N) c_unit_n = c_unit_n || m.s1 == 1;
    s.gemm_swap = c->gemm_orient >= 0 ? c->gemm_orient == 1 : (c_unit_m && !c_unit_n);
    int t = s.tile >= 0 ? s.tile : c->gemm_tile;
    if (t >= PQ_GEMM_TILES) return fail(c, PQ_ERR_ARG, "GEMM tile out of range");
    const uint64_t m = s.gemm_swap ? Nt : Mt, n = s.gemm_swap ? Mt : Nt;
    if (t >= 0)
      s.tile = t;
    else if (kc == PQ_KC_GEMM16)   // the split's conversions favour square 64x64 tiles
      s.tile = n <= 16 ? 6 : (m <= 16 ? 9 : 3);
    else
      s.tile = pick_tile(m, n, esize);
    s.splitk = kc == PQ_KC_GEMM ? gemm_split(s, m, n) : 1;
  } else if (kc == PQ_KC_REDUCE) {
    if (Mt * Nt > 16) return fail(c, PQ_ERR_UNSUPPORTED, "reduce kernel needs M*N <= 16");
  } else if (kc != PQ_KC_GENERIC) {
    return fail(c, PQ_ERR_UNSUPPORTED, "kernel class not available for this shape");
  }
  if (s.L.total < 1 || s.L.total > 65535) return fail(c, PQ_ERR_UNSUPPORTED, "batch too large");
  s.kc = kc;
  return PQ_OK;
}

int prepare_contract(pq_ctx* c, const Tensor& A, int ra, const int32_t* ia, const Tensor& B, int rb,
                     const int32_t* ib, Tensor& Ct, StepSpec& s, int64_t& kt) {
  if (ra != A.rank || rb != B.rank) return fail(c, PQ_ERR_ARG, "label count does not match tensor rank");
  if ((ra && !ia) || (rb && !ib)) return fail(c, PQ_ERR_ARG, "null label array");
  for (int i = 0; i < ra; ++i) {
    if (ia[i] == 0) return fail(c, PQ_ERR_ARG, "label 0 is not allowed");
    for (int j = i + 1; j < ra; ++j)
      if (ia[i] == ia[j]) return fail(c, PQ_ERR_ARG, "repeated label in A");
  }
  for (int i = 0; i < rb; ++i) {
    if (ib[i] == 0) return fail(c, PQ_ERR_ARG, "label 0 is not allowed");
    for (int j = i + 1; j < rb; ++j)
      if (ib[i] == ib[j]) return fail(c, PQ_ERR_ARG, "repeated label in B");
  }
  std::vector<int> a_open, b_open;
  std::vector<std::pair<int, int>> common;
  for (int i = 0; i < ra; ++i) {
    if (ia[i] < 0) {
      for (int j = 0; j < rb; ++j)
        if (ib[j] == ia[i]) return fail(c, PQ_ERR_ARG, "open label shared by A and B");
      a_open.push_back(i);
    } else {
      int jj = -1;
      for (int j = 0; j < rb; ++j)
        if (ib[j] == ia[i]) jj = j;
      if (jj < 0) return fail(c, PQ_ERR_ARG, "summed label missing from B");
      if (A.dims[i] != B.dims[jj]) return fail(c, PQ_ERR_ARG, "summed dimensions differ");
      common.push_back({i, jj});
    }
  }
  for (int j = 0; j < rb; ++j) {
    if (ib[j] < 0) {
      b_open.push_back(j);
    } else {
      bool found = false;
      for (int i = 0; i < ra; ++i)
        if (ia[i] == ib[j]) found = true;
      if (!found) return fail(c, PQ_ERR_ARG, "summed label missing from A");
    }
  }
  int rc = (int)(a_open.size() + b_open.size());
  if (rc > PQ_MAX_RANK) return fail(c, PQ_ERR_UNSUPPORTED, "result rank too large");
  Ct = Tensor();
  Ct.rank = rc;
  int p = 0;
  for (int i : a_open) Ct.dims[p++] = A.dims[i];
  for (int j : b_open) Ct.dims[p++] = B.dims[j];
  set_contiguous(Ct);
  const int64_t nc = numel(Ct);
  if (nc >= (1ll << 31)) return fail(c, PQ_ERR_UNSUPPORTED, "result has 2^31 or more elements");
  kt = 1;
  for (auto& pr : common) kt *= A.dims[pr.first];
  if (nc == 0 || kt == 0) return PQ_OK;
  s = StepSpec();
  p = 0;
  for (int i : a_open) s.M.push_back({A.dims[i], A.strides[i], Ct.strides[p++]});
  for (int j : b_open) s.N.push_back({B.dims[j], B.strides[j], Ct.strides[p++]});
  for (auto& pr : common) s.K.push_back({A.dims[pr.first], A.strides[pr.first], B.strides[pr.second]});
  s.M = merge_modes(s.M);
  s.N = merge_modes(s.N);
  s.K = merge_modes(s.K);
  s.L = single_batch();
  return resolve_step(c, s, c->esize);
}

size_t step_scratch_bytes(const StepSpec& s, size_t esize) {
  if (s.kc == PQ_KC_GEMM && s.splitk > 1) return (size_t)s.splitk * s.Mt * s.Nt * esize;
  if (s.kc != PQ_KC_REDUCE) return 0;
  return (size_t)reduce_blocks_for((uint32_t)s.Kt) * s.Mt * s.Nt * s.L.total * esize;
}

template <typename C>
static int launch_typed(pq_ctx* c, const C* pa, const C* pb, C* pc, const StepSpec& s, void* scratch) {
  const double es = (double)sizeof(C);
  const double nb = (double)s.L.total;
  const double alg_bytes = nb * es * (double)(s.Mt * s.Kt + s.Kt * s.Nt + s.Mt * s.Nt);
  const double alg_flops = nb * 8.0 * (double)s.Mt * (double)s.Nt * (double)s.Kt;
  hipError_t e = hipSuccess;
  if (s.kc == PQ_KC_STREAM) {
    StreamArgs sa;
    memset(&sa, 0, sizeof(sa));
    sa.L = s.L;
    if (s.stream_ab) {
      if (!fill_group(sa.M, s.M) || !fill_group(sa.K, s.K) || !fill_group(sa.N, s.N))
        return fail(c, PQ_ERR_UNSUPPORTED, "too many modes");
      ProfScope ps(c, PQ_KC_STREAM, alg_bytes, alg_flops);
      e = launch_stream<C>(pa, pb, pc, sa, c->stream);
    } else {
      // X = B (big), Y = A (small): K and batch strides swap roles.
      std::vector<Mode> K2;
      for (auto& m : s.K) K2.push_back({m.ext, m.s1, m.s0});
      for (int j = 0; j < sa.L.nmodes; ++j) std::swap(sa.L.sA[j], sa.L.sB[j]);
      if (!fill_group(sa.M, s.N) || !fill_group(sa.K, K2) || !fill_group(sa.N, s.M))
        return fail(c, PQ_ERR_UNSUPPORTED, "too many modes");
      ProfScope ps(c, PQ_KC_STREAM, alg_bytes, alg_flops);
      e = launch_stream<C>(pb, pa, pc, sa, c->stream);
    }
  } else if (s.kc == PQ_KC_GEMM || s.kc == PQ_KC_GEMM16) {
    // Operand roles as launched: transposed product swaps A/B, M/N, the K
    // strides and the batch strides.
    const std::vector<Mode>& gM = s.gemm_swap ? s.N : s.M;
    const std::vector<Mode>& gN = s.gemm_swap ? s.M : s.N;
    std::vector<Mode> gK = s.K;
    Batch L = s.L;
    const C* xa = pa;
    const C* xb = pb;
    if (s.gemm_swap) {
      for (auto& m : gK) std::swap(m.s0, m.s1);
      for (int j = 0; j < L.nmodes; ++j) std::swap(L.sA[j], L.sB[j]);
      std::swap(xa, xb);
    }
    GemmArgs ga;
    memset(&ga, 0, sizeof(ga));
    if (!fill_group(ga.M, gM) || !fill_group(ga.N, gN) || !fill_group(ga.K, gK))
      return fail(c, PQ_ERR_UNSUPPORTED, "too many modes");
    ga.L = L;
    for (auto& m : gK) {
      if (m.s0 == 1) ga.a_kfast = 1;
      if (m.s1 == 1) ga.b_kfast = 1;
    }
    ga.tile = s.tile;
    // Vector-load fast path: K runs contiguously (unit stride) in both
    // operands and every element offset keeps 16-byte alignment.
    {
      const int64_t vec = 16 / (int64_t)sizeof(C);
      bool ok = !gK.empty() && gK[0].s0 == 1 && gK[0].s1 == 1 && gK[0].ext % 8 == 0 &&
                ((uintptr_t)xa % 16) == 0 && ((uintptr_t)xb % 16) == 0;
      for (size_t q = 1; q < gK.size() && ok; ++q) ok = gK[q].s0 % vec == 0 && gK[q].s1 % vec == 0;
      for (auto& m : gM) ok = ok && m.s0 % vec == 0;
      for (auto& m : gN) ok = ok && m.s0 % vec == 0;
      for (int j = 0; j < L.nmodes && ok; ++j) ok = L.sA[j] % vec == 0 && L.sB[j] % vec == 0;
      ga.contig = ok ? 1 : 0;
    }
    bool split = false;
    if constexpr (sizeof(C) == 8) split = s.kc == PQ_KC_GEMM16 && gemm16_applicable(ga);
    if (s.kc == PQ_KC_GEMM && s.splitk > 1) {
      // partial products: the slowest K mode is cut into splitk parts (a
      // batch mode); part z writes a compact [m][n] block of scratch
      if (!scratch || L.total != 1 || gK.empty() || gK.back().ext % s.splitk != 0)
        return fail(c, PQ_ERR_UNSUPPORTED, "split-K needs scratch and an unbatched step");
      const int parts = s.splitk;
      std::vector<Mode> pK = gK;
      Mode& km = pK.back();
      km.ext /= parts;
      const int64_t mt = (int64_t)(s.gemm_swap ? s.Nt : s.Mt), mn = (int64_t)(s.Mt * s.Nt);
      std::vector<Mode> pM = gM, pN = gN, rM, rN;
      int64_t run = 1;
      for (auto& m : pM) {
        rM.push_back({m.ext, run, m.s1});   // partial offset, C offset
        m.s1 = run;
        run *= m.ext;
      }
      run = mt;
      for (auto& m : pN) {
        rN.push_back({m.ext, run, m.s1});
        m.s1 = run;
        run *= m.ext;
      }
      GemmArgs gp = ga;
      if (!fill_group(gp.M, pM) || !fill_group(gp.N, pN) || !fill_group(gp.K, pK))
        return fail(c, PQ_ERR_UNSUPPORTED, "too many modes");
      gp.L = single_batch();
      gp.L.nmodes = 1;
      gp.L.ext[0] = make_fdiv((uint32_t)parts);
      gp.L.sA[0] = (int64_t)km.ext * km.s0;
      gp.L.sB[0] = (int64_t)km.ext * km.s1;
      gp.L.sC[0] = mn;
      gp.L.total = (uint32_t)parts;
      {
        const int64_t vec = 16 / (int64_t)sizeof(C);
        if (gp.contig && (gp.L.sA[0] % vec || gp.L.sB[0] % vec)) gp.contig = 0;
      }
      Group sm, sn;
      if (!fill_group(sm, rM) || !fill_group(sn, rN)) return fail(c, PQ_ERR_UNSUPPORTED, "too many modes");
      ProfScope ps(c, PQ_KC_GEMM, alg_bytes, alg_flops);
      C* part = reinterpret_cast<C*>(scratch);
      e = launch_gemm<C>(xa, xb, part, gp, c->stream);
      if (e == hipSuccess) e = launch_sum_partials<C>(part, pc, sm, sn, (uint32_t)mt, (uint64_t)mn, parts, c->stream);
    } else if (split) {
      ProfScope ps(c, PQ_KC_GEMM16, alg_bytes, alg_flops);
      if constexpr (sizeof(C) == 8) e = launch_gemm16(xa, xb, pc, ga, c->stream);
    } else {
      ProfScope ps(c, PQ_KC_GEMM, alg_bytes, alg_flops);
      e = launch_gemm<C>(xa, xb, pc, ga, c->stream);
    }
  } else {
    GenArgs ga;
    memset(&ga, 0, sizeof(ga));
    if (!fill_group(ga.M, s.M) || !fill_group(ga.N, s.N) || !fill_group(ga.K, s.K))
      return fail(c, PQ_ERR_UNSUPPORTED, "too many modes");
    ga.L = s.L;
    if (s.kc == PQ_KC_REDUCE) {
      ProfScope ps(c, PQ_KC_REDUCE, alg_bytes, alg_flops);
      e = launch_reduce<C>(pa, pb, pc, ga, reinterpret_cast<C*>(scratch), reduce_blocks_for((uint32_t)s.Kt),
                           c->stream);
    } else {
      ProfScope ps(c, PQ_KC_GENERIC, alg_bytes, alg_flops);
      e = launch_generic<C>(pa, pb, pc, ga, c->stream);
    }
  }
  if (e != hipSuccess) return hipfail(c, e, "contraction kernel");
  return PQ_OK;
}

int launch_step(pq_ctx* c, const void* A, const void* B, void* Out, const StepSpec& s, void* scratch) {
  const int dt = s.dtype >= 0 ? s.dtype : c->dtype;
  if (dt == PQ_C128)
    return launch_typed<c128>(c, (const c128*)A, (const c128*)B, (c128*)Out, s, scratch);
  return launch_typed<c64>(c, (const c64*)A, (const c64*)B, (c64*)Out, s, scratch);
}

}  // namespace pqrt

using namespace pqrt;

extern "C" {

int pq_ctx_create(int device, int dtype, pq_ctx** out) {
  if (!out) return PQ_ERR_ARG;
  if (dtype != PQ_C128 && dtype != PQ_C64) return PQ_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return PQ_ERR_HIP;
  }
  if (hipSetDevice(device) != hipSuccess) return PQ_ERR_HIP;
  pq_ctx* c = new pq_ctx;
  c->device = device;
  c->dtype = dtype;
  c->esize = dtype == PQ_C128 ? 16 : 8;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PQ_ERR_HIP;
  }
  c->noalloc_scratch = noalloc_scratch(device);
  *out = c;
  return PQ_OK;
}

int pq_ctx_destroy(pq_ctx* c) {
  if (!c) return PQ_ERR_ARG;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  for (auto& kv : c->tensors) unref(c, kv.second.buf);
  c->tensors.clear();
  for (auto& kv : c->free_blocks) hipFree(kv.second);
  c->free_blocks.clear();
  if (c->scratch) hipFree(c->scratch);
  for (hipStream_t& a : c->aux)
    if (a) {
      hipStreamSynchronize(a);
      hipStreamDestroy(a);
      a = nullptr;
    }
  release_blas(c);
  for (auto& p : c->pe)
    for (auto& ev : p.events) {
      hipEventDestroy(ev.first);
      hipEventDestroy(ev.second);
    }
  hipStreamDestroy(c->stream);
  delete c;
  return PQ_OK;
}

const char* pq_last_error(const pq_ctx* c) { return c ? c->err.c_str() : "null context"; }
int pq_ctx_dtype(const pq_ctx* c) { return c ? c->dtype : -1; }

int pq_synchronize(pq_ctx* c) {
  if (!c) return PQ_ERR_ARG;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hipfail(c, e, "hipStreamSynchronize");
  return PQ_OK;
}

namespace {

// Stream-ordered copy of n host elements (host_dtype) into device memory of
// the context's element type, converting on the host when the types differ;
// synchronises, so the host buffer may be reused on return.
hipError_t host_to_device(pq_ctx* c, void* dst, const void* host, int host_dtype, int64_t n) {
  if (n <= 0) return hipSuccess;
  std::vector<unsigned char> tmp;
  const void* src = host;
  if (host_dtype != c->dtype) {
    tmp.resize((size_t)n * c->esize);
    if (c->dtype == PQ_C64) {
      const double* s = (const double*)host;
      float* d = (float*)tmp.data();
      for (int64_t i = 0; i < 2 * n; ++i) d[i] = (float)s[i];
    } else {
      const float* s = (const float*)host;
      double* d = (double*)tmp.data();
      for (int64_t i = 0; i < 2 * n; ++i) d[i] = (double)s[i];
    }
    src = tmp.data();
  }
  hipError_t e = hipMemcpyAsync(dst, src, (size_t)n * c->esize, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

}  // namespace

int pq_tensor_upload(pq_ctx* c, int rank, const int64_t* dims, const void* host, int host_dtype,
                     pq_tensor* out) {
  if (!c || !out || rank < 0 || rank > PQ_MAX_RANK || (rank > 0 && !dims))
    return fail(c, PQ_ERR_ARG, "bad upload arguments");
  if (host_dtype != PQ_C128 && host_dtype != PQ_C64) return fail(c, PQ_ERR_ARG, "bad host dtype");
  Tensor t;
  t.rank = rank;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] < 0) return fail(c, PQ_ERR_ARG, "negative dimension");
    t.dims[i] = dims[i];
  }
  set_contiguous(t);
  int64_t n = numel(t);
  if (n >= (1ll << 31)) return fail(c, PQ_ERR_UNSUPPORTED, "tensor has 2^31 or more elements");
  if (n > 0 && !host) return fail(c, PQ_ERR_ARG, "null host pointer");
  t.buf = new_buf(c, (size_t)n * c->esize);
  if (!t.buf) return fail(c, PQ_ERR_OOM, "device allocation failed");
  hipError_t e = host_to_device(c, t.buf->ptr, host, host_dtype, n);
  if (e != hipSuccess) {
    unref(c, t.buf);
    return hipfail(c, e, "upload");
  }
  *out = add(c, t);
  return PQ_OK;
}

int pq_tensor_write(pq_ctx* c, pq_tensor h, const void* host, int host_dtype) {
  if (!c) return PQ_ERR_ARG;
  if (host_dtype != PQ_C128 && host_dtype != PQ_C64) return fail(c, PQ_ERR_ARG, "bad host dtype");
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  if (!is_contiguous(*t)) return fail(c, PQ_ERR_ARG, "tensor_write needs a contiguous tensor");
  const int64_t n = numel(*t);
  if (n > 0 && !host) return fail(c, PQ_ERR_ARG, "null host pointer");
  hipError_t e = host_to_device(c, (unsigned char*)t->buf->ptr + t->off * c->esize, host, host_dtype, n);
  if (e != hipSuccess) return hipfail(c, e, "tensor write");
  return PQ_OK;
}

int pq_tensor_shape(pq_ctx* c, pq_tensor h, int* rank, int64_t* dims) {
  if (!c || !rank) return PQ_ERR_ARG;
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  *rank = t->rank;
  if (dims)
    for (int i = 0; i < t->rank; ++i) dims[i] = t->dims[i];
  return PQ_OK;
}

int pq_tensor_download(pq_ctx* c, pq_tensor h, void* host) {
  if (!c) return PQ_ERR_ARG;
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  int64_t n = numel(*t);
  if (n == 0) return PQ_OK;
  if (!host) return fail(c, PQ_ERR_ARG, "null host pointer");
  Tensor src = *t;
  Tensor tmp;
  bool own = false;
  if (!is_contiguous(src)) {
    int r = materialize_into(c, src, tmp);
    if (r) return r;
    src = tmp;
    own = true;
  }
  hipError_t e = hipMemcpyAsync(host, reinterpret_cast<unsigned char*>(src.buf->ptr) + src.off * c->esize,
                                (size_t)n * c->esize, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (own) unref(c, tmp.buf);
  if (e != hipSuccess) return hipfail(c, e, "download");
  return PQ_OK;
}

int pq_tensor_free(pq_ctx* c, pq_tensor h) {
  if (!c) return PQ_ERR_ARG;
  auto it = c->tensors.find(h);
  if (it == c->tensors.end()) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  unref(c, it->second.buf);
  c->tensors.erase(it);
  return PQ_OK;
}

int pq_contract(pq_ctx* c, pq_tensor ha, int ra, const int32_t* ia, pq_tensor hb, int rb,
                const int32_t* ib, pq_tensor* out) {
  if (!c || !out) return PQ_ERR_ARG;
  Tensor* A = get(c, ha);
  Tensor* B = get(c, hb);
  if (!A || !B) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  Tensor Ct;
  StepSpec s;
  int64_t kt = 0;
  int r = prepare_contract(c, *A, ra, ia, *B, rb, ib, Ct, s, kt);
  if (r) return r;
  const int64_t nc = numel(Ct);
  Ct.buf = new_buf(c, (size_t)nc * c->esize);
  if (!Ct.buf) return fail(c, PQ_ERR_OOM, "device allocation failed");
  if (nc > 0 && kt == 0) {
    hipError_t e = hipMemsetAsync(Ct.buf->ptr, 0, (size_t)nc * c->esize, c->stream);
    if (e != hipSuccess) {
      unref(c, Ct.buf);
      return hipfail(c, e, "memset");
    }
  } else if (nc > 0) {
    size_t need = step_scratch_bytes(s, c->esize);
    if (need > c->scratch_bytes) {
      if (c->scratch) {
        hipStreamSynchronize(c->stream);
        hipFree(c->scratch);
        c->scratch = nullptr;
        c->scratch_bytes = 0;
      }
      if (hipMalloc(&c->scratch, need) != hipSuccess) {
        (void)hipGetLastError();
        r = fail(c, PQ_ERR_OOM, "scratch allocation failed");
      } else {
        c->scratch_bytes = need;
      }
    }
    if (r == PQ_OK) {
      const unsigned char* pa = (const unsigned char*)A->buf->ptr + A->off * c->esize;
      const unsigned char* pb = (const unsigned char*)B->buf->ptr + B->off * c->esize;
      unsigned char* pc = (unsigned char*)Ct.buf->ptr;
      r = launch_step(c, pa, pb, pc, s, c->scratch);
    }
    if (r) {
      unref(c, Ct.buf);
      return r;
    }
  }
  *out = add(c, Ct);
  return PQ_OK;
}

int pq_permute(pq_ctx* c, pq_tensor h, const int32_t* perm, pq_tensor* out) {
  if (!c || !out) return PQ_ERR_ARG;
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  if (t->rank && !perm) return fail(c, PQ_ERR_ARG, "null permutation");
  std::vector<int> seen(t->rank, 0);
  Tensor r = *t;
  for (int i = 0; i < t->rank; ++i) {
    int a = perm[i] - 1;
    if (a < 0 || a >= t->rank || seen[a]) return fail(c, PQ_ERR_ARG, "not a permutation");
    seen[a] = 1;
    r.dims[i] = t->dims[a];
    r.strides[i] = t->strides[a];
  }
  r.buf->refs++;
  *out = add(c, r);
  return PQ_OK;
}

int pq_reshape(pq_ctx* c, pq_tensor h, int rank, const int64_t* dims) {
  if (!c || rank < 0 || rank > PQ_MAX_RANK || (rank && !dims)) return fail(c, PQ_ERR_ARG, "bad reshape");
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  int64_t n = 1;
  for (int i = 0; i < rank; ++i) n *= dims[i];
  if (n != numel(*t)) return fail(c, PQ_ERR_ARG, "reshape changes the number of elements");
  if (!is_contiguous(*t)) {
    Tensor m;
    int r = materialize_into(c, *t, m);
    if (r) return r;
    unref(c, t->buf);
    *t = m;
  }
  t->rank = rank;
  for (int i = 0; i < rank; ++i) t->dims[i] = dims[i];
  set_contiguous(*t);
  return PQ_OK;
}

int pq_view(pq_ctx* c, pq_tensor h, int axis, int64_t lo, int64_t hi, pq_tensor* out) {
  if (!c || !out) return PQ_ERR_ARG;
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  if (axis < 1 || axis > t->rank) return fail(c, PQ_ERR_ARG, "axis out of range");
  if (lo < 1 || hi > t->dims[axis - 1] || hi < lo - 1) return fail(c, PQ_ERR_ARG, "range out of bounds");
  Tensor r = *t;
  r.off += (lo - 1) * t->strides[axis - 1];
  r.dims[axis - 1] = hi - lo + 1;
  r.buf->refs++;
  *out = add(c, r);
  return PQ_OK;
}

int pq_materialize(pq_ctx* c, pq_tensor h, pq_tensor* out) {
  if (!c || !out) return PQ_ERR_ARG;
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  Tensor m;
  int r = materialize_into(c, *t, m);
  if (r) return r;
  *out = add(c, m);
  return PQ_OK;
}

int pq_tensor_device_ptr(pq_ctx* c, pq_tensor h, void** base, int64_t* off, int64_t* strides) {
  if (!c) return PQ_ERR_ARG;
  Tensor* t = get(c, h);
  if (!t) return fail(c, PQ_ERR_NOT_FOUND, "unknown tensor handle");
  if (base) *base = t->buf->ptr;
  if (off) *off = t->off;
  if (strides)
    for (int i = 0; i < t->rank; ++i) strides[i] = t->strides[i];
  return PQ_OK;
}

int pq_profile_enable(pq_ctx* c, int enable) {
  if (!c) return PQ_ERR_ARG;
  hipStreamSynchronize(c->stream);
  if (enable) {
    for (auto& p : c->pe) {
      for (auto& ev : p.events) {
        hipEventDestroy(ev.first);
        hipEventDestroy(ev.second);
      }
      p.events.clear();
      p.launches = 0;
      p.bytes = p.flops = 0;
    }
  }
  c->prof = enable != 0;
  return PQ_OK;
}

int pq_profile_read(pq_ctx* c, int kc, int64_t* launches, double* total_ms, double* bytes, double* flops) {
  if (!c || kc < 0 || kc >= PQ_KC_COUNT) return PQ_ERR_ARG;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hipfail(c, e, "profile sync");
  double ms = 0;
  for (auto& ev : c->pe[kc].events) {
    float x = 0;
    if (hipEventElapsedTime(&x, ev.first, ev.second) == hipSuccess) ms += x;
  }
  if (launches) *launches = c->pe[kc].launches;
  if (total_ms) *total_ms = ms;
  if (bytes) *bytes = c->pe[kc].bytes;
  if (flops) *flops = c->pe[kc].flops;
  return PQ_OK;
}

int pq_set_gemm_tile(pq_ctx* c, int tile) {
  if (!c || tile < -1 || tile >= PQ_GEMM_TILES) return PQ_ERR_ARG;
  c->gemm_tile = tile;
  return PQ_OK;
}

int pq_set_arithmetic(pq_ctx* c, int mode) {
  if (!c || mode < 0 || mode > 1) return PQ_ERR_ARG;
  c->arith = mode;
  return PQ_OK;
}

int pq_set_gemm_orientation(pq_ctx* c, int mode) {
  if (!c || mode < -1 || mode > 1) return PQ_ERR_ARG;
  c->gemm_orient = mode;
  return PQ_OK;
}

int pq_set_chain_pairs(pq_ctx* c, int mode) {
  if (!c || mode < -1 || mode > 0) return PQ_ERR_ARG;
  c->chain_pairs = mode;
  return PQ_OK;
}

int pq_set_kernel_override(pq_ctx* c, int kc) {
  if (!c || kc < -1 || kc >= PQ_KC_COUNT) return PQ_ERR_ARG;
  c->override_kc = kc;
  return PQ_OK;
}

}  // extern "C"
