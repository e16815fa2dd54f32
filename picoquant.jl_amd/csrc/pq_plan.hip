// Planned contraction executor: a whole contraction plan (the pair plan of
// contract_network!, src/layer2.jl:294-321) over all slices of a sliced
// network (src/layer2/slicing.jl), run as one stream-ordered program that
// can be captured into a single hipGraph.
//
//  * Slots: every leaf and intermediate tensor, each with its own element
//    type (complex128 or complex64).  Leaves are context tensors (holding
//    their sliced bonds); intermediates live in one arena whose offsets come
//    from a liveness pass over the launch sequence.
//  * Steps: contractions (A x B -> C, all three of one element type) and
//    conversions (A -> C, same layout, element type change) -- so small
//    steps can run in FP64 and the bulk in complex64.
//  * Slices: each sliced bond has extent 2.  Slice s fixes bond j to bit j
//    of s.  Steps whose operands do not depend on any sliced bond run once
//    (phase 1); the others (phase 2) run once per slice group, batched over
//    the 2^group_bits slices of the group on grid z, reading batched
//    intermediates through entry strides and slots that keep sliced bonds as
//    indices (leaves, and small intermediates computed once for both values
//    of their bonds) through per-bond strides.
//  * Small steps of one dependency level and element type share one grouped
//    launch (the world lines of all qubits advance together).
//  * The per-slice result is summed over slices in FP64, deterministic in
//    slice order.
#include "pq_rt.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <functional>
#include <set>

using namespace pqrt;

namespace {

// out[j] = (overwrite ? 0 : out[j]) + sum_b C[b * stride + j] for b < nb,
// in FP64 and in a fixed order (the same bits on every run).  A workgroup of 256 lanes takes tj
// consecutive outputs with tb = 256 / tj lanes each (tj a power of two,
// chosen by the host: few outputs with many copies give one output 256
// lanes, many outputs give 256 outputs one lane each): lane (b0, j) sums
// entries b = b0, b0 + tb, ... of its output (the tj lanes of one b0 read
// adjacent elements), then the tb partial sums of each output are added
// pairwise in LDS.
template <typename C>
__global__ __launch_bounds__(256) void k_accum(const C* __restrict__ src, int64_t elems, int64_t stride, int nb,
                                               int log_tj, int overwrite, double* __restrict__ out) {
  __shared__ double pr[256], pi[256];
  const int tj = 1 << log_tj, tb = 256 >> log_tj;
  const int lj = threadIdx.x & (tj - 1), b0 = threadIdx.x >> log_tj;
  const int64_t j = (int64_t)blockIdx.x * tj + lj;
  double re = 0, im = 0;
  if (j < elems) {
    for (int b = b0; b < nb; b += tb) {
      C v = src[b * stride + j];
      re += (double)v.x;
      im += (double)v.y;
    }
  }
  pr[threadIdx.x] = re;
  pi[threadIdx.x] = im;
  __syncthreads();
  for (int h = 128; h >= tj; h >>= 1) {
    if ((int)threadIdx.x < h) {
      pr[threadIdx.x] += pr[threadIdx.x + h];
      pi[threadIdx.x] += pi[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (b0 == 0 && j < elems) {
    // the first slice group of an output batch writes it, the others add
    out[2 * j] = (overwrite ? 0.0 : out[2 * j]) + pr[lj];
    out[2 * j + 1] = (overwrite ? 0.0 : out[2 * j + 1]) + pi[lj];
  }
}

// log2 of the outputs per k_accum workgroup (tj = 256 / tb): tb, the lanes
// per output, is about nb / 4 (up to 256), so each lane adds about four
// entries; with 1024 outputs or more at most 16, so that a workgroup's
// loads cover 16 or more adjacent outputs.
int accum_log_tj(int64_t elems, int nb) {
  int tb = 1;
  while (tb < 256 && (int64_t)tb * 4 < nb) tb <<= 1;
  if (elems >= 1024 && tb > 16) tb = 16;
  int lt = 0;
  while ((256 >> lt) > tb) ++lt;
  return lt;
}

// Element type conversion, n elements, same layout.
template <typename S, typename D>
__global__ void k_convert(const S* __restrict__ src, D* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    S v = src[i];
    D w;
    w.x = (typename real_of<D>::type)v.x;
    w.y = (typename real_of<D>::type)v.y;
    dst[i] = w;
  }
}

size_t esize_of(int dtype) { return dtype == PQ_C128 ? 16 : 8; }

struct Slot {
  int64_t elems = 0;      // per batch entry
  uint64_t mask = 0;      // sliced bonds the slot depends on
  uint64_t low = 0;       // ... of them within a slice group: one copy per value
  bool batched = false;   // mask != 0
  int64_t copies() const { return (int64_t)1 << __builtin_popcountll(low); }
  int dtype = PQ_C128;
  pq_tensor leaf = 0;
  Buf* leaf_buf = nullptr;            // leaves: the buffer, referenced by the plan until it is destroyed
  int64_t leaf_off = 0;               // leaves: element offset of the tensor in leaf_buf
  int64_t leaf_span = 0;              // leaves: largest element offset the tensor's view reaches
  std::vector<int64_t> slice_stride;  // unbatched slots: stride of each sliced bond kept (0 = absent)
  int64_t arena_off = -1;             // intermediates (bytes)
  size_t region = 0;                  // intermediates: bytes of the arena region (all copies)
  int producer = -1;
  int last_use = -1;
};

struct PStep {
  int kind = 0;            // 0 contract, 1 convert, 2 fused chain
  int a, b, c;
  std::vector<int> ops;    // fused chain: operator slot of each stage
  ChainArgs chain;         // fused chain: program (pointers set at launch)
  bool dep;                // depends on a sliced bond (batched / selects per slice)
  bool phase2 = false;     // depends on a bond above the slice group: rerun per group
  StepSpec spec;           // resolved; batch strides for phase 2 steps
  double flops = 0, bytes = 0;
  int level = 0;           // 1 + max level of the producers of A and B
};

// One entry of the launch sequence: a single step, or a grouped launch of
// small steps of one dependency level (descriptors at desc_off).
struct Launch {
  bool grouped = false;
  bool forest = false;     // grouped as independent trees, one workgroup each (k_forest)
  int step = -1;
  std::vector<int> steps;  // grouped: member steps (forest: tree after tree, each in dependency order)
  std::vector<int> tree_start;  // forest: offset of each tree in steps, then steps.size()
  int desc_off = 0;
  int dtype = PQ_C128;
  unsigned blocks_x = 1, batch = 1;
};

// Where a launch of a phase's sequence runs: its stream, the launches of the
// same sequence it waits for on other streams, and whether a later launch
// waits for it (an event is recorded after it).
struct Sched {
  int stream = 0;
  std::vector<int> waits;
  bool record = false;
};

}  // namespace

struct pq_plan {
  std::vector<Slot> slots;
  std::vector<PStep> steps;
  std::vector<int> order1, order2;   // step indices of phase 1 / phase 2
  std::vector<Launch> seq1, seq2;    // launch sequences of phase 1 / phase 2
  MultiDesc* descs = nullptr;        // device copies of the grouped descriptors
  int32_t* chain_tab = nullptr;      // device copy of the fused-chain address tables
  void* forest_buf = nullptr;        // the forest launch's tables (steps, items, K offsets, levels)
  const ForestStep* f_steps = nullptr;
  const int4* f_items = nullptr;
  const int2* f_ktab = nullptr;
  const ForestCopy* f_copies = nullptr;
  const int32_t* f_lv = nullptr;
  size_t f_lds = 0;                  // forest launch: LDS bytes of the largest tree (0: none)
  int n_sliced = 0, group_bits = 0;
  int64_t n_groups = 1;
  struct OutQubit {
    int slot, bit;                   // leaf slot of a closed output qubit, its bit in a bitstring word
  };
  std::vector<OutQubit> outq;        // a bitstring batch's closed output qubits (pq_plan_set_bitstrings)
  int n_out_bits = 0;                // the top n_out_bits sliced bonds number output batches
  int64_t n_out = 1;                 // 2^n_out_bits results of out_elems each
  int out_slot = -1;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  double* out_dev = nullptr;
  int64_t out_elems = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t launches = 0;
  double alg_flops = 0, alg_bytes = 0;
  int nstreams = 1;                  // streams the launches spread over (1 = all on the context's)
  std::vector<Sched> sch[2];         // per launch of seq1 / seq2
  std::vector<hipEvent_t> lev[2];    // recorded after the launches others wait for
  hipEvent_t fork_ev[2] = {}, join_ev[2][PQ_PLAN_STREAMS] = {};
  std::vector<hipEvent_t> ev;  // per-step event pairs for pq_plan_profile
  std::vector<char> ev_rec;    // which pairs the last profiled run recorded
};

namespace {

// Address of slot s for slice group `group`: slots that keep sliced bonds as
// indices are offset by the group's high bits (phase 1 runs with group 0).
void* slot_ptr(pq_ctx* c, pq_plan* p, int s, int64_t group) {
  Slot& sl = p->slots[s];
  int64_t off = 0;
  for (int j = p->group_bits; j < p->n_sliced; ++j)
    if ((group >> (j - p->group_bits)) & 1) off += sl.slice_stride[j];
  if (sl.leaf) return (unsigned char*)sl.leaf_buf->ptr + (sl.leaf_off + off) * c->esize;
  return (unsigned char*)p->arena + sl.arena_off + off * (int64_t)esize_of(sl.dtype);
}

int64_t span_of(const std::vector<Mode>& ms, bool first) {
  int64_t s = 0;
  for (auto& m : ms) s += (m.ext - 1) * (first ? m.s0 : m.s1);
  return s;
}

// Frees the plan's device memory and releases its references on leaf
// buffers (returned to the context's pool when nothing else holds them).
void free_plan(pq_ctx* c, pq_plan* p) {
  for (Slot& sl : p->slots)
    if (sl.leaf_buf) unref(c, sl.leaf_buf);
  if (p->arena) hipFree(p->arena);
  if (p->scratch) hipFree(p->scratch);
  if (p->out_dev) hipFree(p->out_dev);
  if (p->descs) hipFree(p->descs);
  if (p->chain_tab) hipFree(p->chain_tab);
  if (p->forest_buf) hipFree(p->forest_buf);
  for (auto& v : p->lev)
    for (hipEvent_t e : v) hipEventDestroy(e);
  for (int ph = 0; ph < 2; ++ph) {
    if (p->fork_ev[ph]) hipEventDestroy(p->fork_ev[ph]);
    for (hipEvent_t e : p->join_ev[ph])
      if (e) hipEventDestroy(e);
  }
  delete p;
}

// Slots launch L reads and writes, and whether it uses the plan's split-K
// scratch.
void launch_slots(const pq_plan* p, const Launch& L, std::vector<int>& rd, std::vector<int>& wr, bool& scratch,
                  size_t esize_ctx) {
  rd.clear();
  wr.clear();
  scratch = false;
  const std::vector<int> one{L.step};
  for (int i : (L.grouped ? L.steps : one)) {
    const PStep& st = p->steps[i];
    rd.push_back(st.a);
    if (st.kind == 0) rd.push_back(st.b);
    if (st.kind == 2) rd.insert(rd.end(), st.ops.begin(), st.ops.end());
    wr.push_back(st.c);
    if (!L.grouped && st.kind == 0 && step_scratch_bytes(st.spec, esize_of(st.spec.dtype)) > 0) scratch = true;
  }
  (void)esize_ctx;
}

// Schedule of one phase's launch sequence over nstreams streams.  Launch t
// waits for (1) the launches of the sequence that produce what it reads,
// (2) every earlier launch of the sequence that read or wrote an arena slot
// whose region its outputs reuse (the liveness pass only keeps regions of
// slots live at the same position apart), and (3) the previous user of the
// split-K scratch.  Each launch goes on the stream whose last launch is its
// latest dependency (no wait needed for it), else on the least recently used
// stream; the other dependencies become event waits, one per stream (the
// latest: a stream runs its launches in order).
void make_schedule(const pq_plan* p, const std::vector<Launch>& seq, int nstreams, std::vector<Sched>& out) {
  const int n = (int)seq.size();
  out.assign(n, Sched());
  std::map<int, int> prod;                  // slot -> producing launch
  std::map<int, std::vector<int>> touch;    // arena slot -> launches reading or writing it
  std::vector<int> rd, wr, last_on(nstreams, -1);
  int last_scratch = -1;
  for (int t = 0; t < n; ++t) {
    bool scr = false;
    launch_slots(p, seq[t], rd, wr, scr, 0);
    std::set<int> deps;
    for (int x : rd) {
      auto it = prod.find(x);
      if (it != prod.end() && it->second != t) deps.insert(it->second);
    }
    for (int cs : wr) {
      const Slot& C = p->slots[cs];
      if (C.arena_off < 0) continue;
      const size_t c0 = (size_t)C.arena_off, c1 = c0 + C.region;
      for (auto& kv : touch) {
        if (kv.first == cs) continue;
        const Slot& X = p->slots[kv.first];
        const size_t x0 = (size_t)X.arena_off, x1 = x0 + X.region;
        if (x0 < c1 && c0 < x1)
          for (int u : kv.second)
            if (u != t) deps.insert(u);
      }
    }
    if (scr) {
      if (last_scratch >= 0) deps.insert(last_scratch);
      last_scratch = t;
    }
    // stream: the one whose last launch is the latest dependency, if any
    int s = -1, best = -1;
    for (int u : deps)
      if (u > best && last_on[out[u].stream] == u) {
        best = u;
        s = out[u].stream;
      }
    if (s < 0) {
      s = 0;
      for (int q = 1; q < nstreams; ++q)
        if (last_on[q] < last_on[s]) s = q;
    }
    out[t].stream = s;
    std::vector<int> latest(nstreams, -1);
    for (int u : deps)
      if (out[u].stream != s) latest[out[u].stream] = std::max(latest[out[u].stream], u);
    for (int q = 0; q < nstreams; ++q)
      if (latest[q] >= 0) {
        out[t].waits.push_back(latest[q]);
        out[latest[q]].record = true;
      }
    last_on[s] = t;
    for (int x : wr) prod[x] = t;
    for (int x : rd)
      if (p->slots[x].arena_off >= 0) touch[x].push_back(t);
    for (int x : wr)
      if (p->slots[x].arena_off >= 0) touch[x].push_back(t);
  }
}

// Whether `span` (largest element offset read or written) stays inside slot s.
bool slot_holds(const Slot& sl, int64_t span) {
  if (!sl.leaf) return span < sl.elems;
  return span <= sl.leaf_span;
}

// Appends the address table of a bit list (bits [b0, b1)) to tab: entry t
// is, over the set bits b of t, the XOR of lay[position of b] (pos = true:
// an LDS tile address) or the sum of the strides (pos = false: a tensor
// offset).  Returns the table's offset.
int32_t bit_table(std::vector<int32_t>& tab, const std::vector<int32_t>& vals, int b0, int b1, bool pos,
                  const std::vector<int32_t>& lay) {
  const int32_t off = (int32_t)tab.size();
  for (int64_t t = 0; t < ((int64_t)1 << (b1 - b0)); ++t) {
    int64_t v = 0;
    for (int b = b0; b < b1; ++b)
      if ((t >> (b - b0)) & 1) v = pos ? (v ^ lay[vals[b]]) : v + (int64_t)vals[b];
    tab.push_back((int32_t)v);
  }
  return off;
}

// GF(2) rank of the low `bits` bits of vs[0..n).
int low_rank(const int32_t* vs, int n, int bits) {
  int32_t w[16];
  for (int i = 0; i < n; ++i) w[i] = vs[i] & ((1 << bits) - 1);
  int r = 0;
  for (int b = 0; b < bits && r < n; ++b) {
    int piv = -1;
    for (int i = r; i < n; ++i)
      if ((w[i] >> b) & 1) {
        piv = i;
        break;
      }
    if (piv < 0) continue;
    std::swap(w[r], w[piv]);
    for (int i = 0; i < n; ++i)
      if (i != r && ((w[i] >> b) & 1)) w[i] ^= w[r];
    ++r;
  }
  return r;
}

// LDS layout of a fused-chain tile: tile bit p is stored at address bit
// vector lay[p] (the address of an element is the XOR over its set bits),
// with lay[p] = 2^p plus bits below min(p, 6), so the map is invertible and
// stays inside the tile.  Every LDS access of the kernel has its 64 lanes
// varying in six known tile bits (gather: the first six gathered bits;
// scatter: the first six scattered; stage reads: four R bits and K bits 0-1;
// stage writes: four R bits and M bits 2-3).  Lanes 0-15, 0-31 and 0-63 of
// an access then hit distinct 8-byte bank slots exactly when the low 4, 5
// and 6 address bits of those tile bits are linearly independent; a
// coordinate search over the low bits of each lay[p] minimises the lane
// collisions summed over all accesses.
std::vector<int32_t> chain_layout(int tb, const std::vector<std::vector<int32_t>>& sets) {
  std::vector<int32_t> lay(tb);
  for (int p = 0; p < tb; ++p) lay[p] = 1 << p;
  auto cost = [&]() {
    int c = 0;
    int32_t v[16];
    for (const auto& s : sets)
      for (int g = 4; g <= 6; ++g) {
        if ((int)s.size() < g) continue;
        for (int i = 0; i < g; ++i) v[i] = lay[s[i]];
        c += (1 << (g - low_rank(v, g, g))) - 1;
      }
    return c;
  };
  int best = cost();
  for (int round = 0; round < 4 && best > 0; ++round) {
    bool better = false;
    for (int p = 0; p < tb; ++p) {
      const int lowbits = std::min(p, 6);
      const int32_t keep = lay[p];
      int32_t pick = keep;
      for (int32_t l = 0; l < (1 << lowbits); ++l) {
        lay[p] = (1 << p) | l;
        const int c = cost();
        if (c < best) {
          best = c;
          pick = lay[p];
          better = true;
        }
      }
      lay[p] = pick;
    }
    if (!better) break;
  }
  return lay;
}

// Reads and checks the program of fused-chain step i (format: PQ_CHAIN_* in
// pq_hip.h) into st.chain / st.ops and its address tables into tab.
// Returns an error message or nullptr.
const char* parse_chain(pq_ctx* c, pq_plan* p, const pq_plan_desc* d, int i, PStep& st,
                        const std::vector<int>& produced, std::vector<int32_t>& tab) {
  if (!d->step_chain || !d->chain_words) return "fused chain without a program";
  int64_t q = d->step_chain[i];
  bool ok = true;
  auto w = [&]() -> int32_t {
    if (q < 0 || q >= d->n_chain_words) {
      ok = false;
      return 0;
    }
    return d->chain_words[q++];
  };
  ChainArgs& a = st.chain;
  memset(&a, 0, sizeof(a));
  const int L = w(), tb = w(), nin = w(), nout = w(), nspec = w();
  if (!ok || L < 1 || L > CH_ST || tb < 0 || tb > CH_TB || nin < 0 || nin > tb || nout < 0 || nout > tb ||
      nspec < 0 || nspec > PQ_MAX_MODES)
    return "bad fused-chain header";
  a.tb = tb;
  a.nst = L;
  const Slot &X = p->slots[st.a], &Y = p->slots[st.c];
  // bit lists: tile position and stride per bit; no position twice
  auto bits = [&](int n, std::vector<int32_t>& pos, std::vector<int32_t>* str, int64_t& span,
                  uint32_t& used) -> bool {
    pos.resize(n);
    if (str) str->resize(n);
    for (int b = 0; b < n; ++b) {
      pos[b] = w();
      const int32_t sv = str ? w() : 0;
      if (!ok || pos[b] < 0 || pos[b] >= tb || sv < 0 || ((used >> pos[b]) & 1)) return false;
      used |= 1u << pos[b];
      if (str) (*str)[b] = sv;
      span += sv;
    }
    return true;
  };
  std::vector<int32_t> ipos, istr, opos, ostr;
  int64_t xspan = 0, yspan = 0, unused = 0;
  uint32_t used = 0;
  if (!bits(nin, ipos, &istr, xspan, used)) return "bad gathered bit";
  used = 0;
  if (!bits(nout, opos, &ostr, yspan, used)) return "bad scattered bit";
  std::vector<Mode> spec;
  for (int m = 0; m < nspec; ++m) {
    Mode md;
    md.ext = w();
    md.s0 = w();
    md.s1 = w();
    if (md.ext < 1 || md.s0 < 0 || md.s1 < 0) return "bad spectator mode";
    xspan += (md.ext - 1) * md.s0;
    yspan += (md.ext - 1) * md.s1;
    spec.push_back(md);
  }
  if (!ok || !fill_group(a.spec, spec)) return "bad spectator modes";
  if (X.dtype != PQ_C64 || Y.dtype != PQ_C64) return "fused chain needs complex64 tensors";
  // A chain runs once per slice group at most: its tensors may depend on
  // bonds above the group (the step then reruns per group, at the group's
  // addresses) but not be batched over the slices of a group.
  if (X.low || Y.low) return "fused chain on a tensor batched over the slices of a group";
  if (X.mask & ~Y.mask) return "fused chain output independent of a bond its input depends on";
  for (int64_t v : X.slice_stride)
    if (v) return "fused chain input keeps a sliced bond";
  for (int64_t v : Y.slice_stride)
    if (v) return "fused chain output keeps a sliced bond";
  if (!slot_holds(X, xspan)) return "gather out of range";
  if (!slot_holds(Y, yspan)) return "scatter out of range";
  if (((int64_t)a.spec.total << nout) != Y.elems) return "scatter does not cover the output";
  // Each side moves element pairs along one bit (the stride-1 bit if there
  // is one, else the smallest stride; none for a side without bits: the
  // pair is one element twice); the other bits go to the lane and
  // iteration tables.
  auto pair_bit = [&](const std::vector<int32_t>& str, int n) -> int {
    int best = -1;
    for (int bb = 0; bb < n; ++bb)
      if (best < 0 || str[bb] < str[best]) best = bb;
    return best;
  };
  const int gb = pair_bit(istr, nin), sb = pair_bit(ostr, nout);
  auto even_rest = [&](const std::vector<int32_t>& str, int n, int pb, bool xside) -> bool {
    if (pb < 0 || str[pb] != 1) return false;
    for (int bb = 0; bb < n; ++bb)
      if (bb != pb && (str[bb] & 1)) return false;
    for (const Mode& md : spec)
      if (((xside ? md.s0 : md.s1) & 1) && md.ext > 1) return false;
    return true;
  };
  a.gvec_ok = even_rest(istr, nin, gb, true) ? 1 : 0;
  a.svec_ok = even_rest(ostr, nout, sb, false) ? 1 : 0;
  a.gps = gb >= 0 ? istr[gb] : 0;
  a.sps = sb >= 0 ? ostr[sb] : 0;
  std::vector<int32_t> ipos2, istr2, opos2, ostr2;
  for (int bb = 0; bb < nin; ++bb)
    if (bb != gb) {
      ipos2.push_back(ipos[bb]);
      istr2.push_back(istr[bb]);
    }
  for (int bb = 0; bb < nout; ++bb)
    if (bb != sb) {
      opos2.push_back(opos[bb]);
      ostr2.push_back(ostr[bb]);
    }
  const int nin2 = (int)ipos2.size(), nout2 = (int)opos2.size();
  a.ilo = std::min(nin2, CH_LO_BITS);
  a.ihi = nin2 - a.ilo;
  a.olo = std::min(nout2, CH_LO_BITS);
  a.ohi = nout2 - a.olo;
  std::vector<std::vector<int32_t>> kps(L), kss(L), mps(L), mss(L), rps(L);
  st.ops.clear();
  for (int t = 0; t < L; ++t) {
    ChainStage& g = a.st[t];
    const int os = w();
    const int nk = w(), nm = w(), nr = w();
    if (!ok || os < 0 || os >= (int)p->slots.size() || nk < 0 || nk > CH_OB || nm < 0 || nm > CH_OB || nr < 0 ||
        nr > CH_TB)
      return "bad fused-chain stage";
    const Slot& S = p->slots[os];
    if (!S.leaf && produced[os] < 0) return "operator read before it is produced";
    if (S.dtype != PQ_C64 || S.low) return "fused-chain operator must be a complex64 tensor not batched in a group";
    if (S.mask & ~Y.mask) return "fused chain output independent of a bond its operator depends on";
    // bonds above the group that the operator keeps as indices: read at the
    // group's offset (every group's offset checked below)
    int64_t shigh = 0;
    for (int j = 0; j < p->n_sliced; ++j) {
      if (!S.slice_stride[j]) continue;
      if (j < p->group_bits) return "fused-chain operator keeps a sliced bond of the slice group";
      if (!((Y.mask >> j) & 1)) return "fused chain output independent of a bond its operator keeps";
      shigh += S.slice_stride[j];
    }
    std::vector<int32_t> kpos, kstr, mpos, mstr, rpos;
    int64_t sspan = 0;
    uint32_t kr = 0, mr = 0, rr = 0;
    if (!bits(nk, kpos, &kstr, sspan, kr)) return "bad K bit";
    if (!bits(nm, mpos, &mstr, sspan, mr)) return "bad M bit";
    if (!bits(nr, rpos, nullptr, unused, rr)) return "bad R bit";
    // R must be disjoint from K and from M: every (r, k) and (r, m) address
    // is then distinct across r, which the in-place update relies on
    if ((rr & kr) || (rr & mr)) return "R bits overlap K or M bits";
    if (!slot_holds(S, sspan + shigh)) return "operator access out of range";
    g.nk = nk;
    g.nm = nm;
    g.rlo = std::min(nr, 4);   // R bits from the lane (16 r's per MFMA block)
    g.rhi = nr - g.rlo;
    kps[t] = kpos;
    kss[t] = kstr;
    mps[t] = mpos;
    mss[t] = mstr;
    rps[t] = rpos;
    g.sd = a.sdn;
    a.sdn += 1 << (nk + nm);
    g.ko = a.idn;
    g.mo = a.idn + (1 << nk);
    g.ro = a.idn + (1 << nk) + (1 << nm);
    g.rho = g.ro + (1 << g.rlo);
    a.idn += (1 << nk) + (1 << nm) + (1 << g.rlo) + (1 << g.rhi);
    if ((int64_t)a.sdn * 8 + (int64_t)a.idn * 4 > CH_STAGE_BYTES) return "fused-chain operators exceed the LDS staging area";
    st.ops.push_back(os);
    st.level = std::max(st.level, 1 + (S.leaf ? 0 : p->steps[S.producer].level));
    p->slots[os].last_use = i;
    st.flops += 8.0 * (double)((int64_t)1 << (nr + nk + nm)) * (double)a.spec.total;
    st.bytes += 8.0 * (double)S.elems;
  }
  // A chain whose every operator has at most 4 chunk-blocks runs on the
  // kernel's small-stage form, which also has the lane-row stage: operators
  // of at most 4 x 4 with at least 64 r's take 6 R bits from the lane (one
  // r per lane on 4 x 4 x 1 MFMAs), when the larger R tables still fit the
  // LDS staging area.
  {
    bool small_chain = true;
    for (int t = 0; t < L; ++t) {
      const int nkc = ((1 << a.st[t].nk) + 3) / 4, nmb = ((1 << a.st[t].nm) + 15) / 16;
      small_chain = small_chain && nkc * nmb <= 4;
    }
    auto rows = [&](const ChainStage& g) { return g.nk <= 2 && g.nm <= 2 && g.rlo + g.rhi >= 6; };
    int32_t idn = 0;
    for (int t = 0; t < L; ++t) {
      const ChainStage& g = a.st[t];
      const int nr = g.rlo + g.rhi, rl = small_chain && rows(g) ? 6 : g.rlo;
      idn += (1 << g.nk) + (1 << g.nm) + (1 << rl) + (1 << (nr - rl));
    }
    if (small_chain && (int64_t)a.sdn * 8 + (int64_t)idn * 4 <= CH_STAGE_BYTES) {
      a.idn = 0;
      for (int t = 0; t < L; ++t) {
        ChainStage& g = a.st[t];
        const int nr = g.rlo + g.rhi;
        if (rows(g)) {
          g.rlo = 6;
          g.rhi = nr - 6;
        }
        g.ko = a.idn;
        g.mo = a.idn + (1 << g.nk);
        g.ro = g.mo + (1 << g.nm);
        g.rho = g.ro + (1 << g.rlo);
        a.idn += (1 << g.nk) + (1 << g.nm) + (1 << g.rlo) + (1 << g.rhi);
      }
    }
  }
  // tile layout for the accesses' lane bits, then the address tables
  std::vector<std::vector<int32_t>> sets;
  sets.push_back(std::vector<int32_t>(ipos2.begin(), ipos2.begin() + std::min(nin2, 6)));
  sets.push_back(std::vector<int32_t>(opos2.begin(), opos2.begin() + std::min(nout2, 6)));
  for (int t = 0; t < L; ++t) {
    const int rl = a.st[t].rlo;
    std::vector<int32_t> rd(rps[t].begin(), rps[t].begin() + rl), wr = rd;
    if (rl == 4) {
      for (int b = 0; b < std::min<int>(2, (int)kps[t].size()); ++b) rd.push_back(kps[t][b]);
      for (int b = 2; b < std::min<int>(4, (int)mps[t].size()); ++b) wr.push_back(mps[t][b]);
    }
    sets.push_back(rd);
    sets.push_back(wr);
  }
  const std::vector<int32_t> lay = chain_layout(tb, sets);
  a.gpa = gb >= 0 ? lay[ipos[gb]] : 0;
  a.spa = sb >= 0 ? lay[opos[sb]] : 0;
  a.ilp = bit_table(tab, ipos2, 0, a.ilo, true, lay);
  a.ilg = bit_table(tab, istr2, 0, a.ilo, false, lay);
  a.ihp = bit_table(tab, ipos2, a.ilo, nin2, true, lay);
  a.ihg = bit_table(tab, istr2, a.ilo, nin2, false, lay);
  a.olp = bit_table(tab, opos2, 0, a.olo, true, lay);
  a.olg = bit_table(tab, ostr2, 0, a.olo, false, lay);
  a.ohp = bit_table(tab, opos2, a.olo, nout2, true, lay);
  a.ohg = bit_table(tab, ostr2, a.olo, nout2, false, lay);
  for (int t = 0; t < L; ++t) {
    ChainStage& g = a.st[t];
    const int nk = g.nk, nm = g.nm, nr = g.rlo + g.rhi;
    g.kp = bit_table(tab, kps[t], 0, nk, true, lay);
    g.sk = bit_table(tab, kss[t], 0, nk, false, lay);
    g.mp = bit_table(tab, mps[t], 0, nm, true, lay);
    g.sm = bit_table(tab, mss[t], 0, nm, false, lay);
    g.rl = bit_table(tab, rps[t], 0, g.rlo, true, lay);
    g.rh = bit_table(tab, rps[t], g.rlo, nr, true, lay);
  }
  // per-tile spectator offsets (X, Y), decoded here once instead of per tile
  // on the device; tile t's pair at tab[tof + 2 t]
  a.tof = -1;
  if (a.spec.total <= CH_TOF_TILES) {
    a.tof = (int32_t)tab.size();
    std::vector<int64_t> ext(spec.size());
    for (size_t m = 0; m < spec.size(); ++m) ext[m] = spec[m].ext;
    for (int64_t t = 0; t < (int64_t)a.spec.total; ++t) {
      int64_t r = t, ox = 0, oy = 0;
      for (size_t m = 0; m < spec.size(); ++m) {
        const int64_t v = r % ext[m];
        r /= ext[m];
        ox += v * spec[m].s0;
        oy += v * spec[m].s1;
      }
      tab.push_back((int32_t)ox);
      tab.push_back((int32_t)oy);
    }
  }
  st.bytes += 8.0 * (double)a.spec.total * (double)(((int64_t)1 << nin) + ((int64_t)1 << nout));
  st.spec.kc = PQ_KC_CHAIN;
  st.spec.tile = -1;
  st.dep = Y.batched;
  st.phase2 = (Y.mask >> p->group_bits) != 0;
  if (st.phase2) {
    st.flops *= (double)p->n_groups;
    st.bytes *= (double)p->n_groups;
  }
  return nullptr;
}

int run_program(pq_ctx* c, pq_plan* p, bool profile_steps) {
  // (no fill of the FP64 result: the first slice group of every output
  // batch writes it, k_accum below)
  auto do_launch = [&](const Launch& L, int64_t group) -> int {
    if (!L.grouped) {
      const int i = L.step;
      PStep& st = p->steps[i];
      if (profile_steps) hipEventRecord(p->ev[2 * i], c->stream);
      int r = PQ_OK;
      if (st.kind == 1) {
        const Slot& S = p->slots[st.a];
        const Slot& D = p->slots[st.c];
        int64_t n = D.elems * D.copies();
        unsigned gc = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
        const void* src = slot_ptr(c, p, st.a, group);
        void* dst = slot_ptr(c, p, st.c, group);
        ProfScope ps(c, PQ_KC_COPY, (double)n * (esize_of(S.dtype) + esize_of(D.dtype)), 0);
        if (S.dtype == PQ_C128 && D.dtype == PQ_C64)
          hipLaunchKernelGGL((k_convert<c128, c64>), dim3(std::max(1u, gc)), dim3(256), 0, c->stream,
                             (const c128*)src, (c64*)dst, n);
        else if (S.dtype == PQ_C64 && D.dtype == PQ_C128)
          hipLaunchKernelGGL((k_convert<c64, c128>), dim3(std::max(1u, gc)), dim3(256), 0, c->stream,
                             (const c64*)src, (c128*)dst, n);
        else if (S.dtype == PQ_C128)
          hipLaunchKernelGGL((k_convert<c128, c128>), dim3(std::max(1u, gc)), dim3(256), 0, c->stream,
                             (const c128*)src, (c128*)dst, n);
        else
          hipLaunchKernelGGL((k_convert<c64, c64>), dim3(std::max(1u, gc)), dim3(256), 0, c->stream,
                             (const c64*)src, (c64*)dst, n);
        if (hipGetLastError() != hipSuccess) r = fail(c, PQ_ERR_HIP, "convert kernel");
      } else if (st.kind == 2) {
        ChainArgs a = st.chain;
        a.tab = p->chain_tab;
        a.X = (const c64*)slot_ptr(c, p, st.a, group);
        a.Y = (c64*)slot_ptr(c, p, st.c, group);
        for (int t = 0; t < a.nst; ++t) a.st[t].S = (const c64*)slot_ptr(c, p, st.ops[t], group);
        a.gvec = c->chain_pairs != 0 && a.gvec_ok && ((uintptr_t)a.X % 16) == 0;
        a.svec = c->chain_pairs != 0 && a.svec_ok && ((uintptr_t)a.Y % 16) == 0;
        ProfScope ps(c, PQ_KC_CHAIN, st.bytes, st.flops);
        hipError_t e = launch_chain(a, c->stream);
        if (e != hipSuccess) r = hipfail(c, e, "fused chain kernel");
      } else {
        r = launch_step(c, slot_ptr(c, p, st.a, group), slot_ptr(c, p, st.b, group), slot_ptr(c, p, st.c, group),
                        st.spec, p->scratch);
      }
      if (profile_steps) {
        hipEventRecord(p->ev[2 * i + 1], c->stream);
        p->ev_rec[i] = 1;
      }
      return r;
    }
    // Grouped: the events of the first member step bracket the whole launch.
    if (profile_steps) hipEventRecord(p->ev[2 * L.steps[0]], c->stream);
    double fl = 0, by = 0;
    for (int i : L.steps) {
      fl += p->steps[i].flops;
      by += p->steps[i].bytes;
    }
    hipError_t e;
    {
      ProfScope ps(c, PQ_KC_BATCHED, by, fl);
      if (L.forest)
        e = launch_forest(p->f_steps, p->f_items, p->f_ktab, p->f_copies, p->f_lv, (int)L.tree_start.size() - 1,
                          p->f_lds,
                          c->stream);
      else
        e = (L.dtype == PQ_C128)
                ? launch_multi<c128>(p->descs + L.desc_off, (int)L.steps.size(), L.blocks_x, L.batch, group,
                                     c->stream)
                : launch_multi<c64>(p->descs + L.desc_off, (int)L.steps.size(), L.blocks_x, L.batch, group,
                                    c->stream);
    }
    if (profile_steps) {
      hipEventRecord(p->ev[2 * L.steps[0] + 1], c->stream);
      p->ev_rec[L.steps[0]] = 1;
    }
    if (e != hipSuccess) return hipfail(c, e, "grouped kernel");
    return PQ_OK;
  };
  // Launches spread over the schedule's streams (not when timing steps):
  // each phase (and each slice group of phase 2) forks the extra streams off
  // the context's stream and joins them back, so a captured graph holds the
  // independent launches as parallel branches.
  const bool conc = !profile_steps && p->nstreams > 1;
  hipStream_t main = c->stream;
  auto strm = [&](int q) { return q == 0 ? main : c->aux[q - 1]; };
  auto run_seq = [&](int ph, const std::vector<Launch>& seq, int64_t group) -> int {
    if (conc) {
      hipEventRecord(p->fork_ev[ph], main);
      for (int q = 1; q < p->nstreams; ++q) hipStreamWaitEvent(strm(q), p->fork_ev[ph], 0);
    }
    for (size_t t = 0; t < seq.size(); ++t) {
      if (conc) {
        const Sched& sc = p->sch[ph][t];
        for (int u : sc.waits) hipStreamWaitEvent(strm(sc.stream), p->lev[ph][u], 0);
        c->stream = strm(sc.stream);
      }
      const int r = do_launch(seq[t], group);
      if (conc) {
        if (p->sch[ph][t].record) hipEventRecord(p->lev[ph][t], c->stream);
        c->stream = main;
      }
      if (r) return r;
    }
    if (conc)
      for (int q = 1; q < p->nstreams; ++q) {
        hipEventRecord(p->join_ev[ph][q], strm(q));
        hipStreamWaitEvent(main, p->join_ev[ph][q], 0);
      }
    if (hipGetLastError() != hipSuccess) return fail(c, PQ_ERR_HIP, "stream scheduling");
    return PQ_OK;
  };
  {
    int r = run_seq(0, p->seq1, 0);
    if (r) return r;
  }
  const Slot& os = p->slots[p->out_slot];
  int nbatch = (int)os.copies();
  for (int64_t grp = 0; grp < p->n_groups; ++grp) {
    {
      int r = run_seq(1, p->seq2, grp);
      if (r) return r;
    }
    const void* src = slot_ptr(c, p, p->out_slot, grp);
    // the group's output batch: its bits above the summed slice bonds; the
    // batch's first group writes it
    const int summed = p->n_sliced - p->n_out_bits - p->group_bits;
    double* dst = p->out_dev + 2 * p->out_elems * (grp >> summed);
    const int first = (grp & (((int64_t)1 << summed) - 1)) == 0 ? 1 : 0;
    const int lt = accum_log_tj(os.elems, nbatch);
    const unsigned ga = (unsigned)((os.elems + (1 << lt) - 1) >> lt);
    ProfScope ps(c, PQ_KC_ACCUM, (double)os.elems * nbatch * esize_of(os.dtype) + 16.0 * os.elems, 0);
    if (os.dtype == PQ_C128)
      hipLaunchKernelGGL((k_accum<c128>), dim3(ga), dim3(256), 0, c->stream, (const c128*)src,
                         os.elems, os.elems, nbatch, lt, first, dst);
    else
      hipLaunchKernelGGL((k_accum<c64>), dim3(ga), dim3(256), 0, c->stream, (const c64*)src,
                         os.elems, os.elems, nbatch, lt, first, dst);
    if (hipGetLastError() != hipSuccess) return fail(c, PQ_ERR_HIP, "accumulate kernel");
  }
  return PQ_OK;
}

}  // namespace

extern "C" {

int pq_plan_create(pq_ctx* c, const pq_plan_desc* d, pq_plan** out) {
  if (!c || !d || !out) return PQ_ERR_ARG;
  if (d->n_slots <= 0 || d->n_steps < 0 || d->n_sliced < 0 || d->n_sliced > 40 || d->group_bits < 0 ||
      d->group_bits > d->n_sliced || d->group_bits > 16 || d->out_slot < 0 || d->out_slot >= d->n_slots)
    return fail(c, PQ_ERR_ARG, "bad plan header");
  pq_plan* p = new pq_plan;
  p->n_sliced = d->n_sliced;
  p->group_bits = d->group_bits;
  p->n_groups = (int64_t)1 << (d->n_sliced - d->group_bits);
  if (d->n_out_bits < 0 || d->n_out_bits > d->n_sliced - d->group_bits) {
    delete p;
    return fail(c, PQ_ERR_ARG, "output batch bonds must lie above the slice group");
  }
  p->n_out_bits = d->n_out_bits;
  p->n_out = (int64_t)1 << d->n_out_bits;
  p->out_slot = d->out_slot;
  auto bad = [&](const std::string& m) {
    free_plan(c, p);
    return fail(c, PQ_ERR_ARG, m);
  };
  p->slots.resize(d->n_slots);
  for (int s = 0; s < d->n_slots; ++s) {
    Slot& sl = p->slots[s];
    sl.elems = d->slot_elems[s];
    const uint64_t m = d->slot_batch_mask ? (uint64_t)d->slot_batch_mask[s] : 0;
    if (d->n_sliced < 64 && (m >> d->n_sliced) != 0) return bad("batch mask names a bond that is not sliced");
    sl.mask = m;
    sl.low = m & (((uint64_t)1 << d->group_bits) - 1);
    sl.batched = m != 0;
    sl.leaf = d->slot_leaf[s];
    sl.dtype = d->slot_dtype ? d->slot_dtype[s] : c->dtype;
    if (sl.dtype != PQ_C128 && sl.dtype != PQ_C64) return bad("bad slot element type");
    if (sl.elems < 1 || sl.elems >= (1ll << 31)) return bad("slot size out of range");
    sl.slice_stride.assign(d->n_sliced, 0);
    int64_t extra = 0;
    for (int j = 0; j < d->n_sliced; ++j) {
      const int64_t v = d->slot_slice_stride ? d->slot_slice_stride[(int64_t)s * d->n_sliced + j] : 0;
      if (v < 0) return bad("negative slice stride");
      if (v && sl.batched) return bad("a batched slot cannot keep sliced bonds");
      sl.slice_stride[j] = v;
      extra += v;
    }
    if (sl.leaf) {
      Tensor* t = get(c, sl.leaf);
      if (!t) return bad("unknown leaf tensor");
      if (sl.batched) return bad("a leaf cannot be batched");
      if (sl.dtype != c->dtype) return bad("a leaf must have the context's element type");
      int64_t span = 0;
      for (int i = 0; i < t->rank; ++i) span += (t->dims[i] - 1) * t->strides[i];
      if (span + 1 > (int64_t)(t->buf->bytes / c->esize) - t->off) return bad("leaf view exceeds its buffer");
      // the plan reads this buffer at this offset on every run: hold it, so
      // that freeing or reshaping the tensor later leaves the plan intact
      sl.leaf_buf = t->buf;
      sl.leaf_buf->refs++;
      sl.leaf_off = t->off;
      sl.leaf_span = span;
      if (extra > span) return bad("slice strides exceed the leaf");
    } else if (extra >= sl.elems) {
      return bad("slice strides exceed the slot");
    }
  }
  // Closed output qubits of a bitstring batch: contiguous [2, 2^n_out_bits]
  // leaves, one column per bitstring.
  if (d->n_out_qubits < 0 || (d->n_out_qubits > 0 && (!d->out_qubit_slot || !d->out_qubit_bit)))
    return bad("bad output-qubit list");
  for (int q = 0; q < d->n_out_qubits; ++q) {
    const int s = d->out_qubit_slot[q], b = d->out_qubit_bit[q];
    if (s < 0 || s >= d->n_slots || !p->slots[s].leaf) return bad("an output qubit's slot is not a leaf");
    if (b < 0 || b > 63) return bad("output qubit bit out of range");
    const Tensor* t = get(c, p->slots[s].leaf);
    if (!is_contiguous(*t) || numel(*t) != ((int64_t)2 << d->n_out_bits))
      return bad("an output qubit's leaf must be a contiguous [2, 2^n_out_bits] tensor");
    p->outq.push_back({s, b});
  }
  // Steps.
  int64_t mpos = 0;
  std::vector<int32_t> chain_tab;   // fused-chain address tables (uploaded below)
  std::vector<int> produced(d->n_slots, -1);
  p->steps.resize(d->n_steps);
  for (int i = 0; i < d->n_steps; ++i) {
    PStep& st = p->steps[i];
    st.kind = d->step_kind ? d->step_kind[i] : 0;
    if (st.kind < 0 || st.kind > 2) return bad("bad step kind");
    st.a = d->step_slots[3 * i];
    st.b = d->step_slots[3 * i + 1];
    st.c = d->step_slots[3 * i + 2];
    if (st.kind == 1) st.b = st.a;
    for (int x : {st.a, st.b, st.c})
      if (x < 0 || x >= d->n_slots) return bad("step slot out of range");
    if (p->slots[st.c].leaf || produced[st.c] >= 0) return bad("step output must be a fresh intermediate");
    for (int x : {st.a, st.b})
      if (!p->slots[x].leaf && produced[x] < 0) return bad("step reads a slot before it is produced");
    st.level = 1 + std::max(p->slots[st.a].leaf ? 0 : p->steps[p->slots[st.a].producer].level,
                            p->slots[st.b].leaf ? 0 : p->steps[p->slots[st.b].producer].level);
    produced[st.c] = i;
    p->slots[st.c].producer = i;
    p->slots[st.a].last_use = i;
    p->slots[st.b].last_use = i;
    const int nm = d->step_nmodes[3 * i], nn = d->step_nmodes[3 * i + 1], nk = d->step_nmodes[3 * i + 2];
    if (st.kind == 2 && (nm || nn || nk)) return bad("a fused chain has no modes of its own");
    auto take = [&](int n) {
      std::vector<Mode> v;
      for (int k = 0; k < n; ++k, ++mpos) v.push_back({d->step_modes[3 * mpos], d->step_modes[3 * mpos + 1],
                                                       d->step_modes[3 * mpos + 2]});
      return v;
    };
    st.spec.M = merge_modes(take(nm));
    st.spec.N = merge_modes(take(nn));
    st.spec.K = merge_modes(take(nk));
    for (auto* v : {&st.spec.M, &st.spec.N, &st.spec.K})
      for (auto& m : *v)
        if (m.ext < 1 || m.s0 < 0 || m.s1 < 0) return bad("bad mode");
    const Slot &A = p->slots[st.a], &B = p->slots[st.b], &Cs = p->slots[st.c];
    st.dep = Cs.batched;
    st.phase2 = (Cs.mask >> d->group_bits) != 0;
    if ((A.mask | B.mask) & ~Cs.mask) return bad("result independent of a sliced bond its operand depends on");
    if (st.kind == 2) {
      const char* why = parse_chain(c, p, d, i, st, produced, chain_tab);
      if (why) return bad(std::string("step ") + std::to_string(i) + ": " + why);
      p->alg_flops += st.flops;
      p->alg_bytes += st.bytes;
      (st.phase2 ? p->order2 : p->order1).push_back(i);
      continue;
    }
    if (st.kind == 1) {
      if (A.leaf || A.elems != Cs.elems || A.mask != Cs.mask || A.slice_stride != Cs.slice_stride)
        return bad("convert needs a same-shaped intermediate");
      const double n = (double)Cs.elems * (double)Cs.copies() * (st.phase2 ? (double)p->n_groups : 1.0);
      st.bytes = n * (esize_of(A.dtype) + esize_of(Cs.dtype));
      p->alg_bytes += st.bytes;
      (st.phase2 ? p->order2 : p->order1).push_back(i);
      continue;
    }
    if (A.dtype != Cs.dtype || B.dtype != Cs.dtype) return bad("contraction operands differ in element type");
    st.spec.dtype = Cs.dtype;
    // In-bounds check of every operand for one batch entry; a batched step
    // also reads slots that keep sliced bonds at every slice offset.
    auto fits = [&](const Slot& sl, int64_t span) -> bool {
      int64_t extra = 0;
      for (int j = 0; j < d->n_sliced; ++j)
        if ((Cs.mask >> j) & 1) extra += sl.slice_stride[j];
      if (!sl.leaf) return span + extra < sl.elems;
      return span + extra <= sl.leaf_span;
    };
    if (!fits(A, span_of(st.spec.M, true) + span_of(st.spec.K, true))) return bad("A access out of range");
    if (!fits(B, span_of(st.spec.N, true) + span_of(st.spec.K, false))) return bad("B access out of range");
    if (!fits(Cs, span_of(st.spec.M, false) + span_of(st.spec.N, false))) return bad("C access out of range");
    // Batch group over the slices of a group.
    // Batch group: one entry per value of the result's sliced bonds within
    // the slice group.  A batched operand is read at its own copy for those
    // of the bonds it depends on (stride 0 for the others); a slot that
    // keeps a bond as an index is read at the bond's stride.
    st.spec.L = single_batch();
    if (Cs.low) {
      Batch& L = st.spec.L;
      auto stride_of = [&](const Slot& sl, int j) -> int64_t {
        if ((sl.low >> j) & 1)
          return sl.elems << __builtin_popcountll(sl.low & (((uint64_t)1 << j) - 1));
        return sl.slice_stride[j];
      };
      L.nmodes = 0;
      for (int j = 0; j < d->group_bits; ++j) {
        if (!((Cs.low >> j) & 1)) continue;
        const int q = L.nmodes++;
        L.ext[q] = make_fdiv(2);
        L.sA[q] = stride_of(A, j);
        L.sB[q] = stride_of(B, j);
        L.sC[q] = stride_of(Cs, j);
      }
      L.total = (uint32_t)Cs.copies();
      // every batch entry stays inside its operand's copies
      int64_t offA = 0, offB = 0, offC = 0;
      for (int q = 0; q < L.nmodes; ++q) {
        offA += L.sA[q];
        offB += L.sB[q];
        offC += L.sC[q];
      }
      auto inside = [&](const Slot& sl, int64_t off, int64_t span) -> bool {
        if (sl.leaf) return off + span <= sl.leaf_span;
        return off + span < sl.elems * sl.copies();
      };
      if (!inside(A, offA, span_of(st.spec.M, true) + span_of(st.spec.K, true)) ||
          !inside(B, offB, span_of(st.spec.N, true) + span_of(st.spec.K, false)) ||
          !inside(Cs, offC, span_of(st.spec.M, false) + span_of(st.spec.N, false)))
        return bad("batch entries out of range");
      // A batch mode that one operand does not depend on is a free mode of
      // the other side (the shared operand is then read once, in one larger
      // product); modes that both operands or neither depend on stay batch
      // modes.  The offsets read and written are the same.
      Batch keep = single_batch();
      keep.nmodes = 0;
      for (int q = 0; q < L.nmodes; ++q) {
        if (L.sB[q] == 0 && L.sA[q] != 0 && (int)st.spec.M.size() < PQ_MAX_MODES) {
          st.spec.M.push_back({2, L.sA[q], L.sC[q]});
        } else if (L.sA[q] == 0 && L.sB[q] != 0 && (int)st.spec.N.size() < PQ_MAX_MODES) {
          st.spec.N.push_back({2, L.sB[q], L.sC[q]});
        } else {
          const int k = keep.nmodes++;
          keep.ext[k] = L.ext[q];
          keep.sA[k] = L.sA[q];
          keep.sB[k] = L.sB[q];
          keep.sC[k] = L.sC[q];
          keep.total *= 2;
        }
      }
      L = keep;
    }
    st.spec.arith = d->step_arith ? d->step_arith[i] : 0;
    if (st.spec.arith < 0 || st.spec.arith > 1) return bad("bad step arithmetic");
    st.spec.kc = d->step_kc ? d->step_kc[i] : -1;
    st.spec.tile = d->step_tile ? d->step_tile[i] : -1;
    int r = resolve_step(c, st.spec, esize_of(st.spec.dtype));
    if (r) {
      std::string msg = c->err;
      free_plan(c, p);
      return fail(c, r, "step " + std::to_string(i) + ": " + msg);
    }
    const double nbat = (double)st.spec.L.total * (st.phase2 ? (double)p->n_groups : 1.0);
    const double es = (double)esize_of(st.spec.dtype);
    st.flops = 8.0 * st.spec.Mt * st.spec.Nt * st.spec.Kt * nbat;
    st.bytes = es * (st.spec.Mt * st.spec.Kt + st.spec.Kt * st.spec.Nt + st.spec.Mt * st.spec.Nt) * nbat;
    p->alg_flops += st.flops;
    p->alg_bytes += st.bytes;
    (st.phase2 ? p->order2 : p->order1).push_back(i);
    p->scratch_bytes = std::max(p->scratch_bytes, step_scratch_bytes(st.spec, esize_of(st.spec.dtype)));
  }
  if (!p->slots[p->out_slot].leaf && produced[p->out_slot] < 0) return bad("output slot never produced");
  if (d->n_sliced > 0 && p->slots[p->out_slot].mask != ((d->n_sliced >= 64) ? ~0ull : (((uint64_t)1 << d->n_sliced) - 1)))
    return bad("sliced plan whose output does not depend on every sliced bond");
  // Launch sequences: by dependency level; the small contraction steps of a
  // level and element type share one grouped launch, the others launch alone.
  // A small complex128 -> complex64 conversion joins the complex128 group
  // of its level (the grouped kernel converts it element by element).
  auto conv_elems = [&](const PStep& st) -> int64_t {
    return p->slots[st.c].elems * p->slots[st.c].copies();
  };
  auto small_conv = [&](const PStep& st) {
    return st.kind == 1 && p->slots[st.a].dtype == PQ_C128 && p->slots[st.c].dtype == PQ_C64 &&
           conv_elems(st) <= 65536;
  };
  // Grouped: contractions of at most 2^17 outputs and K <= PQ_MULTI_KMAX,
  // whatever kernel they would run alone (small GEMMs too: two or more of
  // them per level share one launch instead of one launch each; on the 6x6
  // depth-14 plan 49 -> 36 launches and 1.40 -> 1.33 ms; at 2^18 / 2^19
  // outputs the grouped kernel's per-output loop loses more than the saved
  // launches: 1.38 / 1.50 ms).
  auto small = [&](const PStep& st) {
    return (st.kind == 0 &&
            (st.spec.kc == PQ_KC_STREAM || st.spec.kc == PQ_KC_GENERIC ||
             (st.spec.kc == PQ_KC_GEMM && st.spec.splitk <= 1)) &&
            st.spec.Mt * st.spec.Nt <= (1u << 17) && st.spec.Kt <= PQ_MULTI_KMAX) ||
           small_conv(st);
  };
  auto build_seq = [&](const std::vector<int>& order, std::vector<Launch>& seq) {
    std::vector<int> o = order;
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) { return p->steps[x].level < p->steps[y].level; });
    size_t i = 0;
    while (i < o.size()) {
      const int lvl = p->steps[o[i]].level;
      std::vector<int> grp[2];
      size_t j = i;
      for (; j < o.size() && p->steps[o[j]].level == lvl; ++j) {
        const PStep& st = p->steps[o[j]];
        const int ti = (small_conv(st) || st.spec.dtype == PQ_C128) ? 0 : 1;
        if (small(st) && (int)grp[ti].size() < 65535) {
          grp[ti].push_back(o[j]);
        } else {
          Launch L;
          L.step = o[j];
          seq.push_back(L);
        }
      }
      for (int ti = 0; ti < 2; ++ti) {
        if (grp[ti].size() == 1) {
          Launch L;
          L.step = grp[ti][0];
          seq.push_back(L);
        } else if (grp[ti].size() > 1) {
          Launch L;
          L.grouped = true;
          L.steps = grp[ti];
          L.dtype = ti == 0 ? PQ_C128 : PQ_C64;
          seq.push_back(L);
        }
      }
      i = j;
    }
  };
  // Phase 1's forest: the small unbatched steps whose operands are leaves or
  // results of other forest steps -- the world lines, which every qubit
  // contracts on its own.  They form independent trees (each result has one
  // reader), which one launch runs with a workgroup per tree, in place of a
  // grouped launch per dependency level.
  Launch forest;
  {
    std::vector<int> in_forest(d->n_steps, 0);
    std::vector<int> rest;
    std::vector<int> o = p->order1;
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) { return p->steps[x].level < p->steps[y].level; });
    std::vector<int> fs;
    // one workgroup runs a whole tree, so only steps of a few thousand
    // outputs (the world lines' gates and the first merges) join it; larger
    // small steps keep a grouped launch per level, with many workgroups each
    constexpr int64_t kForestOut = 2048, kForestK = 16;
    for (int i : o) {
      const PStep& st = p->steps[i];
      bool ok = small(st) && (st.kind == 1 ? conv_elems(st) <= kForestOut
                                           : st.spec.L.total == 1 && st.spec.Mt * st.spec.Nt <= (uint64_t)kForestOut &&
                                                 st.spec.Kt <= (uint64_t)kForestK);
      for (int x : {st.a, st.b}) {
        const Slot& sl = p->slots[x];
        if (!sl.leaf && !in_forest[sl.producer]) ok = false;
        if (sl.leaf && st.kind == 0 && sl.dtype != st.spec.dtype) ok = false;
      }
      if (ok) {
        in_forest[i] = 1;
        fs.push_back(i);
      } else {
        rest.push_back(i);
      }
    }
    if (fs.size() >= 2) {
      // trees: union of each step with the forest steps that produce its operands
      std::vector<int> parent(d->n_steps);
      for (int i = 0; i < d->n_steps; ++i) parent[i] = i;
      std::function<int(int)> root = [&](int x) { return parent[x] == x ? x : parent[x] = root(parent[x]); };
      for (int i : fs)
        for (int x : {p->steps[i].a, p->steps[i].b}) {
          const Slot& sl = p->slots[x];
          if (!sl.leaf) parent[root(sl.producer)] = root(i);
        }
      std::map<int, std::vector<int>> trees;   // by root; steps in dependency order (fs is level-sorted)
      for (int i : fs) trees[root(i)].push_back(i);
      forest.grouped = true;
      forest.forest = true;
      for (auto& kv : trees) {
        forest.tree_start.push_back((int)forest.steps.size());
        forest.steps.insert(forest.steps.end(), kv.second.begin(), kv.second.end());
      }
      forest.tree_start.push_back((int)forest.steps.size());
      p->seq1.push_back(forest);
      build_seq(rest, p->seq1);
    } else {
      build_seq(p->order1, p->seq1);
    }
  }
  build_seq(p->order2, p->seq2);
  // Liveness over the launch sequence [phase 1][phase 2 of one group]; the
  // steps of one launch share a position; anything a phase-2 step reads
  // stays live to the end (every group reads it again).
  std::vector<int> pos(d->n_steps);
  int t = 0;
  for (const Launch& L : p->seq1) {
    if (L.grouped)
      for (int i : L.steps) pos[i] = t;
    else
      pos[L.step] = t;
    ++t;
  }
  const int phase2_start = t;
  for (const Launch& L : p->seq2) {
    if (L.grouped)
      for (int i : L.steps) pos[i] = t;
    else
      pos[L.step] = t;
    ++t;
  }
  const int end = t;
  struct Live { int s, start, stop; size_t bytes; };
  std::vector<Live> lives;
  for (int s = 0; s < d->n_slots; ++s) {
    Slot& sl = p->slots[s];
    if (sl.leaf) continue;
    if (sl.producer < 0) continue;
    int start = pos[sl.producer];
    int stop = sl.last_use >= 0 ? pos[sl.last_use] : start;
    if (s == p->out_slot) stop = end;
    bool read_in_2 = false;
    for (int i : p->order2) {
      const PStep& q = p->steps[i];
      if (q.a == s || q.b == s || std::find(q.ops.begin(), q.ops.end(), s) != q.ops.end()) read_in_2 = true;
    }
    if (read_in_2 && start < phase2_start) stop = end;
    size_t bytes = (size_t)sl.elems * esize_of(sl.dtype) * (size_t)sl.copies();
    bytes = (bytes + 255) / 256 * 256;
    lives.push_back({s, start, stop, bytes});
  }
  // First-fit over time-ordered intervals.
  std::sort(lives.begin(), lives.end(), [](const Live& x, const Live& y) { return x.start < y.start; });
  std::vector<Live> active;
  size_t top = 0;
  for (auto& L : lives) {
    active.erase(std::remove_if(active.begin(), active.end(), [&](const Live& a) { return a.stop < L.start; }),
                 active.end());
    std::vector<std::pair<size_t, size_t>> used;
    for (auto& a : active) used.push_back({(size_t)p->slots[a.s].arena_off, a.bytes});
    std::sort(used.begin(), used.end());
    size_t off = 0;
    for (auto& u : used) {
      if (off + L.bytes <= u.first) break;
      off = std::max(off, u.first + u.second);
    }
    p->slots[L.s].arena_off = (int64_t)off;
    p->slots[L.s].region = L.bytes;
    top = std::max(top, off + L.bytes);
    active.push_back(L);
  }
  p->arena_bytes = std::max<size_t>(top, 256);
  p->out_elems = p->slots[p->out_slot].elems;
  if (hipMalloc(&p->arena, p->arena_bytes) != hipSuccess ||
      (p->scratch_bytes && hipMalloc(&p->scratch, p->scratch_bytes) != hipSuccess) ||
      hipMalloc((void**)&p->out_dev, sizeof(double) * 2 * p->out_elems * p->n_out) != hipSuccess) {
    (void)hipGetLastError();
    free_plan(c, p);
    return fail(c, PQ_ERR_OOM, "plan memory allocation failed");
  }
  // Grouped descriptors (pointers of slice group 0 + per-bit leaf offsets).
  if (d->n_sliced - d->group_bits > PQ_MAX_HIGH_BITS) {
    bool any = false;
    for (auto* seq : {&p->seq1, &p->seq2})
      for (const Launch& L : *seq) any = any || L.grouped;
    if (any) {
      free_plan(c, p);
      return fail(c, PQ_ERR_UNSUPPORTED, "too many slice groups for grouped launches");
    }
  }
  std::vector<MultiDesc> hd;
  for (auto* seq : {&p->seq1, &p->seq2}) {
    for (Launch& L : *seq) {
      if (!L.grouped || L.forest) continue;
      L.desc_off = (int)hd.size();
      uint64_t maxw = 1;
      unsigned maxb = 1;
      for (int i : L.steps) {
        const PStep& st = p->steps[i];
        MultiDesc m;
        memset(&m, 0, sizeof(m));
        m.c128 = p->slots[st.a].dtype == PQ_C128 ? 1 : 0;
        m.A = slot_ptr(c, p, st.a, 0);
        m.B = slot_ptr(c, p, st.b, 0);
        m.C = slot_ptr(c, p, st.c, 0);
        if (st.kind == 1) {
          // conversion: conv_elems consecutive elements (every copy), one entry
          const int64_t n = conv_elems(st);
          m.conv = 1;
          if (!fill_group(m.M, {{n, 1, 1}}) || !fill_group(m.N, {}) || !fill_group(m.K, {})) {
            free_plan(c, p);
            return fail(c, PQ_ERR_UNSUPPORTED, "grouped conversion too large");
          }
          m.L = single_batch();
          maxw = std::max<uint64_t>(maxw, (uint64_t)n);
        } else {
          if (!fill_group(m.M, st.spec.M) || !fill_group(m.N, st.spec.N) || !fill_group(m.K, st.spec.K)) {
            free_plan(c, p);
            return fail(c, PQ_ERR_UNSUPPORTED, "too many modes in a grouped step");
          }
          m.L = st.spec.L;
          maxw = std::max<uint64_t>(maxw, st.spec.Mt * st.spec.Nt);
          maxb = std::max<unsigned>(maxb, st.spec.L.total);
        }
        m.nhigh = st.phase2 ? d->n_sliced - d->group_bits : 0;
        for (int j = 0; j < m.nhigh; ++j) {
          const int bit = d->group_bits + j;
          m.hA[j] = p->slots[st.a].slice_stride[bit];
          m.hB[j] = p->slots[st.b].slice_stride[bit];
          // a conversion's output keeps the same sliced bonds as its input
          m.hC[j] = st.kind == 1 ? p->slots[st.c].slice_stride[bit] : 0;
        }
        hd.push_back(m);
      }
      // about one output per lane: a grouped problem's outputs are
      // latency-bound chains of K loads (capped at 64 workgroups, a problem
      // of 2^17 outputs took 8 per lane: 21-23 us against 10.5-11 us)
      L.blocks_x = (unsigned)std::min<uint64_t>((maxw + 255) / 256, 512);
      L.batch = maxb;
    }
  }
  if (!hd.empty()) {
    if (hipMalloc((void**)&p->descs, hd.size() * sizeof(MultiDesc)) != hipSuccess ||
        hipMemcpy(p->descs, hd.data(), hd.size() * sizeof(MultiDesc), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipGetLastError();
      free_plan(c, p);
      return fail(c, PQ_ERR_OOM, "descriptor upload failed");
    }
  }
  for (const Launch& L : p->seq1) {
    if (!L.forest) continue;
    // Forest tables: a record per step, an item per output element (offsets
    // decoded here from the step's modes), the K offsets, and per tree its
    // levels' item ranges.
    auto decode = [](const std::vector<Mode>& ms, int64_t i, int64_t& o0, int64_t& o1) {
      o0 = o1 = 0;
      for (const Mode& m : ms) {
        const int64_t r = i % m.ext;
        i /= m.ext;
        o0 += r * m.s0;
        o1 += r * m.s1;
      }
    };
    std::vector<ForestStep> fsv;
    std::vector<int> fstep;   // the plan step of each record
    std::vector<int4> items;
    std::vector<int2> ktab;
    std::vector<ForestCopy> copies;
    const int nt = (int)L.tree_start.size() - 1;
    std::vector<int32_t> lv(nt, 0), hdrs(nt);
    std::vector<size_t> tbytes(nt);
    size_t f_lds = 0;
    for (int b = 0; b < nt; ++b) {
      lv[b] = (int32_t)lv.size();
      std::map<int, std::vector<int>> lvl;   // level -> steps
      for (int q = L.tree_start[b]; q < L.tree_start[b + 1]; ++q)
        lvl[p->steps[L.steps[q]].level].push_back(L.steps[q]);
      lv.push_back((int32_t)lvl.size());
      const size_t hdr = lv.size();
      hdrs[b] = (int32_t)hdr;
      const int32_t fs0 = (int32_t)fsv.size(), k0 = (int32_t)ktab.size();
      // first step record, records, first K entry, K entries, first leaf copy, leaf copies
      lv.insert(lv.end(), 6, 0);
      lv.push_back((int32_t)items.size());
      for (auto& kv : lvl) {
        for (int i : kv.second) {
          const PStep& st = p->steps[i];
          ForestStep f;
          memset(&f, 0, sizeof(f));
          fstep.push_back(i);
          f.A = slot_ptr(c, p, st.a, 0);
          f.B = slot_ptr(c, p, st.b, 0);
          f.C = slot_ptr(c, p, st.c, 0);
          f.c128 = p->slots[st.a].dtype == PQ_C128 ? 1 : 0;
          const int32_t si = (int32_t)fsv.size();
          if (st.kind == 1) {
            f.conv = 1;
            for (int64_t o = 0; o < conv_elems(st); ++o) items.push_back(make_int4(si, (int32_t)o, 0, (int32_t)o));
          } else {
            f.kt = (int32_t)ktab.size();
            f.nk = (int32_t)st.spec.Kt;
            for (int64_t k = 0; k < (int64_t)st.spec.Kt; ++k) {
              int64_t ka, kb;
              decode(st.spec.K, k, ka, kb);
              ktab.push_back(make_int2((int32_t)ka, (int32_t)kb));
            }
            const int64_t mt = (int64_t)st.spec.Mt;
            for (int64_t o = 0; o < (int64_t)(st.spec.Mt * st.spec.Nt); ++o) {
              int64_t pa, pc1, pb, pc2;
              decode(st.spec.M, o % mt, pa, pc1);
              decode(st.spec.N, o / mt, pb, pc2);
              items.push_back(make_int4(si, (int32_t)pa, (int32_t)pb, (int32_t)(pc1 + pc2)));
            }
          }
          fsv.push_back(f);
        }
        lv.push_back((int32_t)items.size());
      }
      const int32_t nfs = (int32_t)fsv.size() - fs0, nkt = (int32_t)ktab.size() - k0;
      lv[hdr] = fs0;
      lv[hdr + 1] = nfs;
      lv[hdr + 2] = k0;
      lv[hdr + 3] = nkt;
      // LDS of the tree's records and K offsets (k_forest: records as 32-bit
      // words, then the K entries 8-byte aligned)
      const size_t words = (size_t)nfs * (sizeof(ForestStep) / 4);
      tbytes[b] = 4 * (words + (words & 1)) + sizeof(int2) * (size_t)nkt;
      f_lds = std::max(f_lds, tbytes[b]);
    }
    if (items.size() >= ((size_t)1 << 31)) {
      free_plan(c, p);
      return fail(c, PQ_ERR_UNSUPPORTED, "forest too large");
    }
    // With the tables in LDS, each tree also keeps in LDS (its data area,
    // after the tables) the results that it reads itself and the small
    // leaves that it reads, while they fit: a result lives from its level
    // to its reader's level, a leaf from the start to its reader's level;
    // first fit, 16-byte aligned.
    if (f_lds <= PQ_FOREST_LDS_BYTES) {
      // readers of each slot over the whole plan: only a result with one
      // reader, in its own tree, and not the plan's output moves to LDS
      std::vector<int> nread(p->slots.size(), 0);
      for (const PStep& st : p->steps) {
        ++nread[st.a];
        if (st.kind == 0) ++nread[st.b];
        for (int os : st.ops) ++nread[os];
      }
      nread[p->out_slot] += 1;
      for (int b = 0; b < nt; ++b) {
        const int32_t fs0 = lv[hdrs[b]], nfs = lv[hdrs[b] + 1];
        const size_t base = (tbytes[b] + 15) & ~(size_t)15;
        if (base >= PQ_FOREST_LDS_BYTES) continue;
        const int64_t budget = (int64_t)(PQ_FOREST_LDS_BYTES - base);
        // the record and operand (0: A, 1: B) reading each result of the tree
        std::map<int, std::pair<int, int>> reader;
        for (int r = fs0; r < fs0 + nfs; ++r) {
          const PStep& st = p->steps[fstep[r]];
          for (int o = 0; o < (st.kind == 1 ? 1 : 2); ++o) {
            const int x = o ? st.b : st.a;
            if (!p->slots[x].leaf && nread[x] == 1) reader[x] = {r, o};
          }
        }
        struct Live {
          int64_t off, end;
          int last;   // level of the reader
        };
        std::vector<Live> live;
        auto place = [&](int64_t bytes, int now, int last) -> int64_t {
          live.erase(std::remove_if(live.begin(), live.end(), [&](const Live& a) { return a.last < now; }),
                     live.end());
          std::sort(live.begin(), live.end(), [](const Live& x, const Live& y) { return x.off < y.off; });
          int64_t at = 0;
          for (const Live& a : live) {
            if (a.off - at >= bytes) break;
            at = std::max(at, a.end);
          }
          if (at + bytes > budget) return -1;
          live.push_back({at, at + ((bytes + 15) & ~(int64_t)15), last});
          return at;
        };
        const int32_t c0 = (int32_t)copies.size();
        int64_t pieces = 0, high = 0;
        const int64_t max_pieces = 256 * PQ_FOREST_COPY_PER_LANE;
        for (int r = fs0; r < fs0 + nfs; ++r) {
          const PStep& st = p->steps[fstep[r]];
          for (int o = 0; o < (st.kind == 1 ? 1 : 2); ++o) {
            const Slot& sl = p->slots[o ? st.b : st.a];
            if (!sl.leaf) continue;
            bool fixed = true;   // read at one offset (keeps no sliced bond)
            for (int64_t v : sl.slice_stride) fixed = fixed && v == 0;
            // the elements the leaf's view reaches
            const int64_t bytes = (sl.leaf_span + 1) * (sl.dtype == PQ_C128 ? 16 : 8);
            if (!fixed || bytes > 4096 || pieces + bytes / 8 > max_pieces) continue;
            const int64_t off = place(bytes, INT32_MIN, st.level);
            if (off < 0) continue;
            const char* src = (const char*)(o ? fsv[r].B : fsv[r].A);
            for (int64_t w = 0; w < bytes / 8; ++w) copies.push_back({src + 8 * w, (int32_t)(off + 8 * w), 0});
            pieces += bytes / 8;
            high = std::max(high, off + bytes);
            fsv[r].lds |= 1 << o;
            if (o) fsv[r].B = (const void*)(uintptr_t)off;
            else fsv[r].A = (const void*)(uintptr_t)off;
          }
        }
        for (int r = fs0; r < fs0 + nfs; ++r) {
          const PStep& st = p->steps[fstep[r]];
          auto rd = reader.find(st.c);
          if (rd == reader.end()) continue;   // the tree's root: read by a later launch
          const Slot& sl = p->slots[st.c];
          // every copy of a batched result (a forest step covers them all
          // when the batch modes fold into its free modes)
          const int64_t bytes = sl.elems * sl.copies() * (sl.dtype == PQ_C128 ? 16 : 8);
          const int64_t off = place(bytes, st.level, p->steps[fstep[rd->second.first]].level);
          if (off < 0) continue;
          high = std::max(high, off + bytes);
          fsv[r].lds |= 4;
          fsv[r].C = (void*)(uintptr_t)off;
          ForestStep& u = fsv[rd->second.first];
          u.lds |= 1 << rd->second.second;
          if (rd->second.second) u.B = (const void*)(uintptr_t)off;
          else u.A = (const void*)(uintptr_t)off;
        }
        lv[hdrs[b] + 4] = c0;
        lv[hdrs[b] + 5] = (int32_t)copies.size() - c0;
        if (high > 0) f_lds = std::max(f_lds, base + (size_t)high);
      }
    }
    if (ktab.empty()) ktab.push_back(make_int2(0, 0));
    if (copies.empty()) copies.push_back({nullptr, 0, 0});
    const size_t b_steps = fsv.size() * sizeof(ForestStep), b_items = items.size() * sizeof(int4),
                 b_k = ktab.size() * sizeof(int2), b_cp = copies.size() * sizeof(ForestCopy),
                 b_lv = lv.size() * sizeof(int32_t);
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t o_items = up(b_steps), o_k = o_items + up(b_items), o_cp = o_k + up(b_k), o_lv = o_cp + up(b_cp),
                 total = o_lv + up(b_lv);
    std::vector<unsigned char> blob(total, 0);
    memcpy(blob.data(), fsv.data(), b_steps);
    memcpy(blob.data() + o_items, items.data(), b_items);
    memcpy(blob.data() + o_k, ktab.data(), b_k);
    memcpy(blob.data() + o_cp, copies.data(), b_cp);
    memcpy(blob.data() + o_lv, lv.data(), b_lv);
    if (hipMalloc(&p->forest_buf, total) != hipSuccess ||
        hipMemcpy(p->forest_buf, blob.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipGetLastError();
      free_plan(c, p);
      return fail(c, PQ_ERR_OOM, "forest table upload failed");
    }
    unsigned char* fb = (unsigned char*)p->forest_buf;
    p->f_steps = (const ForestStep*)fb;
    p->f_items = (const int4*)(fb + o_items);
    p->f_ktab = (const int2*)(fb + o_k);
    p->f_copies = (const ForestCopy*)(fb + o_cp);
    p->f_lv = (const int32_t*)(fb + o_lv);
    // the tables in LDS when the largest tree's fit, else read from memory
    p->f_lds = f_lds <= PQ_FOREST_LDS_BYTES ? f_lds : 0;
  }
  if (!chain_tab.empty()) {
    if (hipMalloc((void**)&p->chain_tab, chain_tab.size() * sizeof(int32_t)) != hipSuccess ||
        hipMemcpy(p->chain_tab, chain_tab.data(), chain_tab.size() * sizeof(int32_t), hipMemcpyHostToDevice) !=
            hipSuccess) {
      (void)hipGetLastError();
      free_plan(c, p);
      return fail(c, PQ_ERR_OOM, "fused-chain table upload failed");
    }
  }
  p->launches = (int64_t)p->seq1.size() + p->n_groups * ((int64_t)p->seq2.size() + 1);
  // Streams: independent launches of a phase run concurrently
  // (pq_set_plan_streams; 1 = every launch on the context's stream).
  p->nstreams = std::max(1, std::min(c->plan_streams, PQ_PLAN_STREAMS));
  if (p->nstreams > 1 && ensure_aux_streams(c) != PQ_OK) p->nstreams = 1;
  if (p->nstreams > 1) {
    const std::vector<Launch>* seqs[2] = {&p->seq1, &p->seq2};
    bool ok = true;
    for (int ph = 0; ph < 2 && ok; ++ph) {
      make_schedule(p, *seqs[ph], p->nstreams, p->sch[ph]);
      p->lev[ph].assign(seqs[ph]->size(), nullptr);
      for (size_t t = 0; t < seqs[ph]->size() && ok; ++t)
        if (p->sch[ph][t].record) ok = hipEventCreateWithFlags(&p->lev[ph][t], hipEventDisableTiming) == hipSuccess;
      ok = ok && hipEventCreateWithFlags(&p->fork_ev[ph], hipEventDisableTiming) == hipSuccess;
      for (int q = 1; q < p->nstreams && ok; ++q)
        ok = hipEventCreateWithFlags(&p->join_ev[ph][q], hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
      (void)hipGetLastError();
      free_plan(c, p);
      return fail(c, PQ_ERR_HIP, "event creation failed");
    }
  }
  *out = p;
  return PQ_OK;
}

int pq_plan_execute(pq_ctx* c, pq_plan* p, int use_graph) {
  if (!c || !p) return PQ_ERR_ARG;
  if (!use_graph) return run_program(c, p, false);
  if (!p->exec) {
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hipfail(c, e, "begin capture");
    c->capturing = true;
    int r = run_program(c, p, false);
    c->capturing = false;
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(c->stream, &g);
    if (r) {
      if (g) hipGraphDestroy(g);
      return r;
    }
    if (e != hipSuccess) return hipfail(c, e, "end capture");
    p->graph = g;
    e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) return hipfail(c, e, "graph instantiate");
  }
  hipError_t e = hipGraphLaunch(p->exec, c->stream);
  if (e != hipSuccess) return hipfail(c, e, "graph launch");
  return PQ_OK;
}

int pq_plan_result(pq_ctx* c, pq_plan* p, double* host) {
  if (!c || !p || !host) return PQ_ERR_ARG;
  hipError_t e = hipMemcpyAsync(host, p->out_dev, sizeof(double) * 2 * p->out_elems * p->n_out,
                                hipMemcpyDeviceToHost,
                                c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hipfail(c, e, "plan result");
  return PQ_OK;
}

int pq_plan_stats(const pq_plan* p, int64_t* arena_bytes, int64_t* launches, double* alg_flops, double* alg_bytes,
                  int64_t* out_elems) {
  if (!p) return PQ_ERR_ARG;
  if (arena_bytes) *arena_bytes = (int64_t)p->arena_bytes;
  if (launches) *launches = p->launches;
  if (alg_flops) *alg_flops = p->alg_flops;
  if (alg_bytes) *alg_bytes = p->alg_bytes;
  if (out_elems) *out_elems = p->out_elems * p->n_out;
  return PQ_OK;
}

int pq_plan_step_info(const pq_plan* p, int step, int* kclass, int* tile, int* dep, double* flops, double* bytes) {
  if (!p || step < 0 || step >= (int)p->steps.size()) return PQ_ERR_ARG;
  const PStep& s = p->steps[step];
  int kc = s.kind == 1 ? PQ_KC_COPY : s.spec.kc;
  for (auto* seq : {&p->seq1, &p->seq2})
    for (const Launch& L : *seq)
      if (L.grouped && std::find(L.steps.begin(), L.steps.end(), step) != L.steps.end()) kc = PQ_KC_BATCHED;
  if (kclass) *kclass = kc;
  if (tile) *tile = s.spec.tile;
  if (dep) *dep = s.dep ? 1 : 0;
  if (flops) *flops = s.flops;
  if (bytes) *bytes = s.bytes;
  return PQ_OK;
}

int pq_plan_profile(pq_ctx* c, pq_plan* p, double* step_ms) {
  if (!c || !p || !step_ms) return PQ_ERR_ARG;
  const size_t n = p->steps.size();
  if (p->ev.empty()) {
    p->ev.resize(2 * n);
    for (auto& e : p->ev) hipEventCreate(&e);
  }
  p->ev_rec.assign(n, 0);
  for (size_t i = 0; i < n; ++i) step_ms[i] = 0;
  hipStreamSynchronize(c->stream);
  // Phase 1 and one slice group with events around every launch; phase-2
  // times are scaled by the number of groups.
  int64_t ng = p->n_groups;
  p->n_groups = 1;
  int r = run_program(c, p, true);
  p->n_groups = ng;
  if (r) return r;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hipfail(c, e, "profile sync");
  // Grouped launches are timed as a whole on their first member step.
  for (size_t i = 0; i < n; ++i) {
    if (!p->ev_rec[i]) continue;
    float ms = 0;
    e = hipEventElapsedTime(&ms, p->ev[2 * i], p->ev[2 * i + 1]);
    if (e != hipSuccess) return hipfail(c, e, "profile event");
    // phase-2 launches (those of seq2, grouped or not) run once per slice
    // group; phase-1 launches once, batched over a group's slices or not
    step_ms[i] = p->steps[i].phase2 ? ms * (double)ng : ms;
  }
  return PQ_OK;
}

int pq_set_plan_streams(pq_ctx* c, int n) {
  if (!c) return PQ_ERR_ARG;
  if (n < 1 || n > PQ_PLAN_STREAMS) return fail(c, PQ_ERR_ARG, "plan streams out of range");
  // Concurrent plan runs are switched off: a test run of plans on four
  // streams failed mixed-precision plans and then ended in a host
  // segmentation fault inside pq_plan_execute, with the cause not found.
  if (n > 1) return fail(c, PQ_ERR_UNSUPPORTED, "concurrent plan streams are switched off in this build");
  c->plan_streams = n;
  return PQ_OK;
}

int pq_plan_cost(const pq_plan* p, pq_cost* total, double* step_macs, double* step_bytes) {
  if (!p) return PQ_ERR_ARG;
  double macs = 0, bytes = 0;
  for (size_t i = 0; i < p->steps.size(); ++i) {
    const PStep& st = p->steps[i];
    // algorithmic flops are 8 real flops per complex multiply-add; a
    // conversion moves data only
    const double m = st.kind == 1 ? 0.0 : st.flops / 8.0;
    if (step_macs) step_macs[i] = m;
    if (step_bytes) step_bytes[i] = st.bytes;
    macs += m;
    bytes += st.bytes;
  }
  if (total) {
    total->macs = macs;
    total->bytes = bytes;
    total->steps = (int64_t)p->steps.size();
    total->launches = p->launches;
  }
  return PQ_OK;
}

int pq_plan_set_bitstrings(pq_ctx* c, pq_plan* p, int n, const uint64_t* bits) {
  if (!c || !p) return PQ_ERR_ARG;
  if (p->outq.empty()) return fail(c, PQ_ERR_ARG, "the plan has no bitstring batch");
  const int64_t nb = (int64_t)1 << p->n_out_bits;
  if (n < 1 || (int64_t)n > nb || !bits) return fail(c, PQ_ERR_ARG, "bitstring count out of range");
  const size_t es = esize_of(c->dtype);
  std::vector<unsigned char> col((size_t)(2 * nb) * es);
  for (const auto& q : p->outq) {
    std::fill(col.begin(), col.end(), 0);
    for (int64_t b = 0; b < nb; ++b) {
      const uint64_t w = bits[b < n ? b : 0];
      const int64_t e = 2 * b + (int64_t)((w >> q.bit) & 1);   // column b: [1, 0] or [0, 1]
      if (c->dtype == PQ_C128)
        reinterpret_cast<double*>(col.data())[2 * e] = 1.0;
      else
        reinterpret_cast<float*>(col.data())[2 * e] = 1.0f;
    }
    const Slot& sl = p->slots[q.slot];
    void* dst = (unsigned char*)sl.leaf_buf->ptr + sl.leaf_off * (int64_t)es;
    hipError_t e = hipMemcpyAsync(dst, col.data(), col.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hipfail(c, e, "bitstring upload");
  }
  return PQ_OK;
}

int pq_plan_run(pq_ctx* c, pq_plan* p, int n, const uint64_t* bits, double* out) {
  if (!c || !p) return PQ_ERR_ARG;
  if (n > 0) {
    const int r = pq_plan_set_bitstrings(c, p, n, bits);
    if (r) return r;
  }
  int r = pq_plan_execute(c, p, 1);
  if (r) return r;
  return out ? pq_plan_result(c, p, out) : PQ_OK;
}

int pq_plan_destroy(pq_ctx* c, pq_plan* p) {
  if (!c || !p) return PQ_ERR_ARG;
  hipStreamSynchronize(c->stream);
  for (hipStream_t a : c->aux)
    if (a) hipStreamSynchronize(a);
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  for (auto& e : p->ev) hipEventDestroy(e);
  free_plan(c, p);
  return PQ_OK;
}

}  // extern "C"
