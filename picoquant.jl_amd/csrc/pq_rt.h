// Runtime internals shared by the tensor API (pq_runtime.hip) and the plan
// executor (pq_plan.hip): context, refcounted device buffers, strided
// tensors, and the lowering of one pairwise step to a kernel launch.
#pragma once
#include "pq_internal.h"
#include "../../include/pq_hip.h"

#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#define PQ_PLAN_STREAMS 4   // streams a planned run spreads independent launches over

namespace pqrt {

struct Buf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int refs = 0;
};

struct Tensor {
  Buf* buf = nullptr;
  int64_t off = 0;  // element offset into buf
  int rank = 0;
  int64_t dims[PQ_MAX_RANK];
  int64_t strides[PQ_MAX_RANK];
};

struct ProfEntry {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  int64_t launches = 0;
  double bytes = 0, flops = 0;
};

// A mode before lowering: extent and strides in two tensors.
struct Mode {
  int64_t ext;
  int64_t s0, s1;
};

// One pairwise step ready to launch.
//   M: (ext, sA, sC)   N: (ext, sB, sC)   K: (ext, sA, sB)   (merged)
struct StepSpec {
  std::vector<Mode> M, N, K;
  Batch L;           // L.total >= 1
  int kc = -1;       // PQ_KC_* after resolve_step
  int tile = -1;     // GEMM tile after resolve_step
  bool stream_ab = true;  // stream kernel: X = A (true) or X = B
  bool gemm_swap = false; // GEMM runs as C^T = B^T A^T (C's unit stride on the M side)
  int arith = -1;         // complex64 GEMM: 0 FP32, 1 FP16 hi/lo; -1 = the context's
  int dtype = -1;         // PQ_C128 / PQ_C64; -1 = the context's
  int splitk = 1;         // GEMM: K cut into this many partial products in scratch, then summed
  uint64_t Mt = 1, Nt = 1, Kt = 1;
};

int64_t numel(const Tensor& t);
bool is_contiguous(const Tensor& t);
void set_contiguous(Tensor& t);
std::vector<Mode> merge_modes(const std::vector<Mode>& in);
bool fill_group(Group& g, const std::vector<Mode>& modes);
Batch single_batch();

}  // namespace pqrt

struct pq_ctx {
  int device = 0;
  int dtype = PQ_C128;
  size_t esize = 16;
  hipStream_t stream = nullptr;
  std::string err;
  std::unordered_map<int64_t, pqrt::Tensor> tensors;
  int64_t next_id = 1;
  std::multimap<size_t, void*> free_blocks;
  size_t cached = 0;
  bool prof = false;
  bool capturing = false;
  pqrt::ProfEntry pe[PQ_KC_COUNT];
  int override_kc = -1;
  int gemm_tile = -1;
  int gemm_orient = -1;
  int chain_pairs = -1;   // fused chains: -1 automatic 16-byte pair moves, 0 never
  int arith = 0;
  void* scratch = nullptr;  // split-K partials for eager calls
  size_t scratch_bytes = 0;
  void* blas = nullptr;     // rocBLAS handle on this stream (rocSOLVER SVD), created on first use
  void* noalloc_scratch = nullptr;  // this device's scratch array of pq_contract_noalloc
  // Extra streams for concurrent branches of a planned run (default priority,
  // created on first use); stream 0 of a plan's schedule is `stream`.
  hipStream_t aux[PQ_PLAN_STREAMS - 1] = {};
  int plan_streams = 1;   // read when a plan is created (above 1 is switched off: pq_set_plan_streams)
};

namespace pqrt {

int fail(pq_ctx* c, int code, const std::string& msg);
int hipfail(pq_ctx* c, hipError_t e, const char* where);
void* dev_alloc(pq_ctx* c, size_t bytes, size_t* got);
void dev_release(pq_ctx* c, void* p, size_t bytes);
Buf* new_buf(pq_ctx* c, size_t bytes);
void unref(pq_ctx* c, Buf* b);
Tensor* get(pq_ctx* c, pq_tensor h);
pq_tensor add(pq_ctx* c, const Tensor& t);
int materialize_into(pq_ctx* c, const Tensor& src, Tensor& dst);
// Creates the context's extra streams if they do not exist yet.
int ensure_aux_streams(pq_ctx* c);
// Destroys the context's rocBLAS handle, if one was created.
void release_blas(pq_ctx* c);

// Choose kernel class and tile for a step (honouring the context overrides
// unless the step already names them).  Returns PQ_OK or PQ_ERR_UNSUPPORTED.
int resolve_step(pq_ctx* c, StepSpec& s, size_t esize);
// Scratch bytes the step needs (split-K partials), 0 for the others.
size_t step_scratch_bytes(const StepSpec& s, size_t esize);
// Checks the ncon labels of a pairwise contraction (pq_contract's rules)
// and builds it: Ct gets the result's dims and contiguous strides (no
// buffer), kt the summed extent, and -- when the result is not empty and kt
// > 0 -- s the resolved step writing Ct's layout.
int prepare_contract(pq_ctx* c, const Tensor& A, int ra, const int32_t* ia, const Tensor& B, int rb,
                     const int32_t* ib, Tensor& Ct, StepSpec& s, int64_t& kt);
// Address of pq_contract_noalloc's device scratch on `device` (looked up
// once per device, with that device current); nullptr if unavailable.
void* noalloc_scratch(int device);
// Launch one resolved step on the context's stream.
int launch_step(pq_ctx* c, const void* A, const void* B, void* Out, const StepSpec& s, void* scratch);

struct ProfScope {
  pq_ctx* c;
  int kc;
  hipEvent_t e1 = nullptr;
  bool on;
  ProfScope(pq_ctx* c_, int kc_, double bytes, double flops);
  ~ProfScope();
};

}  // namespace pqrt
