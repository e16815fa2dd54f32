/*
 * libpq_hip -- MI355X (gfx950) backend for PicoQuant's pairwise tensor
 * contraction path.  Plain C ABI: opaque context, integer tensor handles,
 * plain pointers and sizes.  Every function returns 0 (PQ_OK) on success or a
 * PQ_ERR_* code; pq_last_error() returns the message of the last failure on
 * that context.
 *
 * Conventions (src/backends/interactive.jl, src/layer1.jl):
 *   - tensors are column-major; dims are int64 and ordered as in the Node
 *   - element type is fixed per context (InteractiveBackend{T}'s T):
 *     PQ_C128 = ComplexF64, PQ_C64 = ComplexF32
 *   - host buffers are interleaved (re, im), contiguous, column-major
 *   - one context owns one HIP stream; calls on a context are not re-entrant
 *   - device work is stream-ordered; pq_tensor_download synchronises
 */
#ifndef PQ_HIP_H
#define PQ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PQ_OK 0
#define PQ_ERR_ARG 1          /* invalid argument (shapes, labels, handles) */
#define PQ_ERR_HIP 2          /* HIP runtime failure */
#define PQ_ERR_NOT_FOUND 3    /* unknown tensor handle (KeyError in the reference) */
#define PQ_ERR_OOM 4          /* device allocation failed */
#define PQ_ERR_UNSUPPORTED 5  /* shape outside this build's limits */

#define PQ_C128 0
#define PQ_C64 1

#define PQ_MAX_RANK 64

typedef struct pq_ctx pq_ctx;
typedef int64_t pq_tensor; /* > 0 for a live tensor */

/* Context = one backend instance: InteractiveBackend{T}() (src/backends/interactive.jl:10-23). */
int pq_ctx_create(int device, int dtype, pq_ctx** out);
int pq_ctx_destroy(pq_ctx* ctx);
const char* pq_last_error(const pq_ctx* ctx);
int pq_ctx_dtype(const pq_ctx* ctx);
int pq_synchronize(pq_ctx* ctx);

/* save_tensor_data: store a converted copy (src/backends/interactive.jl:32-36).
 * host_dtype (PQ_C128 / PQ_C64) is converted to the context's element type. */
int pq_tensor_upload(pq_ctx* ctx, int rank, const int64_t* dims, const void* host,
                     int host_dtype, pq_tensor* out);

/* Overwrite a contiguous tensor's elements in place with host data of the
 * same shape (host_dtype converted to the context's type); stream-ordered
 * after earlier work on the context, synchronising.  Compiled plans and
 * captured graphs that read the tensor see the new values on their next
 * run: this is how a plan's output bitstrings change without recompiling.
 * The reference's save_tensor_data on an existing label
 * (src/backends/interactive.jl:32-36) stores a new array instead. */
int pq_tensor_write(pq_ctx* ctx, pq_tensor t, const void* host, int host_dtype);

/* Shape query with no data transfer (partition_network_on_virtual_bonds reads
 * only the shape, src/layer2/slicing.jl:46-48). dims must hold PQ_MAX_RANK. */
int pq_tensor_shape(pq_ctx* ctx, pq_tensor t, int* rank, int64_t* dims);

/* load_tensor_data (src/backends/interactive.jl:44-49): copy out, column-major
 * contiguous, in the context's element type.  Synchronising. */
int pq_tensor_download(pq_ctx* ctx, pq_tensor t, void* host);

/* delete_tensor! (src/backends/interactive.jl:159-161). */
int pq_tensor_free(pq_ctx* ctx, pq_tensor t);

/* contract_tensors (src/backends/interactive.jl:60-75 -> src/layer1.jl:85-92):
 * ncon labels, positive = summed (each appears once in A and once in B),
 * negative = open; C's index order is symdiff(ia, ib): A's open indices in A's
 * order, then B's open indices in B's order.  A and B are not consumed; the
 * caller frees them (the reference deletes both after the contraction). */
int pq_contract(pq_ctx* ctx, pq_tensor A, int rankA, const int32_t* ia,
                pq_tensor B, int rankB, const int32_t* ib, pq_tensor* C);

/* pq_contract writing into an existing result tensor, with no device
 * allocation during the call.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  C must
 * be a contiguous tensor of the result's dims (as pq_contract would
 * create it) sharing no buffer with A or B; it is overwritten.  Split-K and
 * split-reduction partials go to a device array of PQ_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends (looked up by pq_ctx_create); a
 * contraction that needs more scratch returns PQ_ERR_UNSUPPORTED.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (two
 * streams, two contexts, two threads) race.  With profiling enabled
 * (pq_profile_enable) the call creates timing events. */
#define PQ_NOALLOC_SCRATCH_BYTES (32u << 20)
int pq_contract_noalloc(pq_ctx* ctx, pq_tensor A, int rankA, const int32_t* ia,
                        pq_tensor B, int rankB, const int32_t* ib, pq_tensor C);

/* permute_tensor (src/backends/interactive.jl:111-115 -> src/layer1.jl:111-114):
 * perm is 1-based as in permutedims.  Zero-copy: the result is a strided view
 * that consumers read through; download/reshape materialise it. */
int pq_permute(pq_ctx* ctx, pq_tensor t, const int32_t* perm, pq_tensor* out);

/* reshape_tensor (src/backends/interactive.jl:97-102): new dims, same element
 * count, column-major; in place on the handle (metadata only when contiguous). */
int pq_reshape(pq_ctx* ctx, pq_tensor t, int rank, const int64_t* dims);

/* view_tensor! -> tensor_view (src/backends/interactive.jl:169-172,
 * src/layer1.jl:191-194): restrict 1-based axis to lo:hi (inclusive); the
 * axis is kept with extent hi-lo+1.  Zero-copy. */
int pq_view(pq_ctx* ctx, pq_tensor t, int axis, int64_t lo, int64_t hi, pq_tensor* out);

/* Materialise a strided tensor into a new contiguous one (used by save_output
 * and by callers that need dense storage). */
int pq_materialize(pq_ctx* ctx, pq_tensor t, pq_tensor* out);

/* decompose_tensor! -> decompose_tensor (src/backends/interactive.jl:130-152,
 * src/layer1.jl:146-190): SVD split of t into B (dims of the 1-based `left`
 * axes, then chi) and C (chi, then the `right` axes) with B C = t permuted to
 * (left, right) up to the discarded singular values; S is split as
 * sqrt(S) into both.  chi counts the singular values whose ratio to the
 * spectrum's 2-norm exceeds max(threshold, sqrt(eps)) of the element type,
 * capped by max_rank when max_rank > 0.  t is not consumed (the caller
 * frees it, as the backend verb does).  On the GPU: rocSOLVER gesdd;
 * synchronises the context's stream (chi is needed on the host).  A tensor
 * with no singular value above the threshold gives PQ_ERR_UNSUPPORTED. */
int pq_decompose(pq_ctx* ctx, pq_tensor t, int nleft, const int32_t* left, int nright, const int32_t* right,
                 double threshold, int max_rank, pq_tensor* B, pq_tensor* C, int* chi);

/* Device pointer and element offset of a tensor (for in-process callers that
 * share device memory, e.g. tests comparing against torch). */
int pq_tensor_device_ptr(pq_ctx* ctx, pq_tensor t, void** base, int64_t* elem_offset,
                         int64_t* strides);

/* Per-kernel-class launch timing with HIP events on the context's stream.
 * enable=1 starts recording (clears totals), 0 stops.  Classes: see PQ_KC_*. */
#define PQ_KC_STREAM 0   /* big x small operand (bandwidth-bound) */
#define PQ_KC_GEMM 1     /* tiled complex GEMM on MFMA */
#define PQ_KC_GENERIC 2  /* general gathered contraction */
#define PQ_KC_REDUCE 3   /* split-K reduction for tiny outputs */
#define PQ_KC_COPY 4     /* strided copy / permutation */
#define PQ_KC_BATCHED 5  /* batched small contractions (one group per problem) */
#define PQ_KC_ACCUM 6    /* FP64 cross-slice accumulation */
#define PQ_KC_GEMM16 7   /* complex64 GEMM on FP16 hi/lo matrix cores, FP32 accumulation */
#define PQ_KC_CHAIN 8    /* fused chain of small operators on one large tensor (plan steps of kind 2) */
#define PQ_KC_COUNT 9
int pq_profile_enable(pq_ctx* ctx, int enable);
int pq_profile_read(pq_ctx* ctx, int kclass, int64_t* launches, double* total_ms,
                    double* alg_bytes, double* alg_flops);

/* Kernel selection override for tests and the autotuner: -1 = automatic,
 * otherwise force PQ_KC_STREAM / PQ_KC_GEMM / PQ_KC_GENERIC / PQ_KC_REDUCE
 * when the shape allows it (PQ_ERR_UNSUPPORTED otherwise). */
int pq_set_kernel_override(pq_ctx* ctx, int kclass);

/* GEMM workgroup tile override (-1 = automatic): 0:32x32 1:64x32 2:32x64 3:64x64
 * 4:128x64 5:64x128 6:64x16 7:128x16 8:256x16 9:16x64 10:128x32 11:128x128
 * 12:16x128 13:16x256 14:32x128, rows x columns of the product as launched
 * (complex128 tiles are capped at 4096 elements and 128 rows or columns). */
#define PQ_GEMM_TILES 15
int pq_set_gemm_tile(pq_ctx* ctx, int tile);

/* Arithmetic of complex64 GEMM steps for pq_contract / pq_time_step:
 * 0 = FP32 matrix cores (default), 1 = FP16 hi/lo split with FP32
 * accumulation where the shape allows it (K contiguous in both operands, a
 * multiple of 32; other steps keep FP32).  Plans set it per step. */
int pq_set_arithmetic(pq_ctx* ctx, int mode);

/* GEMM orientation override: -1 = automatic (the product is launched
 * transposed, C^T = B^T A^T, when C's unit stride lies on A's open side, so
 * that adjacent lanes store adjacent outputs), 0 = as given, 1 = transposed. */
int pq_set_gemm_orientation(pq_ctx* ctx, int mode);

/* Fused-chain element moves: -1 = automatic (a gather or scatter element
 * pair with stride 1, every other offset even and an aligned tensor moves
 * as one 16-byte access), 0 = always two 8-byte accesses.  Same results;
 * read when a plan run launches its chains (eager runs, and the capture of
 * a graph run). */
int pq_set_chain_pairs(pq_ctx* ctx, int mode);

/* ------------------------------------------------------------------------
 * Planned execution: a whole pair plan (contract_network!, src/layer2.jl:
 * 294-321) over every slice of a sliced network (src/layer2/slicing.jl),
 * summed over slices in FP64, as one stream-ordered program that
 * pq_plan_execute can capture into a single hipGraph and replay.
 * Slots are tensors: leaves (context tensors that keep their sliced bonds)
 * or intermediates (in a plan-owned arena).  Each sliced bond has extent 2;
 * slice s sets bond j to bit j of s.  Slots that depend on a sliced bond are
 * "batched": one tensor per slice of a slice group (2^group_bits slices).
 * Modes are (extent, stride, stride) triples in elements of one slice:
 * M = A-open (stride in A, in C), N = B-open (in B, in C), K = summed (in A,
 * in B).  There is no counterpart in the reference, whose slices are
 * rebuilt and contracted one by one and summed on the host
 * (test/layer2_tests.jl:436-450).
 * ---------------------------------------------------------------------- */
typedef struct pq_plan pq_plan;
typedef struct {
  int32_t n_slots;
  const int64_t* slot_elems;        /* per-slice elements (leaves: all elements) */
  const int64_t* slot_batch_mask;   /* bit j set: the slot holds one copy per value of sliced
                                       bond j within a slice group (for j < group_bits) and is
                                       recomputed per slice group (j >= group_bits); 0 = the
                                       slot does not depend on any sliced bond */
  const pq_tensor* slot_leaf;       /* leaf tensor, 0 for intermediates */
  int32_t n_sliced;                 /* number of sliced bonds (extent 2 each) */
  const int64_t* slot_slice_stride; /* [n_slots * n_sliced] stride of sliced bond j in an unbatched
                                       slot that keeps it as an index (a leaf, or an intermediate
                                       computed for both values of the bond), 0 if absent; batched
                                       steps read such a slot at each slice's offset */
  int32_t group_bits;               /* slices per launch = 2^group_bits (bonds 0..group_bits-1) */
  int32_t n_steps;
  const int32_t* step_slots;        /* [n_steps * 3] A, B, C slots */
  const int32_t* step_kc;           /* [n_steps] PQ_KC_* or -1 (may be NULL) */
  const int32_t* step_tile;         /* [n_steps] GEMM tile or -1 (may be NULL) */
  const int32_t* step_nmodes;       /* [n_steps * 3] mode counts of M, N, K */
  const int64_t* step_modes;        /* concatenated (extent, s0, s1) triples */
  int32_t out_slot;                 /* per-slice result slot */
  const int32_t* slot_dtype;        /* [n_slots] PQ_C128 / PQ_C64 (NULL = context type);
                                       leaves must have the context's type */
  const int32_t* step_kind;         /* [n_steps] 0 = contract A x B -> C, 1 = convert A -> C
                                       (element type change, same layout), 2 = fused chain A -> C
                                       (program in chain_words; B = its first operator slot);
                                       NULL = all contract */
  const int32_t* step_arith;        /* [n_steps] complex64 GEMM arithmetic as for pq_set_arithmetic
                                       (NULL = 0, FP32 matrix cores) */
  const int64_t* step_chain;        /* [n_steps] kind-2 steps: offset of the step's program in
                                       chain_words (NULL when no step has kind 2) */
  int64_t n_chain_words;
  const int32_t* chain_words;       /* fused-chain programs (PQ_CHAIN_* below) */
  int32_t n_out_bits;               /* the last n_out_bits sliced bonds (all above the slice group)
                                       number output batches instead of slices: the result holds
                                       2^n_out_bits blocks of the output slot's elements, block b
                                       summed over the slices whose values of those bonds spell b
                                       (bond n_sliced - n_out_bits + k = bit k of b).  A batch of
                                       bitstrings: each closed output qubit is a leaf with one
                                       column per bitstring, read through these bonds' strides */
  int32_t n_out_qubits;             /* closed output qubits of such a batch (0: none), for
                                       pq_plan_set_bitstrings */
  const int32_t* out_qubit_slot;    /* [n_out_qubits] the qubit's leaf slot: a contiguous
                                       [2, 2^n_out_bits] tensor, column b = the qubit's output
                                       vector for bitstring b ([1, 0] for 0, [0, 1] for 1,
                                       src/layer3.jl:292-313) */
  const int32_t* out_qubit_bit;     /* [n_out_qubits] the qubit's bit in a bitstring word (0-63) */
} pq_plan_desc;

/* A fused chain applies L small complex64 operators one after the other to
 * one large complex64 tensor, as L contraction steps would (each step sums
 * an operator's K bits against bits of the running tensor and adds its M
 * bits), without writing the L - 1 intermediates: one workgroup per
 * configuration of the untouched ("spectator") modes gathers a tile of
 * 2^TB elements into LDS, applies the operators on matrix cores in place,
 * and scatters the tile to the output.  Every extent is split into bits
 * (extent-2 factors); a bit has a tile position and strides.  Program words:
 *   L, TB, NIN, NOUT, NSPEC,
 *   NIN  x (tile position, stride in A)       gathered bits, A-stride order
 *   NOUT x (tile position, stride in C)       scattered bits, C-stride order
 *   NSPEC x (extent, stride in A, stride in C)
 *   L x (operator slot, NK, NM, NR,
 *        NK x (tile position, stride in the operator),
 *        NM x (tile position, stride in the operator),
 *        NR x tile position)
 * An operator element is S[sum of k-bit strides + sum of m-bit strides]; the
 * stage computes T[r, m] = sum_k S[m, k] T[r, k] for every r (bits at the R
 * positions), writing m at the M positions.  The host checks every gather,
 * scatter and operator access against its slot. */
#define PQ_CHAIN_MAX_STAGES 8
#define PQ_CHAIN_MAX_TILE_BITS 13
#define PQ_CHAIN_MAX_OP_BITS 5
/* A chain's operators (8 bytes per element) and its per-stage K, M and R
 * offsets (4 bytes each; 2^NK + 2^NM + 2^min(NR, 4) + 2^(NR - min(NR, 4))
 * per stage) are staged in LDS beside the tiles; together they may take this
 * many bytes. */
#define PQ_CHAIN_STAGE_BYTES 16384

/* The plan holds a reference on every leaf's device buffer and reads the
 * leaf at the offset and strides it had here: freeing or reshaping a leaf
 * tensor afterwards does not affect the plan.  Destroy plans before their
 * context. */
int pq_plan_create(pq_ctx* ctx, const pq_plan_desc* desc, pq_plan** out);
/* Enqueue one full run (all slices + FP64 accumulation); use_graph=1 captures
 * the run into a hipGraph on first use and replays it afterwards. */
int pq_plan_execute(pq_ctx* ctx, pq_plan* plan, int use_graph);
/* Copy the FP64 result (2 * out_elems doubles, interleaved; out_elems as
 * pq_plan_stats reports it, all output batches) to the host. */
int pq_plan_result(pq_ctx* ctx, pq_plan* plan, double* host);
int pq_plan_stats(const pq_plan* plan, int64_t* arena_bytes, int64_t* launches_per_run,
                  double* alg_flops, double* alg_bytes, int64_t* out_elems);
int pq_plan_step_info(const pq_plan* plan, int step, int* kclass, int* tile, int* dependent,
                      double* alg_flops, double* alg_bytes);
/* Exact cost of one run: complex multiply-adds and algorithmic bytes (each
 * operand read once, each result written once) summed over steps and slice
 * groups, as the reference's Metrics ledger counts MACs
 * (src/backends.jl:38-55); per-step figures into step_macs / step_bytes when
 * they are not NULL (pq_plan_stats' step count entries each). */
typedef struct {
  double macs;
  double bytes;
  int64_t steps;
  int64_t launches;
} pq_cost;
int pq_plan_cost(const pq_plan* plan, pq_cost* total, double* step_macs, double* step_bytes);
/* A batch of output bitstrings for a plan compiled with n_out_qubits > 0:
 * bitstring b's word has bit out_qubit_bit[q] = the value closed output
 * qubit q takes (the reference's add_output! per bitstring,
 * src/layer3.jl:292-313).  1 <= n_bitstrings <= 2^n_out_bits; a shorter
 * batch is padded with copies of the first bitstring.  Rewrites the qubits'
 * leaves in place (stream-ordered before later runs), so a captured graph
 * computes the new batch on its next replay. */
int pq_plan_set_bitstrings(pq_ctx* ctx, pq_plan* plan, int n_bitstrings, const uint64_t* bitstrings);
/* One call for the batched path: the bitstrings (if n_bitstrings > 0), one
 * graph-replayed run over every slice, and the FP64 result into out_fp64
 * (if not NULL; as pq_plan_result, one block per bitstring). */
int pq_plan_run(pq_ctx* ctx, pq_plan* plan, int n_bitstrings, const uint64_t* bitstrings, double* out_fp64);
/* One eager run with HIP events around every step; step_ms[i] = time of step i
 * summed over all slice groups. */
int pq_plan_profile(pq_ctx* ctx, pq_plan* plan, double* step_ms);
int pq_plan_destroy(pq_ctx* ctx, pq_plan* plan);
/* Streams a plan created afterwards spreads independent launches over: the
 * context's stream plus extra streams of default priority, with launches
 * that share no data and no reused arena region on different streams (a
 * captured graph then holds them as parallel branches).  Only n = 1 (the
 * default, every launch on the context's stream) is accepted in this build:
 * n > 1 returns PQ_ERR_UNSUPPORTED -- a test run with four streams failed
 * mixed-precision plans and ended in a host segmentation fault whose cause
 * is not found. */
int pq_set_plan_streams(pq_ctx* ctx, int n);

/* ------------------------------------------------------------------------
 * Cost model and autotuning (no counterpart in the reference, whose planners
 * rank plans by FLOP counts only, src/layer2/bgreedy_contraction.jl:53-66).
 * ---------------------------------------------------------------------- */
/* Time one pairwise contraction shape: modes as in pq_plan_desc (M, N, K
 * triples), `batch` entries packed one after another; kclass / tile as for
 * pq_set_kernel_override / pq_set_gemm_tile (-1 = automatic).  Writes the
 * median of `reps` event-timed launches in microseconds. */
int pq_time_step(pq_ctx* ctx, int dtype, int nm, int nn, int nk, const int64_t* modes, int batch,
                 int kclass, int tile, int reps, double* us);

/* 3-layer perceptron (in -> hidden -> hidden -> 1, ReLU, FP32) with HIP
 * forward, backward and Adam kernels.  Parameters as one FP32 vector:
 * W1[in][H] b1[H] W2[H][H] b2[H] w3[H] b3.  X is row-major [n][in]. */
typedef struct pq_mlp pq_mlp;
int pq_mlp_create(pq_ctx* ctx, int in_dim, int hidden, uint64_t seed, pq_mlp** out);
int pq_mlp_destroy(pq_mlp* mlp);
int pq_mlp_num_params(const pq_mlp* mlp);
int pq_mlp_get_params(pq_ctx* ctx, pq_mlp* mlp, float* host);
int pq_mlp_set_params(pq_ctx* ctx, pq_mlp* mlp, const float* host);
int pq_mlp_predict(pq_ctx* ctx, pq_mlp* mlp, const float* X, int n, float* out);
/* Mean squared error over the n samples and its gradient (np floats). */
int pq_mlp_loss_grad(pq_ctx* ctx, pq_mlp* mlp, const float* X, const float* y, int n, float* loss,
                     float* grad);
/* Adam on shuffled minibatches; final_loss = mean loss of the last epoch. */
int pq_mlp_train(pq_ctx* ctx, pq_mlp* mlp, const float* X, const float* y, int n, int epochs, int batch,
                 float lr, uint64_t seed, float* final_loss);

/* ------------------------------------------------------------------------
 * Optimal contraction order for small networks: replaces the
 * TensorOperations.optimaltree call behind netcon
 * (src/layer2/netcon_contraction.jl:14-41).  Host only (no context, no GPU).
 * Tensor t has rank[t] indices; labels / dims hold them all, tensor after
 * tensor.  A label in two tensors is summed, a label in one stays open.
 * Writes the cheapest pair plan by total complex multiply-adds
 * (prod of dims of A's and B's indices per step): plan[2 s], plan[2 s + 1]
 * name the operands of step s, ids < n_tensors are inputs and id
 * n_tensors + s is the result of step s; 2 (n_tensors - 1) entries.
 * PQ_ERR_UNSUPPORTED when the search exceeds max_work pair evaluations
 * (<= 0: 2^32) or n_tensors > 64. */
int pq_netcon(int n_tensors, const int32_t* rank, const int64_t* labels, const int64_t* dims, int64_t max_work,
              int32_t* plan, double* cost);

#ifdef __cplusplus
}
#endif
#endif /* PQ_HIP_H */
