/*
 * jet_hip.h -- C ABI of libjet_hip.so, the MI355X (gfx950) pairwise
 * tensor-contraction path of Jet.
 *
 * All entry points are extern "C", take plain pointers and sizes, return an
 * int status (JET_HIP_SUCCESS == 0) and never throw.  On failure the message
 * is available from jet_hip_last_error() (thread-local), formatted like the
 * reference's Jet::Exception text (include/jet/Abort.hpp:90-97) so that a C++
 * wrapper can rethrow it unchanged.
 *
 * Tensors are dense complex arrays in device memory.  Modes are int32 labels;
 * every entry point takes a layout flag:
 *   JET_HIP_ROW_MAJOR -- last mode fastest, the convention of Jet::Tensor
 *                        (include/jet/Tensor.hpp:94-96);
 *   JET_HIP_COL_MAJOR -- first mode fastest, the convention of
 *                        Jet::CudaTensor (include/jet/CudaTensorHelpers.hpp:117-127).
 *
 * Reentrancy: every function may be called concurrently from several host
 * threads (Taskflow workers call ContractTensors concurrently,
 * include/jet/TaskBasedContractor.hpp:382-394).  Work is ordered only by the
 * hipStream_t passed in (NULL = the context's stream); plans are immutable
 * after creation.
 */
#ifndef JET_HIP_H
#define JET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JET_HIP_VERSION_MAJOR 0
#define JET_HIP_VERSION_MINOR 1
#define JET_HIP_VERSION_PATCH 0

typedef enum {
    JET_HIP_SUCCESS = 0,
    JET_HIP_ERROR_INVALID_VALUE = 1,
    JET_HIP_ERROR_NOT_SUPPORTED = 2,
    JET_HIP_ERROR_OUT_OF_MEMORY = 3,
    JET_HIP_ERROR_DEVICE = 4,
    JET_HIP_ERROR_INTERNAL = 5
} jet_hip_status_t;

typedef enum {
    JET_HIP_C64 = 0,  /* std::complex<float>  / cuComplex       */
    JET_HIP_C128 = 1  /* std::complex<double> / cuDoubleComplex */
} jet_hip_dtype_t;

typedef enum {
    JET_HIP_ROW_MAJOR = 0,
    JET_HIP_COL_MAJOR = 1
} jet_hip_layout_t;

/* Arithmetic of the complex GEMM inside a contraction. */
typedef enum {
    JET_HIP_PREC_DEFAULT = 0, /* FP32 for C64, FP64 for C128 (the reference's
                                 CUTENSOR_COMPUTE_32F / _64F,
                                 include/jet/CudaTensor.hpp:494-502)          */
    JET_HIP_PREC_FP32 = 1,    /* 4M complex GEMM on FP32 MFMA                  */
    JET_HIP_PREC_BF16X3 = 2,  /* 3M complex GEMM, each FP32 operand split into
                                 hi+lo BF16, three BF16 MFMA products per real
                                 product, FP32 accumulation.  C64 only.  Steps
                                 the plan finds HBM-bound run on FP32 MFMA.   */
    JET_HIP_PREC_FP64 = 3,    /* 4M complex GEMM on FP64 MFMA (C128 only)      */
    JET_HIP_PREC_BF16X3_4M = 4, /* as BF16X3, but the complex product in 4M
                                  form (12 BF16 MFMA per real-product triple;
                                  twice the MFMA work per operand byte of 3M
                                  and no Im = T3 - T1 - T2 cancellation)      */
    JET_HIP_PREC_F16X3_4M = 5  /* as BF16X3_4M with FP16 pieces: each operand
                                  row is scaled by a power of two into the
                                  FP16 range, split into hi+lo FP16 (22
                                  significant bits, BF16 keeps 16) and the
                                  scales are undone on the outputs.  Same
                                  MFMA rate; error close to FP32 MFMA.  C64.  */
} jet_hip_precision_t;

typedef struct jet_hip_ctx_s *jet_hip_ctx_t;
typedef struct jet_hip_plan_s *jet_hip_plan_t;
typedef struct jet_hip_network_s *jet_hip_network_t;
typedef struct jet_hip_batch_s *jet_hip_batch_t;

/* ---- library / errors -------------------------------------------------- */

int jet_hip_get_version(int *major, int *minor, int *patch);

/* Message of the last failing call on this thread ("" if none). */
const char *jet_hip_last_error(void);

/* ---- context, memory, streams (replaces the cudaMalloc / cudaMemcpy /
 *      CudaScopedDevice plumbing of include/jet/CudaTensor.hpp:145-214,
 *      358-394 and include/jet/CudaTensorHelpers.hpp:187-231) ------------- */

int jet_hip_ctx_create(int device, jet_hip_ctx_t *ctx);
int jet_hip_ctx_destroy(jet_hip_ctx_t ctx);
int jet_hip_ctx_device(jet_hip_ctx_t ctx, int *device);
/* The context's own (non-blocking, default-priority) stream. */
int jet_hip_ctx_stream(jet_hip_ctx_t ctx, void **stream);

/* Pool allocation, ordered on the context stream and synchronized before
 * return (usable from any stream).  jet_hip_free is ordered after the work
 * already queued on the context stream; work on other streams that uses the
 * buffer must be synchronized first. */
int jet_hip_malloc(jet_hip_ctx_t ctx, size_t bytes, void **ptr);
int jet_hip_free(jet_hip_ctx_t ctx, void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device.  Synchronous when
 * stream is NULL and async == 0. */
int jet_hip_memcpy(jet_hip_ctx_t ctx, void *dst, const void *src, size_t bytes,
                   int kind, void *stream, int async);
int jet_hip_memset(jet_hip_ctx_t ctx, void *dst, int value, size_t bytes,
                   void *stream);
int jet_hip_stream_synchronize(jet_hip_ctx_t ctx, void *stream);
int jet_hip_device_synchronize(jet_hip_ctx_t ctx);

/* ---- permutation (replaces cutensorPermutation in CudaTensor::Transpose,
 *      include/jet/CudaTensor.hpp:701-813, and the CPU
 *      Permuter<Backend>::Transpose, include/jet/permute/Permuter.hpp:49-119)
 *
 * out has modes out_modes (a permutation of in_modes); extents are those of
 * `in` in in_modes order.  in and out must not overlap.  Any rank <= 64 and
 * any extents; power-of-two extents take the LDS-tiled bit-permutation
 * kernel.  Identity permutations are the library's streaming device copy. */
int jet_hip_permute(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                    jet_hip_layout_t layout, int rank, const int32_t *in_modes,
                    const int64_t *extents, const int32_t *out_modes,
                    const void *in, void *out, void *stream);

/* ---- contraction (replaces CudaTensor::GetCudaContractionPlan /
 *      ContractTensorsWithoutAllocation / CudaContractionPlan,
 *      include/jet/CudaTensor.hpp:470-669)
 *
 * C[modes_c] = sum over modes in A and B but not C of A[modes_a]*B[modes_b].
 * Modes in A, B and C are batch modes; modes only in A (resp. B) must be in
 * C.  rank_c/modes_c may describe any order; ContractTensors' own order is
 * left ++ right (include/jet/Tensor.hpp:714-745).  No conjugation.         */
int jet_hip_contraction_plan_create(
    jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, jet_hip_precision_t precision,
    jet_hip_layout_t layout, int rank_a, const int32_t *modes_a,
    const int64_t *extents_a, int rank_b, const int32_t *modes_b,
    const int64_t *extents_b, int rank_c, const int32_t *modes_c,
    jet_hip_plan_t *plan);
/* Bytes of device workspace jet_hip_contract needs for this plan. */
int jet_hip_plan_workspace_size(jet_hip_plan_t plan, size_t *bytes);
/* Algorithmic real flops (8*M*N*K*batch) and the plan's kernel choice. */
int jet_hip_plan_info(jet_hip_plan_t plan, double *flops, int64_t *m,
                      int64_t *n, int64_t *k, int64_t *batch,
                      int *gemm_kernel);
int jet_hip_plan_destroy(jet_hip_plan_t plan);
/* workspace may be NULL: the context's pool then provides it (stream-
 * ordered).  a, b, c must not overlap. */
int jet_hip_contract(jet_hip_plan_t plan, const void *a, const void *b,
                     void *c, void *workspace, size_t workspace_bytes,
                     void *stream);

/* jet_hip_contract without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The same kernels as jet_hip_contract run with a fixed
 * per-device scratch array of JET_HIP_NOALLOC_SCRATCH_BYTES as the
 * workspace (the library holds it on each GPU it runs on until the process
 * ends); no hipMalloc, hipFree or other allocating call is made.  A plan
 * whose jet_hip_plan_workspace_size exceeds that limit is rejected with
 * JET_HIP_ERROR_INVALID_VALUE before anything is queued.  All calls on one
 * GPU share its scratch array: calls whose executions overlap on that GPU
 * (concurrent streams, or host threads without ordering) race and must be
 * serialized by the caller. */
#define JET_HIP_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int jet_hip_contract_noalloc(jet_hip_plan_t plan, const void *a,
                             const void *b, void *c, void *stream);

/* ---- GEMM kernel autotuning (replaces cuTENSOR's kernel selection behind
 *      cutensorInitContractionFind(CUTENSOR_ALGO_DEFAULT) and
 *      cutensorInitContractionPlan, include/jet/CudaTensor.hpp:616-640).
 *
 * Kernel ids (jet_hip_plan_info's gemm_kernel): 0 VALU per output, 1 VALU
 * block reduction, 2 FP32 MFMA, 3 FP64 MFMA, 4 split-BF16 3M, 5 split-BF16
 * 4M, 6 split-K with FP64 partials, 7 split-FP16 4M.  Variants: kernel 2
 * 1 = 128x128 / 2 = 64x128 / 3 = 128x64 block tiles; kernels 5 and 7
 * 1 = 128x128 / 2 = 128x192 block tiles / 3 = 128x128 tiles walked by a
 * persistent grid of one work-group per CU / 4 = 192x128 block tiles (long
 * side along M); kernel 0 1 = one thread per output
 * / 2 = short-K column blocks / 3 = short-K row threads (K <= 16); every
 * other kernel 1.
 *
 * The library holds a table of (precision mode, log2 M, log2 N, log2 K,
 * log2 batch) -> (kernel, variant), measured on gfx950 and compiled in.
 * Every plan (contraction, network, DAG) takes its GEMM from the entry of
 * its shape, else from the nearest entry of its mode whose choice can run
 * the shape, else from a built-in rule.  Entries are 7 int32 each:
 * {mode, log2 M, log2 N, log2 K, log2 batch, kernel, variant}, mode 0 FP32,
 * 1 split-BF16 3M, 2 FP64, 3 split-BF16 4M, 4 split-FP16 4M.            */

/* The kernel and variant a plan of this shape would use; source: 0 the
 * built-in rule, 1 the shape's own table entry, 2 the nearest entry.
 * Host-only (no device needed). */
int jet_hip_gemm_choice(jet_hip_dtype_t dtype, jet_hip_precision_t precision,
                        int64_t m, int64_t n, int64_t k, int64_t batch,
                        int *kernel, int *variant, int *source);
/* Replace the process-wide table with n entries (n = 0: built-in rule only;
 * n < 0: restore the compiled-in table); plans created afterwards use it. */
int jet_hip_tuning_set(int n, const int32_t *entries);
/* The current table: *n entries, the first `capacity` copied out. */
int jet_hip_tuning_get(int capacity, int32_t *entries, int *n);
/* Time every kernel and variant that can run the batched NT product
 * C[b][m][n] = sum_k A[b][m][k] B[b][n][k] of this shape and precision
 * (random operands in pool memory; the split modes' operand passes are part
 * of the time): mean ms over `reps` launches after one warm-up, per
 * candidate, in kernels / variants / ms (up to capacity), and the number of
 * candidates in *count.  Runs on the context stream and waits for it. */
int jet_hip_autotune_gemm(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                          jet_hip_precision_t precision, int64_t m, int64_t n,
                          int64_t k, int64_t batch, int reps, int capacity,
                          int32_t *kernels, int32_t *variants, float *ms,
                          int *count);

/* ---- elementwise (replaces cutensorElementwiseBinary in
 *      CudaTensor::AddTensors, include/jet/CudaTensor.hpp:38-137, and
 *      Tensor::AddTensors, include/jet/Tensor.hpp:412-454)
 * c[modes_a] = a[modes_a] + b[modes_b] with b transposed to a's order.   */
int jet_hip_add(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                jet_hip_layout_t layout, int rank, const int32_t *modes_a,
                const int64_t *extents_a, const int32_t *modes_b,
                const void *a, const void *b, void *c, void *stream);

/* ---- slicing (replaces CudaTensor::SliceIndex / Tensor::SliceIndex,
 *      include/jet/CudaTensor.hpp:865-905, include/jet/Tensor.hpp:493-526)
 * out = in with mode `axis` fixed at `value`; other modes keep their order. */
int jet_hip_slice(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                  jet_hip_layout_t layout, int rank, const int64_t *extents,
                  int axis, int64_t value, const void *in, void *out,
                  void *stream);

/* ---- seeded fill (replaces curandGenerateUniform in
 *      CudaTensor::FillRandom, include/jet/CudaTensor.hpp:419-446).
 * Counter-based (splitmix64) so that the same seed gives the same values on
 * the CPU: element i, part p (0 re, 1 im) uses counter 2*i+p.
 * distribution 0: uniform [0,1) ; 1: uniform [-1,1) ; 2: normal N(0, 1/2)
 * (re and im each variance 1/2).                                          */
int jet_hip_fill_random(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype, size_t n,
                        uint64_t seed, int distribution, void *out,
                        void *stream);

/* ---- FP64 compensated reduction of per-slice results (the cross-slice
 *      sum of TaskBasedContractor::AddReductionTask,
 *      include/jet/TaskBasedContractor.hpp:258-280).
 * out (device, 2 doubles: re, im) = Neumaier sum of n complex values.     */
int jet_hip_reduce_compensated(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                               size_t n, const void *x, double *out,
                               void *stream);

/* ---- contraction sequences (replaces the per-task ContractTensors calls
 *      that TaskBasedContractor schedules, include/jet/TaskBasedContractor.hpp:
 *      162-248, 382-394, and TensorNetwork::Contract(path),
 *      include/jet/TensorNetwork.hpp:301-328, for a whole path at once)
 *
 * Leaves: n_leaves row-major tensors; leaf i has leaf_ranks[i] modes, and the
 * modes / extents of all leaves are concatenated in leaf_modes / leaf_extents.
 * Path: n_steps pairs (path[2s], path[2s+1]) of SSA ids -- leaves are
 * 0..n_leaves-1 and step s produces node n_leaves+s (PathInfo's convention,
 * include/jet/PathInfo.hpp:262-297).
 * out_modes: the modes of the root, in the order it is written.  A mode in
 * out_modes is never summed by a step, even where both operands hold it: it
 * then becomes a batch mode.  Listing the sliced indices of a sliced network
 * there runs every slice in one sequence and leaves the per-slice results
 * in the root (the batched slice loop).  Modes held by more than two leaves
 * are summed once no remaining tensor holds them.
 * The network owns its intermediate arena and workspace (device memory,
 * allocated at creation).  Executions on one network must not overlap.    */
int jet_hip_network_create(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                           jet_hip_precision_t precision, int n_leaves,
                           const int32_t *leaf_ranks, const int32_t *leaf_modes,
                           const int64_t *leaf_extents, int n_steps,
                           const int32_t *path, int n_out,
                           const int32_t *out_modes, jet_hip_network_t *net);
/* The plan of jet_hip_network_create without a device or any allocation:
 * algorithmic flops and the arena / workspace bytes it would allocate. */
int jet_hip_network_estimate(jet_hip_dtype_t dtype,
                             jet_hip_precision_t precision, int n_leaves,
                             const int32_t *leaf_ranks,
                             const int32_t *leaf_modes,
                             const int64_t *leaf_extents, int n_steps,
                             const int32_t *path, int n_out,
                             const int32_t *out_modes, double *flops,
                             size_t *arena_bytes, size_t *workspace_bytes);
/* Algorithmic real flops (sum of 8*M*N*K*batch), arena and workspace bytes,
 * root size in elements, number of steps. */
int jet_hip_network_info(jet_hip_network_t net, double *flops,
                         size_t *arena_bytes, size_t *workspace_bytes,
                         int64_t *out_size, int *n_steps);
/* GEMM shape, kernel and permutations (bit 0: A, 1: B, 2: C) of one step. */
int jet_hip_network_step_info(jet_hip_network_t net, int step, int64_t *m,
                              int64_t *n, int64_t *k, int64_t *batch,
                              int *gemm_kernel, int *permutes);
int jet_hip_network_execute(jet_hip_network_t net, const void *const *leaves,
                            void *out, void *stream);
/* Capture one execution with these leaf and output pointers as a hipGraph
 * (replaces the cudaFlow capture of examples/heterogeneous_contraction.cu:
 * 91-142); jet_hip_network_launch replays it. */
int jet_hip_network_capture(jet_hip_network_t net, const void *const *leaves,
                            void *out, void *stream);
int jet_hip_network_launch(jet_hip_network_t net, void *stream);
int jet_hip_network_destroy(jet_hip_network_t net);

/* ---- contraction DAGs: sequences whose intermediates feed several steps
 *      (a forward contraction and its adjoint, environment pass).  The GPU
 *      counterpart of a TaskBasedContractor task graph in which tasks share
 *      subtrees by name (include/jet/TaskBasedContractor.hpp:206-240).
 * Step s contracts nodes steps[2s] and steps[2s+1] (leaves 0..n_leaves-1,
 * step s produces node n_leaves+s) over all the modes they share; its
 * result holds the first operand's other modes, then the second's.  A node
 * may feed any number of later steps.  The results of the steps listed in
 * keep (node ids) stay in the arena after an execution, in one contiguous
 * region in keep order; jet_hip_dag_value gives their addresses.  Executed
 * with jet_hip_network_execute / _capture / _launch with out = NULL, and
 * destroyed with jet_hip_network_destroy.                                  */
int jet_hip_dag_create(jet_hip_ctx_t ctx, jet_hip_dtype_t dtype,
                       jet_hip_precision_t precision, int n_leaves,
                       const int32_t *leaf_ranks, const int32_t *leaf_modes,
                       const int64_t *leaf_extents, int n_steps,
                       const int32_t *steps, int n_keep, const int32_t *keep,
                       jet_hip_network_t *net);
/* The plan of jet_hip_dag_create without a device or any allocation. */
int jet_hip_dag_estimate(jet_hip_dtype_t dtype, jet_hip_precision_t precision,
                         int n_leaves, const int32_t *leaf_ranks,
                         const int32_t *leaf_modes, const int64_t *leaf_extents,
                         int n_steps, const int32_t *steps, int n_keep,
                         const int32_t *keep, double *flops,
                         size_t *arena_bytes, size_t *workspace_bytes);
/* Device address, rank and modes / extents (arrays of at least 64 entries,
 * may be NULL) of a kept step result. */
int jet_hip_dag_value(jet_hip_network_t net, int node, void **ptr, int *rank,
                      int32_t *modes, int64_t *extents);

/* ---- uneven batches: executions of several planned networks (different
 *      circuits, bitstrings, slice ranges) captured as one hipGraph and
 *      launched at once.  Replaces the composition of many contraction
 *      task graphs into one executor run (TaskBasedContractor::AddTaskflow,
 *      include/jet/TaskBasedContractor.hpp:147-155, and the cudaFlow of
 *      examples/heterogeneous_contraction.cu:91-142).
 * Execution i runs networks[i] on leaves[i] (that network's leaf pointers)
 * into outs[i].  A network may appear several times: its executions run in
 * the order given and share its arena.  All networks belong to ctx.      */
int jet_hip_batch_capture(jet_hip_ctx_t ctx, int n, const jet_hip_network_t *networks,
                          const void *const *const *leaves, void *const *outs,
                          void *stream, jet_hip_batch_t *batch);
int jet_hip_batch_launch(jet_hip_batch_t batch, void *stream);
int jet_hip_batch_destroy(jet_hip_batch_t batch);

/* ---- dtype conversion (Tensor<complex<float>> <-> Tensor<complex<double>>
 *      data, as used to build the FP64 oracle inputs on the device) ------- */
int jet_hip_convert(jet_hip_ctx_t ctx, jet_hip_dtype_t src_dtype,
                    jet_hip_dtype_t dst_dtype, size_t n, const void *in,
                    void *out, void *stream);

/* ---- events (timing on a stream; replaces the std::chrono timing around
 *      Contract().wait() of the reference's benchmark drivers,
 *      examples/paper_benchmarks/CPU/jet_cpu_m10/jet_sliced.cpp:85-91) ---- */
int jet_hip_event_create(jet_hip_ctx_t ctx, void **event);
int jet_hip_event_record(jet_hip_ctx_t ctx, void *event, void *stream);
int jet_hip_event_elapsed_ms(void *start, void *stop, float *ms);
int jet_hip_event_destroy(void *event);

/* ---- sustained FP16 MFMA rate (no counterpart in the reference: a
 *      measurement of this device).  A loop of v_mfma_f32_16x16x32_f16 --
 *      the instruction of the split GEMMs -- on register operands (uniform
 *      in (-1, 1) with full mantissas, like the split operands: the clock
 *      the chip holds depends on the operand bits), with no memory
 *      traffic, on four work-groups of four waves per CU: one warm-up
 *      and `reps` timed launches of `rounds` (a positive multiple of 4)
 *      rounds of eight MFMAs per wave, on the context stream.  *tflops: FP16
 *      MFMA TFLOP/s over the timed launches; *clock_ghz: the mean shader
 *      clock of the last launch's waves (clock-counter deltas over their
 *      constant-rate timer deltas).  The chip clocks down under a full MFMA
 *      load, so this rate, not the nominal peak at the 2.4 GHz peak clock,
 *      is the practical ceiling of MFMA-bound kernels.  Allocates.        */
int jet_hip_mfma_ceiling(jet_hip_ctx_t ctx, int rounds, int reps, double *tflops,
                         double *clock_ghz);

/* ---- per-kernel profiler (no counterpart in the reference, which has no
 *      tracer).  When enabled, each kernel launch is bracketed by events on
 *      its stream; jet_hip_profiler_read waits for them and returns, for each
 *      kernel family k < *count, launches[k], total device ms[k] and the
 *      algorithmic flops[k] / minimum HBM bytes[k] of those launches.      */
int jet_hip_profiler_enable(int enable);
int jet_hip_profiler_reset(void);
int jet_hip_profiler_read(int capacity, int64_t *launches, double *ms,
                          double *flops, double *bytes, int *count);
const char *jet_hip_profiler_kernel_name(int kernel);

#ifdef __cplusplus
}
#endif

#endif /* JET_HIP_H */
