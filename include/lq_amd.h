/*
 * lq_amd.h -- C-ABI of the MI355X state-vector engine for the lightning.qubit
 * hot path (gate application, expval/probs, adjoint Jacobian).
 *
 * Every entry point returns 0 on success and non-zero on failure; the
 * failure message (format "[file][Line:n][Method:f]: Error in PennyLane
 * Lightning: <msg>", as core/utils/Error.hpp:111-140) is available from
 * lq_last_error() on the calling thread. Plain pointers and sizes only.
 *
 * Reference paths are relative to pennylane_lightning/core/
 * (lq/ = simulators/lightning_qubit/). Wire w is amplitude-index bit n-1-w
 * (lq/StateVectorLQubit.hpp:950). Matrices are row-major with wires[0] the
 * most significant bit of the local index; they are complex64 for LQ_C64
 * states and complex128 for LQ_C128 states, like the reference bindings.
 *
 * Gate application is queued and executed in fused LDS passes when a result
 * is needed (a measurement, lq_sv_get_data, lq_sv_synchronize). Argument
 * errors are reported by the call that queues the gate.
 *
 * Handles are not thread-safe (the reference's state vectors are not
 * either, see lq/StateVectorLQubit.hpp); use one handle per thread.
 */
#ifndef LQ_AMD_H
#define LQ_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lq_sv_s *lq_sv_t;   /* device state vector + stream + op queue */
typedef struct lq_obs_s *lq_obs_t; /* observable tree */
typedef struct lq_ops_s *lq_ops_t; /* operation list (OpsData) */

enum { LQ_C64 = 0, LQ_C128 = 1 };

/* Message of the last failure on this thread. */
const char *lq_last_error(void);

/* Engine identification, replaces backend_info()/runtime_info()
 * (core/bindings/Bindings.hpp:284-367, lq/bindings/LQubitBindings.hpp:464-475). */
const char *lq_backend_info(void);
int lq_device_count(int *out);

/* ---- State vector --------------------------------------------------------
 * Replaces StateVectorLQubitManaged<T>(num_qubits)
 * (lq/StateVectorLQubitManaged.hpp:75-82; binding core/bindings/Bindings.hpp:876). */
int lq_sv_create(uint32_t num_qubits, int dtype, lq_sv_t *out);
int lq_sv_destroy(lq_sv_t sv);
int lq_sv_num_qubits(lq_sv_t sv, uint32_t *out);
int lq_sv_dtype(lq_sv_t sv, int *out);
/* resetStateVector (lq/StateVectorLQubit.hpp:959-963) */
int lq_sv_reset(lq_sv_t sv);
/* setBasisState(state, wires) (lq/StateVectorLQubit.hpp:943-953) */
int lq_sv_set_basis_state(lq_sv_t sv, const uint32_t *bits, const uint32_t *wires, uint32_t nw);
/* setStateVector(state, wires): writes the 2^nw amplitudes where the other
 * wires are 0, leaves the rest (lq/StateVectorLQubit.hpp:1005-1036) */
int lq_sv_set_state_vector(lq_sv_t sv, const void *state, const uint32_t *wires, uint32_t nw);
/* updateData / getState copy in and out (lq/bindings/LQubitBindings.hpp:196-243) */
int lq_sv_update_data(lq_sv_t sv, const void *host, uint64_t len);
int lq_sv_get_data(lq_sv_t sv, void *host, uint64_t len);
/* copy of another state of the same size and precision (device to device) */
int lq_sv_copy(lq_sv_t dst, lq_sv_t src);
/* executes queued gates and waits for the device */
int lq_sv_synchronize(lq_sv_t sv);
/* device address of the amplitudes (executes queued gates first) */
int lq_sv_device_ptr(lq_sv_t sv, void **out);
/* updateData / getState between the state and device memory of the caller
 * (a torch tensor's storage, another library's buffer): len amplitudes of the
 * state's precision, copied on the device. The caller's writes to src must be
 * complete before the call; the copy is complete when the call returns.
 * (Device-side form of lq/bindings/LQubitBindings.hpp:196-243.) */
int lq_sv_update_data_device(lq_sv_t sv, const void *src, uint64_t len);
int lq_sv_get_data_device(lq_sv_t sv, void *dst, uint64_t len);
/* The state's amplitudes as a DLPack tensor without a copy: *out is a
 * DLManagedTensor (dlpack.h v0.8 ABI; device kDLROCM, dtype complex64 or
 * complex128, shape {2^n}, compact). Queued gates are executed and waited
 * for first. The tensor shares the state's memory: gates applied later show
 * in it. It keeps that memory alive after lq_sv_destroy until its deleter
 * runs (the consumer calls it, or lq_dlpack_release for a tensor never
 * consumed). The zero-copy counterpart of the reference's array exchange
 * with numpy (core/bindings/Bindings.hpp:381-456, allocate_aligned_array;
 * LQubitBindings.hpp:196-243, updateData / getState). */
int lq_sv_dlpack(lq_sv_t sv, void **out);
int lq_dlpack_release(void *managed);

/* Recorded circuits (HIP graphs; no counterpart in the reference, whose
 * gates run eagerly on the CPU): lq_sv_record prepares |0...0>, applies
 * `ops` and keeps that program as a HIP graph captured on the state's
 * stream (program upload, state preparation, every fused pass).
 * lq_rec_replay runs the same circuit again from |0...0> with new
 * parameters: `params` holds every op's parameters in op order (nparams =
 * the recorded ops' total, including the 0.0 placeholders of matrix ops).
 * A replay rewrites the pass coefficients and launches the graph; when the
 * new parameters plan differently (specialised kernels compiled since the
 * capture, a rotation that changes form) the circuit is captured again
 * (lq_rec_stats counts captures and replays). Queued gates of the state are
 * discarded by both calls. The recording keeps the state alive; replays of
 * one recording are serialised on the state's stream; the state is not
 * synchronised on return (measurements and lq_sv_synchronize order after
 * it). Intended for parameter loops over one circuit structure (variational
 * training, serving repeated request shapes). */
typedef struct lq_rec_s *lq_rec_t;
int lq_sv_record(lq_sv_t sv, lq_ops_t ops, lq_rec_t *out);
int lq_rec_replay(lq_rec_t rec, const double *params, uint64_t nparams);
int lq_rec_stats(lq_rec_t rec, uint64_t *captures, uint64_t *replays);
int lq_rec_destroy(lq_rec_t rec);

/* applyOperation(name, [controlled_wires, controlled_values,] wires, inverse, params)
 * (lq/StateVectorLQubit.hpp:378-423; binding core/bindings/Bindings.hpp:213-277).
 * nc = 0 for the plain form. "Identity" is a no-op. */
int lq_sv_apply_op(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, const uint32_t *ctrls,
                   const uint8_t *ctrl_vals, uint32_t nc, int inverse, const double *params, uint32_t np);

/* applyOperations(ops, wires, inverses, params) (core/simulators/base/StateVectorBase.hpp:127-172):
 * nops named gates in order, one call. names: the names back to back, each
 * NUL-terminated; op i acts on wires[wire_off[i] .. wire_off[i+1]) with
 * params[param_off[i] .. param_off[i+1]) (param_off and params may be null
 * for parameter-free ops). */
int lq_sv_apply_operations(lq_sv_t sv, uint32_t nops, const char *names, const uint32_t *wire_off,
                           const uint32_t *wires, const uint8_t *inverses, const uint32_t *param_off,
                           const double *params);
/* applyMatrix / applyControlledMatrix (lq/StateVectorLQubit.hpp:571-710;
 * binding core/bindings/Bindings.hpp:152-200); at most 10 target wires */
int lq_sv_apply_matrix(lq_sv_t sv, const void *matrix, const uint32_t *wires, uint32_t nw, const uint32_t *ctrls,
                       const uint8_t *ctrl_vals, uint32_t nc, int inverse);
/* applyPauliRot(wires, inverse, {theta}, word) (lq/StateVectorLQubit.hpp:492-501) */
/* applySparseMatrix / applyControlledSparseMatrix (lq/StateVectorLQubit.hpp:725-860,
 * GateImplementationsLM.hpp:516-574): a CSR 2^nw x 2^nw operator on `wires`
 * (wires[0] the local MSB), values in the state's precision; inverse is
 * rejected as in the reference. Applied through the dense fused path (nw <= 10). */
int lq_sv_apply_sparse_matrix(lq_sv_t sv, const int64_t *indptr, uint64_t indptr_len, const int64_t *indices,
                              const void *values, uint64_t nnz, const uint32_t *wires, uint32_t nw,
                              const uint32_t *ctrls, const uint8_t *ctrl_vals, uint32_t nc, int inverse);
int lq_sv_apply_pauli_rot(lq_sv_t sv, const uint32_t *wires, uint32_t nw, int inverse, double theta,
                          const char *word);
/* applyGenerator (lq/StateVectorLQubit.hpp:527-560): state <- G state, the
 * non-control subspace zeroed; *scale receives s with U = exp(i s theta G) */
int lq_sv_apply_generator(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, const uint32_t *ctrls,
                          const uint8_t *ctrl_vals, uint32_t nc, int adj, double *scale);
/* normalize / collapse (lq/StateVectorLQubit.hpp:881-921) */
int lq_sv_normalize(lq_sv_t sv);
/* <psi|psi> of the state after its queued gates, FP64 compensated sum (the
 * norm computeNorm of normalize, lq/StateVectorLQubit.hpp:908-921); for a
 * complex64 state its distance from 1 is the rounding drift of the run */
int lq_sv_norm2(lq_sv_t sv, double *out);
int lq_sv_collapse(lq_sv_t sv, uint32_t wire, int branch);

/* ---- Measurements (lq/measurements/MeasurementsLQubit.hpp) ---------------- */
/* probs() when all != 0, else probs(wires) (:90-163); out has 2^n or 2^nw doubles */
int lq_probs(lq_sv_t sv, const uint32_t *wires, uint32_t nw, int all, double *out);
/* No counterpart in the reference. A variant of lq_probs(wires) for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops): the state is the caller's device memory (2^num_qubits
 * amplitudes, complex64 for dtype LQ_C64, complex128 for LQ_C128), the
 * result 2^nw doubles in the caller's device memory out_dev (wires[0] the
 * most significant bit of the result index, as lq_probs), enqueued on
 * `stream` (a hipStream_t; NULL: the null stream); the caller synchronizes.
 * No allocation: the per-block partial histograms go to a fixed device array
 * of 512 x 2^10 doubles (4 MiB, held by each GPU the library runs on until
 * the process ends), so nw is limited to 1..10; every call on a GPU shares
 * that array, and calls that overlap on one GPU race. lq_probs_noalloc_init
 * takes the array's address on the current device (optional: the first call
 * on a device does it otherwise). */
int lq_probs_noalloc_init(void);
int lq_probs_noalloc(const void *state_dev, uint32_t num_qubits, int dtype, const uint32_t *wires, uint32_t nw,
                     double *out_dev, void *stream);
/* expval(name, wires) for Identity/PauliX/PauliY/PauliZ/Hadamard (:247-305) */
int lq_expval_named(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, double *out);
/* several one-wire named expvals in one read of the state */
int lq_expval_named_batch(lq_sv_t sv, uint32_t count, const char *const *names, const uint32_t *wires,
                          double *out);
/* expval(matrix, wires) (:213-238) */
int lq_expval_matrix(lq_sv_t sv, const void *matrix, const uint32_t *wires, uint32_t nw, double *out);
/* expval(Observable) (:372-394) */
int lq_expval_obs(lq_sv_t sv, lq_obs_t obs, double *out);
/* var(name, wires) (:461-479): the operation applied by name, |O psi|^2 - Re<psi|O psi>^2 */
int lq_var_named(lq_sv_t sv, const char *name, const uint32_t *wires, uint32_t nw, double *out);
/* var(matrix, wires) (:488-506); matrix in the state's precision, row-major */
int lq_var_matrix(lq_sv_t sv, const void *matrix, const uint32_t *wires, uint32_t nw, double *out);
/* var(Observable) (:430-452, calculateObsVar :716-727) */
int lq_var_obs(lq_sv_t sv, lq_obs_t obs, double *out);
/* expval / var of a CSR matrix over the state index (:323-337, :616-636;
 * lq/utils/SparseLinAlg.hpp:42-63): indptr has 2^n + 1 entries, values are
 * complex in the state's precision */
int lq_expval_sparse(lq_sv_t sv, const int64_t *indptr, uint64_t indptr_len, const int64_t *indices,
                     const void *values, uint64_t nnz, double *out);
int lq_var_sparse(lq_sv_t sv, const int64_t *indptr, uint64_t indptr_len, const int64_t *indices,
                  const void *values, uint64_t nnz, double *out);
/* generate_samples(wires, shots) / generate_samples(shots) when all != 0
 * (:646-677): alias method over probs(wires) (MeasurementKernels.hpp:308-381)
 * drawn with std::mt19937 seeded with `seed` when seeded != 0 (the
 * Measurements seed, re-applied on every call as MeasurementsBase.hpp:77-84
 * does), else from std::random_device. out: shots x nw values 0/1, row-major,
 * column j = wires[j] */
int lq_generate_samples(lq_sv_t sv, const uint32_t *wires, uint32_t nw, int all, uint64_t shots, int seeded,
                        uint64_t seed, uint64_t *out);

/* ---- Observables (core/observables/Observables.hpp:128-584,
 *      lq/observables/ObservablesLQubit.hpp:52-366) ---------------------- */
int lq_obs_named(const char *name, const uint32_t *wires, uint32_t nw, lq_obs_t *out);
/* matrix is complex128, row-major 2^nw x 2^nw */
int lq_obs_hermitian(const void *matrix_c128, const uint32_t *wires, uint32_t nw, lq_obs_t *out);
int lq_obs_tensor(const lq_obs_t *factors, uint32_t count, lq_obs_t *out);
int lq_obs_hamiltonian(const double *coeffs, const lq_obs_t *terms, uint32_t count, lq_obs_t *out);
/* SparseHamiltonian(data, indices, offsets, wires) (lq/observables/ObservablesLQubit.hpp:372-436):
 * CSR over the full state index; values complex128 */
int lq_obs_sparse(const void *values_c128, const int64_t *indices, uint64_t nnz, const int64_t *offsets,
                  uint64_t offsets_len, const uint32_t *wires, uint32_t nw, lq_obs_t *out);
int lq_obs_destroy(lq_obs_t obs);

/* ---- Operation list: OpsData (core/algorithms/JacobianData.hpp:39-253),
 * built like create_ops_list (core/bindings/Bindings.hpp:772-799). Arrays
 * are flattened with offsets of length nops+1; matrices are interleaved
 * complex128 (empty for named gates). */
int lq_ops_create(uint32_t nops, const char *const *names, const uint32_t *param_off, const double *params,
                  const uint32_t *wire_off, const uint32_t *wires, const uint8_t *inverses, const uint32_t *mat_off,
                  const double *matrices, const uint32_t *ctrl_off, const uint32_t *ctrls, const uint8_t *ctrl_vals,
                  lq_ops_t *out);
int lq_ops_destroy(lq_ops_t ops);
int lq_ops_num_par_ops(lq_ops_t ops, uint64_t *out);
/* apply every op of the list to the state (AdjointJacobianBase::applyOperations,
 * core/algorithms/AdjointJacobianBase.hpp:75-100) */
int lq_sv_apply_ops(lq_sv_t sv, lq_ops_t ops, int adjoint);

/* ---- Adjoint Jacobian: AdjointJacobian<SV>::adjointJacobian
 * (lq/algorithms/AdjointJacobianLQubit.hpp:347-491; binding
 * core/bindings/Bindings.hpp:698-729). `sv` holds the final state (or the
 * initial state with apply_operations != 0); trainable indexes the ops with
 * parameters. jac receives nobs*ntp doubles in [obs][param] order; expvals,
 * if not NULL, receives <obs> on the final state. sv is not modified. */
int lq_adjoint_jacobian(lq_sv_t sv, lq_ops_t ops, const lq_obs_t *obs, uint32_t nobs, const uint64_t *trainable,
                        uint32_t ntp, int apply_operations, double *jac, double *expvals);
/* ---- Vector-Jacobian product: VectorJacobianProduct<SV>::operator()
 * (lq/algorithms/VectorJacobianProduct.hpp:43-160; binding
 * lq/bindings/LQubitBindings.hpp:412-457). dy is a host array of len (= 2^n)
 * amplitudes in the handle's precision (the state cotangent); out receives
 * ntp complex values as (re, im) doubles, w_t = <dy | d psi / d theta_t>.
 * sv is not modified. */
int lq_vector_jacobian_product(lq_sv_t sv, lq_ops_t ops, const void *dy, uint64_t len, const uint64_t *trainable,
                               uint32_t ntp, int apply_operations, double *out);
/* ---- Instrumentation (no reference counterpart) ----------------------------
 * When enabled, every tile-pass kernel launch on the handle's stream is
 * bracketed by HIP events; profile_read totals the measured kernel time, the
 * number of passes and their algorithmic HBM bytes (amplitudes read +
 * written). lq_sv_stream gives the handle's HIP stream (hipStream_t). */
int lq_sv_profile(lq_sv_t sv, int enable);
int lq_sv_profile_read(lq_sv_t sv, int reset, uint64_t *passes, double *ms, double *bytes, double *max_ms);
int lq_sv_stream(lq_sv_t sv, void **out);
/* Per-pass records of the profiled passes since the last reset: up to `cap`
 * rows of 7 doubles (kernel ms, algorithmic bytes, ops, register stages, LDS
 * transposes, 1 if specialised, tile-bit mask); *count = rows available. */
int lq_sv_profile_passes(lq_sv_t sv, double *out, uint64_t cap, uint64_t *count);

/* <a|b> of two states of the same size, either precision, accumulated in FP64
 * (no reference counterpart; used for the complex64 fidelity reports) */
int lq_sv_overlap(lq_sv_t a, lq_sv_t b, double *re, double *im);

/* free the adjoint's work states held by the handle */
int lq_sv_release_workspace(lq_sv_t sv);

/* ---- Request server (no reference counterpart) ------------------------------
 * Energy + adjoint-gradient requests (ops list, Pauli-sum Hamiltonian,
 * trainable parameters) served with continuous batching on one GPU: requests
 * of up to 16 qubits run many per launch, one workgroup each, with lambda and
 * H lambda held in LDS (complex128: n <= 12, complex64: n <= 13) or in a
 * global scratch area; larger ones run through the engine on concurrent workers,
 * admitted while 3 x 2^n amplitudes fit the HBM budget. batched = 0 gives the
 * plain path: one request at a time, in arrival order, through the engine. */
typedef struct lq_server_s *lq_server_t;
typedef struct {
    int dtype;               /* LQ_C64 / LQ_C128 */
    int batched;             /* 1: continuous batching; 0: plain FIFO */
    uint64_t hbm_budget;     /* bytes; 0: 250 GiB */
    int small_max_qubits;    /* -1: sizes whose states fit the packed kernel's LDS (12 qubits complex128, 13 complex64); at most 16 */
    int max_batch;           /* requests per packed launch (0: 4096) */
    int batch_streams;       /* packed launches in flight (0: 2) */
    int large_workers;       /* concurrent engine requests (0: 8) */
    int prep_threads;        /* host threads building packed programs (0: 4) */
    int lane_workers;        /* engine workers that take only requests of <= lane_max_qubits, oldest
                                first, besides the large_workers (0: 2; negative: none) */
    int lane_max_qubits;     /* (0: 16) */
} lq_server_config;
int lq_server_create(const lq_server_config *cfg, lq_server_t *out);
int lq_server_destroy(lq_server_t srv);
int lq_server_submit(lq_server_t srv, uint32_t num_qubits, lq_ops_t ops, lq_obs_t ham, const uint64_t *trainable,
                     uint32_t ntp, uint64_t *id);
/* Wait for request `id` (timeout in seconds); *done = 0 on timeout. grad
 * receives ntp values; latency_ms is submit -> result; path 0 = packed,
 * 1 = engine. A failed request returns its error. */
int lq_server_wait(lq_server_t srv, uint64_t id, double timeout_s, int *done, double *energy, double *grad,
                   uint32_t ntp, double *latency_ms, int *path);
/* Submit `count` requests at arrival_s[i] seconds after the call (request i:
 * n[i] qubits, ops[i], ham[i], trainable tp[tp_off[i] .. tp_off[i+1])) and
 * collect every result: energies[i], grads[tp_off[i] ..], latency_ms[i],
 * path[i]; *wall_s = first arrival -> last completion. */
int lq_server_replay(lq_server_t srv, uint32_t count, const uint32_t *n, const lq_ops_t *ops, const lq_obs_t *ham,
                     const uint64_t *tp, const uint64_t *tp_off, const double *arrival_s, double *energies,
                     double *grads, double *latency_ms, int *path, double *wall_s);
/* submitted, completed, packed, engine, packed launches, HBM bytes held */
int lq_server_stats(lq_server_t srv, uint64_t *out6);
/* Engine states the server allocated and evicted for room since it started
 * (the workers share one pool of idle states, reused by qubit count). */
int lq_server_pool_stats(lq_server_t srv, uint64_t *made, uint64_t *freed);
/* Packed-path timing (no reference counterpart: server telemetry), longest
 * since the server started or the last reset != 0 call, in ms: out6[0] a
 * request's wait from its program being ready to its launch, out6[1] the host
 * time of one launch call, out6[2] one launch's time until the batcher saw it
 * done, out6[3..5] that launch's upload, kernel and download times on the
 * GPU (event timestamps). */
int lq_server_timing(lq_server_t srv, double *out6, int reset);

/* ---- Kernel specialisation (no reference counterpart) ----------------------
 * Register tile passes of states with at least LQ_JIT_MIN_QUBITS qubits
 * (default 16) are compiled into straight-line kernels in background threads
 * and cached by structure; until a kernel is ready its pass runs in the
 * generic tile kernel. LQ_JIT=0 disables this, LQ_JIT_SYNC=1 compiles before
 * first use. lq_jit_wait blocks until every queued compile is done;
 * lq_jit_stats reports kernels loaded, compiles failed and compiles pending;
 * lq_sv_pass_counts reports how many passes of the handle ran specialised and
 * generic kernels. lq_jit_shutdown drops queued compiles and waits for those
 * in flight (no new compiles start afterwards); call it before the process
 * exits when compiles may be in flight -- the Python package does so from
 * its exit hook. lq_set_split_fallback chooses what runs while some
 * specialised kernels of a program are missing: 0 the unsplit plan in the
 * generic kernels, 1 (default, or LQ_SPLIT_FALLBACK) the cheaper of the two
 * by an op-count estimate, 2 the split plan with its missing passes in the
 * generic kernel; the previous mode is returned in *previous (may be NULL). */
int lq_jit_wait(void);
int lq_jit_shutdown(void);
int lq_jit_stats(int *ready, int *failed, int *pending);
int lq_sv_pass_counts(lq_sv_t sv, int reset, uint64_t *specialised, uint64_t *generic);
int lq_set_split_fallback(int mode, int *previous);
/* Matrix-core path for complex64 dense operators on exactly the five
 * lowest index bits (wires n-5 .. n-1 in any order) without controls:
 * with on != 0, applyMatrix of such an operator runs as one 64x64 real
 * GEMM per 32-amplitude group on the MFMA units (v_mfma_f32_16x16x4_f32)
 * instead of as a fused tile op; default on (LQ_MFMA_DENSE5). Replaces the
 * dense k-qubit kernel for that case (lq/gates/cpu_kernels/
 * GateImplementationsLM.hpp:454-498, applyNQubitOp). */
int lq_set_mfma_dense5(int on, int *previous);
/* Engine options by name (process-wide; *previous may be NULL):
 * "split_fallback" (0..2, as lq_set_split_fallback), "mfma_dense5" (0/1, as
 * lq_set_mfma_dense5), "product_fold" (0/1: after a reset the leading
 * one-qubit gates of the next program fold into the product state its first
 * pass forms; default 0, LQ_PRODUCT_FOLD), "plan_search" (0/1: states of 26
 * qubits and more plan specialised programs under several pass budgets and
 * keep the plan with the fewest passes; default 0, LQ_PLAN_SEARCH),
 * "adjoint_group" (0: one observable per adjoint sweep; 1: up to three
 * observables share a sweep once that sweep's specialised kernels are
 * compiled and its estimated state traffic is below one sweep per
 * observable; 2: share whenever compiled; default 1, LQ_ADJOINT_GROUP). */
int lq_set_engine_option(const char *name, int value, int *previous);
/* Adjoint backward sweeps run in this process since the last reset: shared
 * by a group of observables (*grouped) and for one observable (*single). */
int lq_adjoint_sweep_counts(int reset, uint64_t *grouped, uint64_t *single);

#ifdef __cplusplus
}
#endif

#endif /* LQ_AMD_H */
