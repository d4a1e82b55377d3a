/*
 * pmcmc_amd.h — C ABI of the MI355X (gfx950) DEER-over-MALA library.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (PMCMC_OK == 0) and never throws across the boundary; the message of the
 * last failure on a context is pmcmc_last_error(ctx).
 *
 * Memory: every array argument is a DEVICE pointer (hipMalloc'd memory or a
 * ROCArray/torch tensor's storage) unless a comment says otherwise.  Matrices
 * are column-major like Julia's: element (i, j) of an m x n matrix is at
 * i + j*m.  A trajectory is D x T, column t is the state after step t+1.
 *
 * Precision: PMCMC_PREC_F64 buffers are double; PMCMC_PREC_F32 and
 * PMCMC_PREC_BF16_JAC buffers are float (BF16 is internal to the Jacobian
 * path of the GEMM targets).
 *
 * Threading: a context owns one HIP stream on one device; calls on one
 * context are serialised by the caller (one context per Julia task / chain),
 * different contexts are independent.
 */
#ifndef PMCMC_AMD_H
#define PMCMC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PMCMC_OK 0
#define PMCMC_ERR_ARGUMENT 1  /* Julia ArgumentError */
#define PMCMC_ERR_DIMENSION 2 /* Julia DimensionMismatch */
#define PMCMC_ERR_HIP 3       /* HIP runtime error */
#define PMCMC_ERR_ALLOC 4     /* device allocation failed */
#define PMCMC_ERR_DEVICE 5    /* a kernel reported an error */
#define PMCMC_ERR_UNSUPPORTED 6

/* ---- enums -------------------------------------------------------------- */
#define PMCMC_PREC_F64 0
#define PMCMC_PREC_F32 1
#define PMCMC_PREC_BF16_JAC 2

#define PMCMC_JAC_STOCH_DIAG 0 /* Hutchinson diagonal, src/DEER/DEER.jl:398-419 */
#define PMCMC_JAC_DIAG 1       /* exact diagonal, src/DEER/DEER.jl:456-462 */
/* Dense Jacobians (dense Gaussian targets, D <= 64); the reference rejects
 * :full (src/DEER/DEER.jl:464-465).  FULL: derivative of the taped step with
 * the accept decision held fixed, a (I + eps H) + (1 - a) I.  FULL_SIGMOID:
 * the surrogate's dense Jacobian, the matrix of the JVPs of
 * src/MALA/MALA.jl:749-795 (adds g(1-g)(y-x) grad(log alpha)^T). */
#define PMCMC_JAC_FULL 2
#define PMCMC_JAC_FULL_SIGMOID 3

typedef struct pmcmc_ctx pmcmc_ctx;
typedef struct pmcmc_target pmcmc_target;
typedef struct pmcmc_deer_ws pmcmc_deer_ws;

/* ---- context ------------------------------------------------------------ */
/* stream: a hipStream_t, or NULL for a stream the library creates.          */
int pmcmc_ctx_create(int device, void* stream, pmcmc_ctx** out);
/* Drops the creator's reference.  Targets, workspaces and batches created
 * on the context hold references of their own, so they may be destroyed
 * before or after it (finalizers of a garbage-collected host language run in
 * no fixed order); the stream and buffers are released with the last one. */
int pmcmc_ctx_destroy(pmcmc_ctx* ctx);
const char* pmcmc_last_error(const pmcmc_ctx* ctx);
int pmcmc_ctx_synchronize(pmcmc_ctx* ctx);
/* Library version string (static storage). */
const char* pmcmc_version(void);

/* ---- measurement --------------------------------------------------------
 * No counterpart in the reference (its benchmarks time whole calls with
 * BenchmarkTools).  Per-kernel durations of DEER solves on this context,
 * from HIP events recorded on the context stream around each launch; only
 * launches of iterations up to convergence are counted.                     */
#define PMCMC_KT_GLM_PASS 0    /* fused GEMM->sigmoid->GEMM^T pass (MFMA) */
#define PMCMC_KT_TILE_REDUCE 1 /* level-0 tile reduce (MALA step + JVP + in-tile scan) */
#define PMCMC_KT_TILE_APPLY 2  /* level-0 tile apply (or the single-tile kernel) */
#define PMCMC_KT_GLM_REDUCE 3  /* split-sum + prior terms of a GEMM pass */
#define PMCMC_KT_SCAN_SPINE 4  /* scan of the tile aggregates (all levels) */
#define PMCMC_KT_ITERATION 5   /* one whole Newton iteration */
#define PMCMC_KT_LAMBDA_PASS 6 /* batched dense-Gaussian Lambda pass (pmcmc_batch_* only) */
int pmcmc_ctx_set_profiling(pmcmc_ctx* ctx, int on); /* on: every kind */
/* Profile only the kinds in `kinds` (bit k = 1u << PMCMC_KT_k; 0 turns
 * profiling off).  Each event pair adds a marker to the stream, so timing a
 * solve with every kind on slows it: a throughput run profiles one kind.   */
int pmcmc_ctx_set_profiling_kinds(pmcmc_ctx* ctx, unsigned kinds);
int pmcmc_ctx_kernel_time(pmcmc_ctx* ctx, int kind, int64_t* launches, double* total_ms);
int pmcmc_ctx_reset_kernel_times(pmcmc_ctx* ctx);

/* ---- device buffers (for hosts without their own device arrays) --------- */
int pmcmc_buffer_alloc(pmcmc_ctx* ctx, int64_t nbytes, void** out);
int pmcmc_buffer_free(pmcmc_ctx* ctx, void* ptr);
/* host <-> device copies; `host` is a HOST pointer.  Synchronous. */
int pmcmc_buffer_upload(pmcmc_ctx* ctx, void* dst_device, const void* host, int64_t nbytes);
int pmcmc_buffer_download(pmcmc_ctx* ctx, void* host, const void* src_device, int64_t nbytes);

/* ---- diagonal affine scan ----------------------------------------------
 * S[:,t] = A[:,t] .* S[:,t-1] + B[:,t], S[:,0] uses s0.  LDS-tiled
 * reduce-then-scan (no inter-workgroup waits).  Replaces
 * DEERScan.solve_affine_scan_diag!
 * (src/DEER/DEERScan.jl:102-173); S must alias neither A nor B (:124-125).
 * T == 0 is a no-op.                                                        */
int pmcmc_scan_diag_f32(pmcmc_ctx* ctx, const float* A, const float* B, const float* s0,
                        float* S, int64_t D, int64_t T);
int pmcmc_scan_diag_f64(pmcmc_ctx* ctx, const double* A, const double* B, const double* s0,
                        double* S, int64_t D, int64_t T);

/* ---- validation helpers of the diagonal scan -----------------------------
 * DEERScan.solve_affine_seq! (src/DEER/DEERScan.jl:51-68): the sequential
 * recurrence, one device thread per row, product and sum rounded separately
 * as the reference's broadcast does.  DEERScan.affine_scan_residual
 * (:203-224): *out = max over (d, t) of |S - (A .* S_prev + B)| (S_prev = s0
 * for t = 1; NaN when any entry is NaN).  The comparison of check_affine_scan
 * (:240-258): out[0] = max |X - Y|, out[1] = max |X - Y| / (atol + rtol |X|)
 * (NaN where both are 0, as the reference's ratio), out[2] = 1 when every
 * |X - Y| <= atol + rtol |X| (its `ok`), else 0; n entries.  out: HOST
 * pointers.  Synchronise the context stream.                                */
int pmcmc_affine_seq_f32(pmcmc_ctx* ctx, const float* A, const float* B, const float* s0, float* S,
                         int64_t D, int64_t T);
int pmcmc_affine_seq_f64(pmcmc_ctx* ctx, const double* A, const double* B, const double* s0, double* S,
                         int64_t D, int64_t T);
int pmcmc_affine_residual_f32(pmcmc_ctx* ctx, const float* S, const float* A, const float* B,
                              const float* s0, int64_t D, int64_t T, double* out);
int pmcmc_affine_residual_f64(pmcmc_ctx* ctx, const double* S, const double* A, const double* B,
                              const double* s0, int64_t D, int64_t T, double* out);
int pmcmc_affine_compare_f32(pmcmc_ctx* ctx, const float* X, const float* Y, int64_t n, double atol,
                             double rtol, double* out);
int pmcmc_affine_compare_f64(pmcmc_ctx* ctx, const double* X, const double* Y, int64_t n, double atol,
                             double rtol, double* out);

/* ---- allocation-free diagonal scan ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same scan as pmcmc_scan_diag_f32 (same kernels,
 * shapes and results), with its scratch (tile aggregates, look-back status
 * words) in a fixed per-device array of PMCMC_NOALLOC_SCRATCH_BYTES instead
 * of memory the context allocates.  It calls no hipMalloc / hipFree.  A D x T
 * whose scratch exceeds that size returns PMCMC_ERR_ARGUMENT (about
 * 12 D^2 T / 4096 bytes at D <= 1024: D = 64 up to T ~ 2.8M).  Every call on a
 * device shares that device's array, which keeps the last call's contents:
 * calls that overlap on one device (two contexts' streams) race.  The array's
 * device address is looked up by pmcmc_ctx_create.                         */
#define PMCMC_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int pmcmc_scan_diag_f32_noalloc(pmcmc_ctx* ctx, const float* A, const float* B, const float* s0,
                                float* S, int64_t D, int64_t T);

/* ---- dense-Jacobian affine scan -----------------------------------------
 * S[:,t] = J_t S[:,t-1] + B[:,t], S[:,0] uses s0, with J_t = J + t*D*D the
 * t-th D x D column-major matrix; D <= 64.  The general form of the :full
 * Newton update for callers that hold their Jacobians; no reference
 * counterpart (the reference rejects :full, src/DEER/DEER.jl:464-465; its
 * diagonal scan is src/DEER/DEERScan.jl:102-173).  Tile maps composed on chip
 * (64 x 64 products per column), a recursive scan of the tile maps, then
 * each tile run from its incoming state.  S aliases neither J nor B.
 * T == 0 is a no-op.  Synchronises the context stream.                     */
int pmcmc_scan_full_f32(pmcmc_ctx* ctx, const float* J, const float* B, const float* s0,
                        float* S, int64_t D, int64_t T);
int pmcmc_scan_full_f64(pmcmc_ctx* ctx, const double* J, const double* B, const double* s0,
                        double* S, int64_t D, int64_t T);

/* ---- owned matmul and reductions ----------------------------------------
 * src/ParallelMCMC.jl:23-25: pmcmc_matmul(A, B) = A * B,
 * pmcmc_dot(a, b) = dot(a, b), pmcmc_dotsum(A, B) = sum(A .* B).
 * Device pointers, column-major; prec PMCMC_PREC_F32 or PMCMC_PREC_F64.
 * pmcmc_matmul: C (m x n, ldc) = op(A) op(B) with op(A) m x k, op(B) k x n,
 * op = transpose when trans_* != 0 (the pullbacks of ext/EnzymeExt.jl:78-284
 * are dY B^T and A^T dY); k = 0 gives zeros; C aliases neither operand;
 * lda >= max(1, rows of A as stored), likewise ldb, ldc >= max(1, m).
 * MFMA tiles (f32 or f64), FP32 accumulation for FP32 operands.
 * pmcmc_dot / pmcmc_dotsum: FP64 accumulation in a fixed order; the result
 * (one element of the operands' type) is written to device memory at out.
 * pmcmc_dotsum sums over an m x n block of A and B (leading dims lda, ldb).
 * All three synchronise the context stream.                                */
int pmcmc_matmul(pmcmc_ctx* ctx, int prec, int trans_a, int trans_b, int64_t m, int64_t n,
                 int64_t k, const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                 int64_t ldc);
int pmcmc_dot(pmcmc_ctx* ctx, int prec, int64_t n, const void* a, const void* b, void* out);
int pmcmc_dotsum(pmcmc_ctx* ctx, int prec, int64_t m, int64_t n, const void* A, int64_t lda,
                 const void* B, int64_t ldb, void* out);

/* ---- targets -------------------------------------------------------------
 * Kernels cannot call user closures, so a DensityModel carries one of these
 * descriptors (the analytic counterpart of its logdensity / grad / hvp
 * callbacks, src/interface.jl:36-70).  Arrays are copied into the target at
 * creation; `dtype_prec` is the PMCMC_PREC_* of the buffers passed in.     */

/* N(mean, diag(1/prec)); prec/mean may be NULL (1 and 0: the standard normal). */
int pmcmc_target_diag_gaussian(pmcmc_ctx* ctx, int64_t D, int dtype_prec, const void* prec,
                               const void* mean, pmcmc_target** out);
/* Bayesian logistic regression with prior N(0, prior_scale^2 I);
 * Xd is N x D column-major, y has N entries in {0,1}.
 * benchmarks/ParallelMCMCBenchmarks/src/models/bayes_logreg.jl:91-206.      */
int pmcmc_target_logistic(pmcmc_ctx* ctx, int64_t N, int64_t D, int dtype_prec, const void* Xd,
                          const void* y, double prior_scale, pmcmc_target** out);
/* Replaces a logistic target's isotropic prior by N(0, P^-1) with the dense
 * precision P (D x D, symmetric), for every kernel that evaluates it
 * (gradient, Hessian-vector products, log-density, the FP64 decisions).  No
 * reference counterpart as a separate call: it carries a mass-matrix
 * preconditioner (ParallelMALASampler cholM, src/MALA/MALA.jl:29-47) into
 * the reparametrized target x = L theta, whose prior precision is
 * L' L / prior_scale^2.                                                     */
int pmcmc_target_logistic_set_prior(pmcmc_target* tgt, const void* P, int dtype_prec);
/* Dense Gaussian N(mean, Lambda^-1); Lambda is D x D symmetric (column-major),
 * mean may be NULL (zero).  The precision-matrix form of the correlated
 * Gaussian models of the reference's GPU tests (grad = -Lambda (x - mean),
 * Hv = -Lambda v), evaluated as MFMA GEMMs.                                  */
int pmcmc_target_gaussian(pmcmc_ctx* ctx, int64_t D, int dtype_prec, const void* Lambda,
                          const void* mean, pmcmc_target** out);
int pmcmc_target_destroy(pmcmc_target* tgt);
int64_t pmcmc_target_dim(const pmcmc_target* tgt);

/* ---- batched MALA step + surrogate JVP -----------------------------------
 * Per column t: the taped MALA step from X[:,t] with noise Xi[:,t] and
 * uniform U[t], and the JVP of the sigmoid surrogate along Z[:,t].
 * Replaces MALA.mala_step_batched_fwd_and_jvp! (src/MALA/MALA.jl:556-628),
 * the fwd_and_jvp_batch(Xbar, Z) plug-in of DEER.TapedRecursion
 * (src/DEER/DEER.jl:17-38).  accepted[t] = (log U[t] < logalpha[t]).
 * logalpha may be NULL; it is double for every precision.                  */
int pmcmc_mala_fwd_and_jvp(pmcmc_ctx* ctx, const pmcmc_target* tgt, int prec, double eps,
                           const void* X, const void* Xi, const void* U, const void* Z, void* FT,
                           void* Jt, uint8_t* accepted, double* logalpha, int64_t D, int64_t T);

/* Sequential MALA on the device for N chains of one target, K steps: step k
 * is mala_step_batched! (src/MALA/MALA.jl:487-528) on the N columns of the
 * previous step's states (X0 first, D x N): proposal, both log-densities and
 * gradients, accept iff log u < log alpha.  Xi and S_out are D x (N K)
 * column-major with step k's N columns in block k, U and accepted have N K
 * entries likewise.  The GPU comparator (and warm-up path) of DEER; the
 * multi-step driver has no reference counterpart (the reference loops over
 * the one-step call).  Synchronises the context stream.                    */
int pmcmc_mala_run_batched(pmcmc_ctx* ctx, const pmcmc_target* tgt, int prec, double eps, const void* X0,
                           const void* Xi, const void* U, void* S_out, uint8_t* accepted, int64_t D, int64_t N,
                           int64_t K);

/* ---- DEER Newton solve ----------------------------------------------------
 * Replaces DEER.solve (src/DEER/DEER.jl:506-595) on the taped MALA
 * recursion built by _build_mala_deer_rec (src/interface.jl:360-473):
 * S starts at s0 in every column (or at `init`), each iteration is one
 * deer_update! (:375-475), metric = max|S_new-S| / (tol_abs + tol_rel
 * max|S_new|), stop at metric <= 1 or maxiter.  Probes are generated on the
 * device from (seed, iteration, probe, t*D+d) instead of the host RNG.     */
typedef struct pmcmc_deer_params {
    double eps;
    double tol_abs;
    double tol_rel;
    double damping; /* in (0, 1] */
    int32_t maxiter;
    int32_t jacobian; /* PMCMC_JAC_* */
    int32_t probes;   /* >= 1; used by PMCMC_JAC_STOCH_DIAG */
    int32_t precision; /* PMCMC_PREC_* */
    uint64_t seed;    /* probe stream */
    int32_t fixed_probes; /* 0: new probes every iteration (src/DEER/DEER.jl:407);
                             1: one probe set for the whole solve (the batched
                             solve always uses one set per window) */
} pmcmc_deer_params;

typedef struct pmcmc_deer_info {
    int32_t converged;
    int32_t iters;
    double metric; /* metric of the last iteration */
} pmcmc_deer_info;

/* Workspace for chains of shape D x T at one precision (DEERWorkspace,
 * src/DEER/DEER.jl:47-90).  Reused across solves of the same shape.        */
int pmcmc_deer_ws_create(pmcmc_ctx* ctx, int64_t D, int64_t T, int precision,
                         pmcmc_deer_ws** out);
int pmcmc_deer_ws_destroy(pmcmc_deer_ws* ws);

/* x0: D; Xi: D x T; U: T; init: D x T or NULL; S_out: D x T (may not alias
 * inputs); accepted_out: T bytes or NULL (hard accept mask of the final
 * iteration); info: HOST pointer or NULL.  Synchronises the context stream. */
int pmcmc_deer_solve(pmcmc_ctx* ctx, pmcmc_deer_ws* ws, const pmcmc_target* tgt,
                     const pmcmc_deer_params* params, const void* x0, const void* Xi,
                     const void* U, const void* init, void* S_out, uint8_t* accepted_out,
                     pmcmc_deer_info* info);

/* One block's MALA tape drawn on the device (opt-in; replaces the host draw
 * and upload of _make_mala_tape_block, src/interface.jl:334-358, with another
 * stream): Xi (D x T column-major) ~ N(0, I) and U (T) ~ U(0, 1) in the
 * precision's state type (FP64 for PMCMC_PREC_F64, else FP32), a pure
 * function of `seed`: Philox4x64-10 with key (seed, 0) for the noise (Box-
 * Muller pairs) and (seed, 1) for the uniforms, counter = call index.
 * Synchronises the context stream.                                          */
int pmcmc_tape_generate(pmcmc_ctx* ctx, uint64_t seed, int precision, void* Xi, void* U, int64_t D, int64_t T);

/* The trajectory and its log-densities from the final Newton iteration of
 * the workspace's last pmcmc_deer_solve on a logistic or dense Gaussian
 * target (replaces the separate evaluation of _trajectory_logps,
 * src/interface.jl:517-524, on the block's trajectory): S_out (D x T, the
 * workspace precision) = FT(S_k), the taped step applied to that iteration's
 * iterate S_k (Y_t where step t accepted, else the state before t; within the
 * solve's tolerance of the returned iterate and exactly the next Picard
 * step), logp_out (T doubles, may be NULL) = log p(S_out[:, t]) as that
 * iteration's passes evaluated it.  PMCMC_ERR_ARGUMENT when the last solve
 * was on another target.  Synchronises the context stream.                 */
int pmcmc_deer_last_trajectory(pmcmc_deer_ws* ws, void* S_out, double* logp_out);

/* Trajectory log-densities logp(S[:,t]) (src/interface.jl:517-524),
 * written as double.                                                        */
int pmcmc_target_logdensity_batch(pmcmc_ctx* ctx, const pmcmc_target* tgt, int prec,
                                  const void* S, double* out, int64_t T);

/* Adjoint of a converged trajectory S (decisions `accepted` held fixed): given
 * G = dL/dS (D x T), the gradients of the statistic L with respect to eps and
 * the target's hyperparameter (the logistic prior scale; 0 for the Gaussian
 * targets) and dL/dx0.  Solves the reverse recurrence
 * lambda_t = G_t + J_{t+1}^T lambda_{t+1} by reverse-time DEER sweeps (batched
 * HVPs + diagonal scan) to params->tol_abs / tol_rel within params->maxiter;
 * params->seed picks the Hutchinson probe.  No reference counterpart: the
 * reference differentiates sequential MALA with Enzyme (ext/EnzymeExt.jl);
 * this replaces that for the DEER path.  grad_out: HOST double[2];
 * dx0_out: D (device, or NULL); lambda_out: D x T (device, or NULL).
 * Synchronises the context stream.                                         */
int pmcmc_deer_adjoint(pmcmc_ctx* ctx, pmcmc_deer_ws* ws, const pmcmc_target* tgt,
                       const pmcmc_deer_params* params, const void* x0, const void* Xi,
                       const void* S, const uint8_t* accepted, const void* G, double* grad_out,
                       void* dx0_out, void* lambda_out, pmcmc_deer_info* info);

/* ---- autotuner -------------------------------------------------------------
 * No reference counterpart.  Level-0 tile shapes of the diagonal scan and of
 * the fused DEER tile kernels (tc columns per tile, `parts` parts per row in
 * the in-tile scan, and for the standalone scan the single-pass look-back
 * form) are looked up in a process-wide table keyed by (kind, precision, D,
 * ceil(log2 T)); without an entry (or with one that does not fit) a fixed
 * rule applies.  single_pass: 0 two-pass reduce/apply, 1 LDS-tiled
 * decoupled look-back, 2 register-resident tiles with look-back
 * (16 <= D <= 256, as row blocks of at most 64 rows; tc = 64, 128 or 256
 * columns, 256 FP32 only; parts unused), 3 the same with persistent
 * workgroups that load the next tile during a tile's look-back (tc = 64 or
 * 128).
 * pmcmc_scan_tune times candidate shapes of the standalone
 * diagonal scan on synthetic operands; pmcmc_deer_tune times the
 * shape-dependent kernels (tile reduce, spine, tile apply) of `iters`
 * Newton iterations of the caller's problem on `ws` (tolerances ignored), and
 * re-plans ws with the winner.  Both record the winner in the table (used by
 * later scans and workspaces; batched solves keep the rule's shapes, which
 * measured faster there) and return it; `us` is its measured
 * time per scan / per iteration.  get/set/count/entry let a host save and
 * restore the table (set rejects a shape that cannot run: UNSUPPORTED);
 * get returns tc = 0 when there is no entry.                               */
#define PMCMC_TUNE_SCAN_DIAG 0
#define PMCMC_TUNE_DEER 1
typedef struct pmcmc_tile_shape {
    int32_t tc;
    int32_t parts;
    int32_t single_pass;
    int32_t pad;
    double us;
} pmcmc_tile_shape;
int pmcmc_scan_tune(pmcmc_ctx* ctx, int prec, int64_t D, int64_t T, int32_t reps,
                    pmcmc_tile_shape* best);
int pmcmc_deer_tune(pmcmc_ctx* ctx, pmcmc_deer_ws* ws, const pmcmc_target* tgt,
                    const pmcmc_deer_params* params, const void* x0, const void* Xi, const void* U,
                    int32_t iters, pmcmc_tile_shape* best);
int pmcmc_tune_get(int kind, int prec, int64_t D, int64_t T, pmcmc_tile_shape* out);
int pmcmc_tune_set(int kind, int prec, int64_t D, int64_t T, const pmcmc_tile_shape* shape);
int pmcmc_tune_clear(void);
int64_t pmcmc_tune_count(void);
int pmcmc_tune_entry(int64_t i, int* kind, int* prec, int64_t* D, int32_t* log2T,
                     pmcmc_tile_shape* out);

/* ---- many chains in one solve ---------------------------------------------
 * Chains of uneven D and T (logistic or dense-Gaussian targets, D <= 256).
 * Each chain iterates over its own sliding window of at most `window`
 * columns: a window kernel at the start of every iteration places it from
 * the chain's front (its first column that is not final), and the apply
 * pass's last workgroup moves the front past the final prefix (tiles whose
 * residual FT - S is exactly 0), or past the whole window once it meets the
 * tolerance.  All chains' Newton iterations run as one set of batched
 * launches (work items over per-chain descriptor tables; each chain stops
 * at the end of its trajectory).  Between chunks of iterations the host
 * picks each chain's window length (two tiles while its front moves about
 * one column per iteration, wider when it moves faster) and rebuilds the
 * work items; the iterations in between are replayed from a captured
 * hipGraph (an iteration pair).  params: jacobian = :stoch_diag with one
 * probe, precision F32 or BF16_JAC, maxiter per window's worth of columns;
 * eps per chain.
 * The batched counterpart of running src/interface.jl:475-515 once per
 * chain and block; the reference has no batched or windowed solve.
 * All pointers in a descriptor are device pointers (FP32); S_out: D x T. */
typedef struct pmcmc_chain_desc {
    const pmcmc_target* target;
    int64_t T;
    double eps;
    const void* x0;
    const void* Xi;
    const void* U;
    void* S_out;
    uint8_t* accepted_out; /* T bytes or NULL */
} pmcmc_chain_desc;

typedef struct pmcmc_batch pmcmc_batch;

int pmcmc_batch_create(pmcmc_ctx* ctx, int32_t n, const pmcmc_chain_desc* chains,
                       const pmcmc_deer_params* params, int64_t window, pmcmc_batch** out);
/* 0: launch the batched kernels directly instead of replaying graphs */
int pmcmc_batch_set_graph(pmcmc_batch* b, int32_t on);
/* infos: n HOST records or NULL (iters: Newton iterations of the chain's
 * whole trajectory).  Synchronises. */
int pmcmc_batch_solve(pmcmc_batch* b, pmcmc_deer_info* infos);
/* Measurement of the dense-Gaussian Lambda passes (no reference counterpart):
 * with profiling on, solves launch directly (no graphs), record HIP events
 * around every Lambda-pass (and GEMM-pass) launch and count the passes' algorithmic flops
 * (2 x rows x D x columns, states and directions) on the device;
 * pmcmc_batch_pass_stats returns the last profiled solve's launches, their
 * summed duration and the flops. */
int pmcmc_batch_set_profiling(pmcmc_batch* b, int32_t on);
int pmcmc_batch_pass_stats(const pmcmc_batch* b, int64_t* launches, double* total_ms, double* flops);
/* The same for the logistic chains' GEMM passes (no reference counterpart;
 * they replace bayes_logreg.jl's gradient / HVP GEMMs): launches, summed
 * duration and algorithmic flops (8 x rows x D x columns: two GEMMs of the
 * states and of the directions) of the last profiled solve. */
int pmcmc_batch_glm_pass_stats(const pmcmc_batch* b, int64_t* launches, double* total_ms, double* flops);
int pmcmc_batch_destroy(pmcmc_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* PMCMC_AMD_H */
