/* fagra_gpu.h -- C ABI of the MI355X (gfx950) bundle-adjustment backend.
 *
 * The library is a drop-in for one path of oleflb/fagra: the LM/GN bundle-adjustment
 * inner loop. It mirrors three reference interfaces:
 *
 *   LeastSquaresBackend (src/optimization.rs:131-169)  -> fg_schur_*   (GPU Schur-PCG)
 *   FactorBatch for the BAL Frame (examples/bal/model.rs:194-237) -> fg_ba_*  (device evaluator)
 *   LevenbergMarquardt / GaussNewton solve_batch
 *     (src/optimization/levenberg_marquardt.rs:252-262, gauss_newton.rs:63-74) -> fg_lm_*, fg_gn_*
 *
 * Every call returns an fg_status whose values correspond one-to-one to the reference's
 * SolverError / EvaluationError variants (src/error.rs:21-76). Device pointers never cross
 * the ABI: arrays are host memory in the reference's coordinate order (camera i at
 * [9i, 9i+9), point j at [9nc+3j, 9nc+3j+3), Jacobian rows of observation k at [2k, 2k+2)
 * with k the camera-stable row order of examples/bal.rs:189-201). One HIP stream per
 * backend; not thread-safe (the reference is &mut self, src/schur.rs:527).
 */
#ifndef FAGRA_GPU_H
#define FAGRA_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/error.rs:21-76 */
typedef enum {
  FG_OK = 0,
  FG_INVALID_OPTIONS = 1,        /* SolverError::InvalidOptions */
  FG_LINEAR_SOLVE_FAILED = 2,    /* SolverError::LinearSolveFailed */
  FG_NO_CONVERGENCE = 3,         /* SolverError::NoConvergence */
  FG_NO_PROGRESS = 4,            /* SolverError::NoProgress */
  FG_INVALID_EVALUATION = 5,     /* EvaluationError::InvalidEvaluation */
  FG_DIMENSION_MISMATCH = 6,     /* EvaluationError::DimensionMismatch */
  FG_INVALID_EMISSION = 7,       /* EvaluationError::InvalidEmission */
  FG_UNSUPPORTED_STRUCTURE = 8,  /* EvaluationError::UnsupportedStructure */
  FG_INVALID_INPUT = 9,          /* BAL text rejected (Problem::parse's Err(String)) */
  FG_DUPLICATE_COVARIANCE_BLOCK = 10, /* SolverError::DuplicateCovarianceBlock */
  FG_SINGULAR_INFORMATION = 11,  /* SolverError::SingularInformation */
  FG_INVALID_INFORMATION = 12,   /* SolverError::InvalidInformation */
  FG_INVALID_RANK_TOLERANCE = 13, /* SolverError::InvalidRankTolerance */
  FG_UNKNOWN_KEY = 14,           /* KeyError::Unknown (a selection that is not a variable block) */
  FG_DEVICE_ERROR = 100          /* HIP runtime failure; no reference counterpart */
} fg_status;

/* Arithmetic of the bulk path. FG_FP32: Jacobians, Schur products and PCG matvecs in FP32,
 * with FP64 cost sums, gain ratio, damping and PCG dot products/norms. */
typedef enum { FG_FP64 = 0, FG_FP32 = 1 } fg_precision;

/* Reduced-camera PCG preconditioner. BLOCK_JACOBI is the reference's (B_i + lambda I)^-1
 * (src/schur.rs:658-682). LEARNED applies L^-T G G^T L^-1 with G from the per-camera MLP
 * and falls back to BLOCK_JACOBI when the residual does not fall by fallback_factor within
 * fallback_iterations. */
typedef enum { FG_PRECOND_BLOCK_JACOBI = 0, FG_PRECOND_LEARNED = 1 } fg_preconditioner;

/* src/schur.rs:32-41 */
typedef enum {
  FG_CG_NONE = -1,
  FG_CG_CONVERGED = 0,
  FG_CG_ITERATION_LIMIT = 1,
  FG_CG_NOT_POSITIVE_DEFINITE = 2,
  FG_CG_NON_FINITE = 3
} fg_schur_termination;

/* SchurDiagnostics (src/schur.rs:44-60); NaN encodes None. */
typedef struct {
  int termination;
  size_t iterations;
  double damping;
  double estimated_relative_residual;
  double reduced_relative_residual;
  double scaled_relative_residual;
  double original_relative_residual;
  int used_fallback;            /* learned preconditioner replaced by block-Jacobi */
} fg_schur_diagnostics;

/* SchurTimings (src/schur.rs:77-91), seconds, from HIP events on the backend stream. */
typedef struct {
  double preparation, factorization, linear_solve, products, preconditioner, verification;
  double linearization, cost;   /* evaluator time; not part of the reference struct */
  size_t matvecs;               /* reduced-system products, including verification */
  double matvec_kernel_seconds; /* summed duration of the two Schur matvec kernels */
} fg_schur_timings;

typedef void (*fg_on_solve_fn)(const fg_schur_diagnostics*, void* user);

/* Schur public fields (src/schur.rs:116-125) plus the device choices. */
typedef struct {
  size_t max_iterations;        /* default 1000 (src/schur.rs:155) */
  double relative_tolerance;    /* default max(1e-10, 128 eps) of the precision (:156-157) */
  int collect_timings;          /* default 0 */
  int precision;                /* fg_precision */
  int preconditioner;           /* fg_preconditioner */
  double fallback_factor;       /* learned: required residual reduction (default 0.2) ... */
  size_t fallback_iterations;   /* ... within this many iterations (default 8) */
  fg_on_solve_fn on_solve;      /* observe every numerically attempted solve */
  void* on_solve_user;
} fg_schur_config;

typedef struct fg_schur fg_schur;  /* GPU Schur backend (src/schur.rs:116) */
typedef struct fg_ba fg_ba;        /* device-resident BAL problem and batch evaluator */

/* ---- batch evaluator: BAL Frame batches (examples/bal/model.rs:194-237) ---- */

/* Upload a problem: cameras nc x 9 [angle-axis, t, f, k1, k2], points np x 3, observations
 * (camera, point, pixel). huber > 0 selects blockwise Huber with that threshold via IRLS
 * (src/factors.rs:42-52); 0 is plain least squares. Builds the camera-stable row order,
 * camera offsets and point permutation on the host (examples/bal.rs:176-201). */
fg_status fg_ba_create(size_t nc, size_t np, size_t no, const double* cameras, const double* points,
                       const int64_t* obs_camera, const int64_t* obs_point, const double* obs_pixel,
                       double huber, int precision, fg_ba** out);
void fg_ba_destroy(fg_ba*);
/* Bit-exact layout: camera-stable row order (no entries: original observation index of row
 * k), camera offsets (nc+1), point permutation of rows (no) and point offsets (np+1). */
fg_status fg_ba_layout(const fg_ba*, int64_t* cam_order, int64_t* cam_offsets, int64_t* pt_perm,
                       int64_t* pt_offsets);
fg_status fg_ba_get_state(fg_ba*, double* cameras, double* points);
fg_status fg_ba_set_state(fg_ba*, const double* cameras, const double* points);
/* Total cost of the accepted (trial=0) or staged (trial=1) state: sum over Frame batches of
 * 0.5||r||^2 or the Huber cost (Frame::cost, model.rs:201-213; checked_cost storage.rs:612). */
fg_status fg_ba_cost(fg_ba*, int trial, double* out);
/* StatePool::stage (src/storage.rs:288-306) with the additive BAL retraction. */
fg_status fg_ba_stage(fg_ba*, const double* delta);
void fg_ba_accept(fg_ba*);  /* StatePool::accept_trial (storage.rs:308-313) */
void fg_ba_reject(fg_ba*);  /* StatePool::reject_trial (storage.rs:315-318) */

/* ---- LeastSquaresBackend mirror (src/optimization.rs:131-169, src/schur.rs:400-546) ---- */

fg_status fg_schur_config_default(fg_schur_config* cfg, int precision);
fg_status fg_schur_create(const fg_schur_config*, fg_schur** out);
void fg_schur_destroy(fg_schur*);
fg_status fg_schur_set_config(fg_schur*, const fg_schur_config*);
/* prepare(dimension): validates the 9/3 partition for the bound problem (schur.rs:409-431). */
fg_status fg_schur_prepare(fg_schur*, fg_ba*, size_t dimension);
void fg_schur_clear(fg_schur*);
/* Bulk replacement of Frame::linearize + Schur::accumulate for every observation of the
 * bound problem at its accepted state (model.rs:214-236, schur.rs:438-482, with the sink's
 * finiteness checks of optimization.rs:556,597). */
fg_status fg_schur_linearize(fg_schur*);
fg_status fg_schur_gradient_norm(fg_schur*, double* out);
fg_status fg_schur_prepare_damping(fg_schur*, double min_column_norm);
/* delta (host, 9nc+3np) stays valid until the next mutating call. */
fg_status fg_schur_solve_damped(fg_schur*, double lambda, const double** delta, double* predicted);
fg_status fg_schur_solve(fg_schur*, const double** delta);
void fg_schur_diagnostics_get(const fg_schur*, fg_schur_diagnostics*);
void fg_schur_statistics(const fg_schur*, size_t* solves, size_t* iterations);
void fg_schur_timings_get(const fg_schur*, fg_schur_timings*);

/* Model and operator access for parity checks (host buffers, reference order). */
fg_status fg_schur_model(fg_schur*, double* gradient, double* scales, double* negative_residual);
fg_status fg_schur_camera_gram(fg_schur*, double* gram /* nc x 81, row-major full */);
/* Reduced::apply (src/schur.rs:606-652): y = S x on the 9nc retained coordinates. */
fg_status fg_schur_reduced_apply(fg_schur*, double lambda, const double* x, double* y);
/* Measurement hook, no reference counterpart: `repeats` back-to-back launches of each of the two
 * Schur-product kernels (point-major elimination, camera-major accumulation) for the scaled
 * camera vector x at lambda, timed with HIP events on the backend stream; mean seconds per
 * launch of each. */
fg_status fg_schur_time_products(fg_schur*, double lambda, const double* x, int repeats, double* sec_points,
                                 double* sec_cameras);
/* Consistency hook, no reference counterpart: after a linearization, the number of
 * observations whose camera-major record or weighted residual (formed by the camera-major
 * pass) differs in any bit from the point-major copy; 0 when the two passes agree, -1 when not
 * linearized or on a device error. */
int fg_schur_record_mismatches(fg_schur*);
/* Jacobian::apply / transpose_apply (src/linear_model.rs:39-87). */
fg_status fg_schur_jacobian_apply(fg_schur*, const double* x, double* y);
fg_status fg_schur_jacobian_transpose_apply(fg_schur*, const double* y, double* x);
/* Damped::apply / transpose_apply (src/lsmr.rs:147-205). */
fg_status fg_schur_damped_apply(fg_schur*, double lambda, int scaled, const double* x, double* y);
fg_status fg_schur_damped_transpose_apply(fg_schur*, double lambda, int scaled, const double* y,
                                          double* x);

/* ---- learned preconditioner (no reference counterpart: the reference's Schur uses plain
 *      block-Jacobi, src/schur.rs:658-682) ---- */

/* Per-camera network 47 -> 128 -> 128 -> 45 (GELU), BF16 operands, FP32 accumulation. */
typedef struct fg_mlp fg_mlp;
size_t fg_mlp_param_count(void);
fg_status fg_mlp_create(uint64_t seed, fg_mlp** out);
void fg_mlp_destroy(fg_mlp*);
fg_status fg_mlp_set_params(fg_mlp*, const float* params);
fg_status fg_mlp_get_params(fg_mlp*, float* params);
/* Input range of the trained network (lo, hi: 47 feature bounds each): a camera whose features
 * leave it -- a damping or a Gram block unlike any the network was trained on -- keeps plain
 * block-Jacobi (G = I). Null lo and hi clear the range. */
fg_status fg_mlp_set_input_range(fg_mlp*, const float* lo, const float* hi);
/* out (n x 45) = network(features n x 47) */
fg_status fg_mlp_forward(fg_mlp*, size_t n, const float* features, float* out);
/* Mean loss ||G^T A G - I||_F^2 over n rows (targets n x 45 packed lower) and its gradient. */
fg_status fg_mlp_loss_grad(fg_mlp*, size_t n, const float* features, const float* targets, double* loss,
                           float* grads);
/* Adam epochs on the GPU; epoch_loss receives the mean batch loss of each epoch. */
fg_status fg_mlp_train(fg_mlp*, size_t n, const float* features, const float* targets, size_t epochs,
                       size_t batch, double lr, uint64_t seed, double* epoch_loss);
/* Attach a network to a backend for FG_PRECOND_LEARNED. */
fg_status fg_schur_set_mlp(fg_schur*, fg_mlp*);
/* Training data of the prepared model at lambda: per camera the network features (nc x 47),
 * the target A = L^-1 S_cc L^-T (nc x 45 packed) and the observation count. */
fg_status fg_schur_capture(fg_schur*, double lambda, float* features, float* targets, int* counts);

/* ---- LSMR backend (the reference's second drop-in, src/lsmr.rs:301-595) ----
 * Same device model and kernels as the Schur backend; the LSMR iteration (faer 0.24.4's
 * matrix_free::lsmr, Fong & Saunders 2011) runs on the device with its scalar recurrences in
 * device state. Vectors are f64; the Jacobian products use the backend's precision. */
typedef struct fg_lsmr fg_lsmr;
/* LsmrTermination (src/lsmr.rs:60-68) */
typedef enum {
  FG_LSMR_NONE = -1,
  FG_LSMR_CONVERGED = 0,
  FG_LSMR_ITERATION_LIMIT = 1,
  FG_LSMR_NON_FINITE = 2
} fg_lsmr_termination;
/* LsmrDiagnostics (src/lsmr.rs:72-105); NaN = None. */
typedef struct {
  int termination; /* fg_lsmr_termination */
  size_t iterations;
  int finite_step;
  double estimated_absolute_residual, estimated_relative_residual;
  double original_absolute_residual, original_relative_residual;
  double scaled_absolute_residual, scaled_relative_residual;
  double min_scale, max_scale, damping;
  size_t block_fallbacks;
} fg_lsmr_diagnostics;
/* LsmrTimings (src/lsmr.rs:40-57), seconds. */
typedef struct {
  double preparation, factorization, linear_solve, forward, transpose, preconditioner, residual_checks;
} fg_lsmr_timings;
typedef void (*fg_lsmr_observer_fn)(const fg_lsmr_diagnostics*, void* user);
/* Lsmr public fields (src/lsmr.rs:273-295) + precision. */
typedef struct {
  size_t max_iterations;          /* default 1000 */
  double relative_tolerance;      /* default max(1e-10, 128 eps) */
  int diagonal_preconditioning;   /* default 1 */
  int block_preconditioning;      /* default 0 */
  int collect_timings;            /* default 0 */
  int precision;                  /* fg_precision of the Jacobian products */
  fg_lsmr_observer_fn on_solve;
  void* on_solve_user;
} fg_lsmr_config;

fg_status fg_lsmr_config_default(fg_lsmr_config*, int precision);
fg_status fg_lsmr_create(const fg_lsmr_config*, fg_lsmr** out);
void fg_lsmr_destroy(fg_lsmr*);
fg_status fg_lsmr_set_config(fg_lsmr*, const fg_lsmr_config*);
/* LeastSquaresBackend for Lsmr (src/lsmr.rs:471-595) over a device problem. */
fg_status fg_lsmr_prepare(fg_lsmr*, fg_ba*, size_t dimension);
void fg_lsmr_clear(fg_lsmr*);
fg_status fg_lsmr_linearize(fg_lsmr*);
fg_status fg_lsmr_gradient_norm(fg_lsmr*, double* out);
fg_status fg_lsmr_prepare_damping(fg_lsmr*, double min_column_norm);
fg_status fg_lsmr_solve_damped(fg_lsmr*, double lambda, const double** delta, double* predicted);
fg_status fg_lsmr_solve(fg_lsmr*, const double** delta);
void fg_lsmr_diagnostics_get(const fg_lsmr*, fg_lsmr_diagnostics*);
void fg_lsmr_statistics(const fg_lsmr*, size_t* solves, size_t* iterations);
void fg_lsmr_timings_get(const fg_lsmr*, fg_lsmr_timings*);

/* ---- nonlinear optimizers ---- */

/* OptimizeOptions (src/optimization.rs:39-67) */
typedef struct {
  size_t max_iterations;
  double gradient_tolerance, step_tolerance, cost_tolerance;
} fg_optimize_options;
/* LmOptions (src/optimization/levenberg_marquardt.rs:17-46) */
typedef struct {
  double initial_damping, min_damping, max_damping, min_column_norm, acceptance_threshold;
  size_t max_trials;
} fg_lm_options;
/* TerminationReason (src/optimization.rs:89-98); TrialFailure (levenberg_marquardt.rs:71-80) */
enum { FG_TERM_NONE = -1, FG_TERM_NO_VARIABLES = 0, FG_TERM_GRADIENT = 1, FG_TERM_STEP = 2,
       FG_TERM_COST = 3 };
enum { FG_REJ_NONE = -1, FG_REJ_LINEAR_SOLVE = 0, FG_REJ_PREDICTION = 1,
       FG_REJ_INVALID_EVALUATION = 2, FG_REJ_POOR_AGREEMENT = 3 };
/* LmOutcome = result + LmStatistics + LmWork (levenberg_marquardt.rs:84-184); NaN = None. */
typedef struct {
  int status;
  int termination;
  size_t iterations;
  double initial_cost, final_cost;
  size_t linearizations, attempts, rejected_steps, accepted_steps, cost_evaluations;
  double damping, cost, gradient_norm, step_norm;
  int last_rejection;
  size_t linear_solves, linear_iterations;
} fg_lm_outcome;

typedef void (*fg_on_accept_fn)(const fg_lm_outcome* statistics, void* user);

fg_status fg_optimize_options_default(fg_optimize_options*);
fg_status fg_lm_options_default(fg_lm_options*);

/* LevenbergMarquardt::<Schur>::solve_batch_with_outcome on the device problem; the problem's
 * state holds the accepted estimates afterwards, also on error. */
fg_status fg_lm_solve_batch(fg_schur*, fg_ba*, const fg_optimize_options*, const fg_lm_options*,
                            fg_on_accept_fn on_accept, void* user, fg_lm_outcome* out);
/* GaussNewton::<Schur>::solve_batch on the device problem. */
fg_status fg_gn_solve_batch(fg_schur*, fg_ba*, const fg_optimize_options*, fg_lm_outcome* out);
/* The same loops with the LSMR backend (LevenbergMarquardt::<Lsmr> / GaussNewton::<Lsmr>). */
fg_status fg_lm_solve_batch_lsmr(fg_lsmr*, fg_ba*, const fg_optimize_options*, const fg_lm_options*,
                                 fg_on_accept_fn on_accept, void* user, fg_lm_outcome* out);
fg_status fg_gn_solve_batch_lsmr(fg_lsmr*, fg_ba*, const fg_optimize_options*, fg_lm_outcome* out);

/* The same LM/GN loops over caller-supplied state and backend callbacks. */
typedef struct {
  void* ctx;
  size_t dimension;
  int (*prepare)(void* ctx);
  int (*cost)(void* ctx, int trial, double* out);
  int (*linearize)(void* ctx);
  int (*gradient_norm)(void* ctx, double* out);
  int (*prepare_damping)(void* ctx, double min_column_norm);
  /* step_norm = infinity norm of the step, NaN when it is nonfinite */
  int (*solve_damped)(void* ctx, double lambda, double* step_norm, double* predicted);
  int (*solve)(void* ctx, double* step_norm);
  int (*stage)(void* ctx);  /* stage the last solved step */
  void (*accept)(void* ctx);
  void (*reject)(void* ctx);
  void (*linear_work)(void* ctx, size_t* solves, size_t* iterations);
} fg_problem;
void fg_lm_optimize(const fg_problem*, const fg_optimize_options*, const fg_lm_options*,
                    fg_on_accept_fn on_accept, void* user, fg_lm_outcome* out);
void fg_gn_optimize(const fg_problem*, const fg_optimize_options*, fg_lm_outcome* out);

/* ---- Schur backend for emissions of any factor family ----
 * Schur::new(retained, retained_width, eliminated_width) (src/schur.rs:116-178) over explicit
 * Jacobian blocks that the caller's own model emits on the host (its FactorBatch::linearize),
 * accumulated in bulk; the damped systems are solved on the device (f64, block-Jacobi PCG).
 * Block widths up to 16 (wider: FG_UNSUPPORTED_STRUCTURE at create). The config must be FP64
 * with block-Jacobi preconditioning. */
typedef struct fg_gschur fg_gschur;
fg_status fg_gschur_create(const fg_schur_config*, size_t retained, size_t retained_width,
                           size_t eliminated_width, fg_gschur** out);
void fg_gschur_destroy(fg_gschur*);
fg_status fg_gschur_set_config(fg_gschur*, const fg_schur_config*);
/* LeastSquaresBackend::prepare / clear (src/schur.rs:409-437) */
fg_status fg_gschur_prepare(fg_gschur*, size_t dimension);
void fg_gschur_clear(fg_gschur*);
/* LeastSquaresBackend::accumulate (src/schur.rs:438-482, src/linear_model.rs:180-210) for n
 * emissions of `rows` rows and n_blocks blocks each: residual[n*rows], cols[n*n_blocks] (block
 * column offsets), widths[n_blocks], values: per emission, its blocks in order, each
 * column-major rows x width. The batch is validated (DimensionMismatch, UnsupportedStructure)
 * before any of it is appended. */
fg_status fg_gschur_accumulate(fg_gschur*, size_t n, size_t rows, const double* residual, size_t n_blocks,
                               const size_t* cols, const size_t* widths, const double* values);
fg_status fg_gschur_gradient_norm(fg_gschur*, double* out);
fg_status fg_gschur_prepare_damping(fg_gschur*, double min_column_norm);
fg_status fg_gschur_solve_damped(fg_gschur*, double lambda, const double** delta, double* predicted);
fg_status fg_gschur_solve(fg_gschur*, const double** delta);
void fg_gschur_diagnostics_get(const fg_gschur*, fg_schur_diagnostics*);
void fg_gschur_statistics(const fg_gschur*, size_t* solves, size_t* iterations);
/* infinity norm of the last solved step (NaN when nonfinite), for fg_problem.solve_damped */
double fg_gschur_last_step_norm(const fg_gschur*);

/* ---- selected covariance (src/covariance.rs) ---- */
/* CovarianceOptions (src/covariance.rs:136-160): default 64 epsilon, finite and in [0, 1). */
typedef struct {
  double relative_pivot_tolerance;
} fg_covariance_options;
fg_status fg_covariance_options_default(fg_covariance_options*);
/* CovarianceBackend for DenseNormalCholesky (src/covariance.rs:275-283): the inverse of a dense
 * information matrix (n x n column-major, lower triangle read) at the selected coordinates,
 * k x k column-major into out, factored on the device with the reference's rank screen
 * (src/normal.rs:372-392). */
fg_status fg_dense_covariance(const double* information, size_t n, const size_t* columns, size_t k,
                              const fg_covariance_options*, double* out);
/* The same as fg_dense_covariance without allocating: no reference counterpart. For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The staging copy of the information and the workspace live
 * in a file-scope device array that each GPU holds from its first call until the process
 * ends (about 39 MB); the call makes no hipMalloc / hipFree and creates no stream (it runs on
 * the null stream). Limits: n <= 2048 and k <= 256, else FG_INVALID_OPTIONS (before any
 * device work). Every call on a GPU reuses that GPU's array: calls that overlap on one GPU
 * (from several host threads) race. Results are those of fg_dense_covariance. */
fg_status fg_dense_covariance_noalloc(const double* information, size_t n, const size_t* columns, size_t k,
                                      const fg_covariance_options*, double* out);
/* CovarianceBackend for Schur (src/schur.rs:548-556): undamped information J^T J of the current
 * linearization, selected variable blocks given by (offset, width) in reference coordinates
 * (cameras 9c/9, points 9 nc + 3p/3; generic: the block grid of the partition), in selection
 * order; out is (sum of widths)^2 column-major. Dense: n^2 doubles of device memory. */
fg_status fg_schur_covariance(fg_schur*, const size_t* offsets, const size_t* widths, size_t n_blocks,
                              const fg_covariance_options*, double* out);
fg_status fg_gschur_covariance(fg_gschur*, const size_t* offsets, const size_t* widths, size_t n_blocks,
                               const fg_covariance_options*, double* out);

/* ---- BAL problem text (examples/bal/model.rs:59-113) ---- */
typedef struct {
  size_t n_cameras, n_points, n_observations;
} fg_bal_sizes;
/* Problem::parse's validation over the whole text (no values stored): FG_OK with the sizes,
 * or FG_INVALID_INPUT with the reference's message in err. */
fg_status fg_bal_scan(const char* text, size_t len, fg_bal_sizes* sizes, char* err, size_t err_len);
/* Problem::parse into caller arrays sized from fg_bal_scan: cameras 9 n_c, points 3 n_p,
 * observation camera / point indices n_o, pixels 2 n_o, observations in input order. */
fg_status fg_bal_parse(const char* text, size_t len, const fg_bal_sizes* sizes, double* cameras, double* points,
                       int64_t* obs_camera, int64_t* obs_point, double* pixels, char* err, size_t err_len);
/* BAL text that fg_bal_parse reads back bit for bit (shortest round-trip decimals); no
 * reference counterpart (the reference only reads). *written = bytes needed; FG_OK once
 * they fit in cap, FG_DIMENSION_MISMATCH when buf is null or too small, FG_INVALID_OPTIONS
 * for out-of-range indices or nonfinite values. */
fg_status fg_bal_format(const fg_bal_sizes* sizes, const double* cameras, const double* points,
                        const int64_t* obs_camera, const int64_t* obs_point, const double* pixels, char* buf,
                        size_t cap, size_t* written);

/* ---- device information ---- */
fg_status fg_device_info(char* name, size_t name_len, int* cus, size_t* lds_bytes);
const char* fg_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
