/* fg_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU f64 restatement of oleflb/fagra's bundle-adjustment inner loop, used as the
 * parity checker for the HIP path and as the timed CPU baseline. Nothing in the
 * product library links, loads or calls this code; only tests/, the smoke check
 * and bench.py's cpu_baseline leg do.
 *
 * Status codes mirror src/error.rs one-to-one with the product header's fg_status.
 */
#ifndef FG_ORACLE_H
#define FG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_INVALID_OPTIONS = 1,
  ORC_LINEAR_SOLVE_FAILED = 2,
  ORC_NO_CONVERGENCE = 3,
  ORC_NO_PROGRESS = 4,
  ORC_INVALID_EVALUATION = 5,
  ORC_DIMENSION_MISMATCH = 6,
  ORC_INVALID_EMISSION = 7,
  ORC_UNSUPPORTED_STRUCTURE = 8,
  /* SolverError variants of the covariance query (src/error.rs:43-58) and KeyError::Unknown */
  ORC_DUPLICATE_COVARIANCE_BLOCK = 10,
  ORC_SINGULAR_INFORMATION = 11,
  ORC_INVALID_INFORMATION = 12,
  ORC_INVALID_RANK_TOLERANCE = 13,
  ORC_UNKNOWN_KEY = 14
};

/* Schur termination (src/schur.rs:32-41); -1 = none. */
enum { ORC_CG_NONE = -1, ORC_CG_CONVERGED = 0, ORC_CG_ITERATION_LIMIT = 1,
       ORC_CG_NOT_POSITIVE_DEFINITE = 2, ORC_CG_NON_FINITE = 3 };

typedef struct {
  int termination;            /* ORC_CG_* */
  size_t iterations;
  double damping;
  double estimated_relative_residual;   /* NaN = None */
  double reduced_relative_residual;
  double scaled_relative_residual;
  double original_relative_residual;
} orc_schur_diagnostics;

/* LsmrTermination (src/lsmr.rs:60-68); -1 = none. */
enum { ORC_LSMR_NONE = -1, ORC_LSMR_CONVERGED = 0, ORC_LSMR_ITERATION_LIMIT = 1, ORC_LSMR_NON_FINITE = 2 };

/* LsmrDiagnostics (src/lsmr.rs:72-105); NaN = None. */
typedef struct {
  int termination;
  size_t iterations;
  int finite_step;
  double estimated_absolute_residual, estimated_relative_residual;
  double original_absolute_residual, original_relative_residual;
  double scaled_absolute_residual, scaled_relative_residual;
  double min_scale, max_scale, damping;
  size_t block_fallbacks;
} orc_lsmr_diagnostics;

/* ---- projection (examples/bal/model.rs:126-184) ---- */
int orc_projection(const double cam[9], double rotation[9], double right[9]);
int orc_project(const double cam[9], const double point[3], double pixel[2]);
void orc_jacobians(const double cam[9], const double point[3], double jc[18], double jp[6]);

/* ---- generic Schur backend (src/schur.rs) ---- */
typedef struct orc_schur orc_schur;
orc_schur* orc_schur_new(size_t retained, size_t retained_width, size_t eliminated_width);
void orc_schur_free(orc_schur*);
void orc_schur_set_options(orc_schur*, size_t max_iterations, double relative_tolerance);
int orc_schur_prepare(orc_schur*, size_t dimension);
void orc_schur_clear(orc_schur*);
/* One residual block: rows entries of residual, nblocks Jacobian blocks with column
 * offsets cols[i], widths widths[i], column-major values concatenated in values. */
int orc_schur_accumulate(orc_schur*, const double* residual, size_t rows, size_t nblocks,
                         const size_t* cols, const size_t* widths, const double* values);
int orc_schur_gradient_norm(const orc_schur*, double* out);
int orc_schur_prepare_damping(orc_schur*, double min_column_norm);
int orc_schur_solve_damped(orc_schur*, double lambda, double* delta, double* predicted);
int orc_schur_solve(orc_schur*, double* delta);
void orc_schur_diagnostics_get(const orc_schur*, orc_schur_diagnostics*);
void orc_schur_statistics(const orc_schur*, size_t* solves, size_t* iterations);
/* Model internals for parity checks. */
size_t orc_schur_dimension(const orc_schur*);
size_t orc_schur_rows(const orc_schur*);
void orc_schur_gradient(const orc_schur*, double* out);
void orc_schur_scales(const orc_schur*, double* out);
void orc_schur_negative_residual(const orc_schur*, double* out);
/* Reduced::apply (src/schur.rs:606-652) at damping lambda; requires prepare_damping. */
int orc_schur_reduced_apply(const orc_schur*, double lambda, const double* x, double* y);
/* Reduced right-hand side of the last solve, and camera Gram blocks (9x9 lower, col-major). */
void orc_schur_reduced_rhs(const orc_schur*, double* out);
int orc_schur_gram_block(const orc_schur*, size_t col, double* out, size_t* width);
/* Jacobian::apply / transpose_apply (src/linear_model.rs:39-87). */
void orc_schur_jacobian_apply(const orc_schur*, const double* x, double* y);
void orc_schur_jacobian_transpose_apply(const orc_schur*, const double* y, double* x);
/* Damped::apply / transpose_apply (src/lsmr.rs:147-205) with diagonal sqrt(lambda) (scaled)
 * or sqrt(lambda)*D (unscaled); x has n entries, y has rows+n entries. */
void orc_schur_damped_apply(const orc_schur*, double lambda, int scaled, const double* x, double* y);
void orc_schur_damped_transpose_apply(const orc_schur*, double lambda, int scaled, const double* y, double* x);
/* Block factors of the last prepare_damping at lambda (block_factors.rs:113-160). */
size_t orc_schur_factor(orc_schur*, double lambda);
void orc_schur_apply_factor_inverse(const orc_schur*, double* rhs, int transpose);
/* Test hooks: BlockFactors::prepare with explicit scales, Gram poke, prepare_scales alone,
 * and LinearizedModel::verify_step for an explicit step (src/linear_model.rs:254-282). */
int orc_schur_prepare_blocks(orc_schur*, const double* scales);
void orc_schur_gram_set(orc_schur*, size_t index, double value);
int orc_schur_prepare_scales(orc_schur*, double floor);
double orc_schur_verify(orc_schur*, const double* step, double lambda, double* normal_residual,
                        double* reference, double* residual_norm, double* reference_norm);

/* ---- LSMR backend over an orc_schur's linearized model (src/lsmr.rs); mode 0 identity,
 *      1 diagonal scaling, 2 block-Jacobi (blocks prepared by orc_schur_prepare_damping) ---- */
int orc_lsmr_solve_damped(orc_schur*, double lambda, int mode, size_t max_iterations, double relative_tolerance,
                          double* delta, double* predicted, orc_lsmr_diagnostics* diagnostics);
int orc_lsmr_solve(orc_schur*, size_t max_iterations, double relative_tolerance, double* delta,
                   orc_lsmr_diagnostics* diagnostics);

/* ---- BA problem + batch evaluator (examples/bal/model.rs, examples/bal.rs:176-201) ---- */
typedef struct orc_ba orc_ba;
orc_ba* orc_ba_new(size_t nc, size_t np, size_t no, const double* cameras, const double* points,
                   const int64_t* obs_camera, const int64_t* obs_point, const double* obs_pixel,
                   double huber_delta);
void orc_ba_free(orc_ba*);
/* Layout arrays (Appendix A of the build notes): row order, camera offsets, point permutation
 * and offsets. */
void orc_ba_layout(const orc_ba*, int64_t* cam_order, int64_t* cam_offsets, int64_t* pt_perm,
                   int64_t* pt_offsets);
int orc_ba_cost(const orc_ba*, double* out);
void orc_ba_get_state(const orc_ba*, double* cameras, double* points);
void orc_ba_set_state(orc_ba*, const double* cameras, const double* points);
/* Linearize the current state into a Schur backend (clears it first). */
int orc_ba_linearize(orc_ba*, orc_schur*);
/* the same linearization one emission at a time through orc_schur_accumulate */
int orc_ba_linearize_sequential(orc_ba*, orc_schur*);
/* OpenMP threads of the oracle's block loops (results do not depend on it) */
void orc_set_threads(int n);
int orc_get_threads(void);
/* Retraction of a full-length step into the state (additive). */
void orc_ba_retract(orc_ba*, const double* delta);

/* ---- LM / GN (src/optimization/levenberg_marquardt.rs:330-472, gauss_newton.rs:101-176) ---- */
typedef struct {
  size_t max_iterations;
  double gradient_tolerance, step_tolerance, cost_tolerance;
} orc_optimize_options;
typedef struct {
  double initial_damping, min_damping, max_damping, min_column_norm, acceptance_threshold;
  size_t max_trials;
} orc_lm_options;
enum { ORC_TERM_NONE = -1, ORC_TERM_NO_VARIABLES = 0, ORC_TERM_GRADIENT = 1, ORC_TERM_STEP = 2,
       ORC_TERM_COST = 3 };
enum { ORC_REJ_NONE = -1, ORC_REJ_LINEAR_SOLVE = 0, ORC_REJ_PREDICTION = 1,
       ORC_REJ_INVALID_EVALUATION = 2, ORC_REJ_POOR_AGREEMENT = 3 };
typedef struct {
  int status;                  /* ORC_* of the call */
  int termination;             /* ORC_TERM_* when status == OK */
  size_t iterations;           /* report.iterations */
  double initial_cost, final_cost;
  size_t linearizations, attempts, rejected_steps, accepted_steps, cost_evaluations;
  double damping;
  double cost;                 /* NaN = None */
  double gradient_norm;        /* NaN = None */
  double step_norm;            /* NaN = None */
  int last_rejection;          /* ORC_REJ_* */
  size_t linear_solves, linear_iterations;
} orc_lm_outcome;

/* Generic problem callbacks so the LM/GN loops can drive toy factor graphs too. */
typedef struct {
  void* ctx;
  size_t dimension;
  int (*prepare)(void* ctx);                        /* backend.prepare(dimension) */
  int (*cost)(void* ctx, int trial, double* out);   /* accepted (0) or staged (1) state */
  int (*linearize)(void* ctx);
  int (*gradient_norm)(void* ctx, double* out);
  int (*prepare_damping)(void* ctx, double min_column_norm);
  int (*solve_damped)(void* ctx, double lambda, const double** delta, double* predicted);
  int (*solve)(void* ctx, const double** delta);
  int (*stage)(void* ctx, const double* delta);
  void (*accept)(void* ctx);
  void (*reject)(void* ctx);
  void (*linear_work)(void* ctx, size_t* solves, size_t* iterations);
  void (*on_accept)(void* ctx, const orc_lm_outcome* statistics);
} orc_problem;

void orc_lm_optimize(const orc_problem*, const orc_optimize_options*, const orc_lm_options*,
                     orc_lm_outcome*);
void orc_gn_optimize(const orc_problem*, const orc_optimize_options*, orc_lm_outcome*);

/* Selected covariance (src/covariance.rs:163-262, src/normal.rs:341-392): the inverse of the
 * information H (n x n, column-major, lower triangle read) at the selected coordinates, k x k
 * column-major into out. */
int orc_dense_covariance(const double* information, size_t n, const size_t* columns, size_t k, double tolerance,
                         double* out);
/* LinearizedModel::covariance (src/linear_model.rs:122-167) of a Schur model: H = J^T J
 * replayed from the cached blocks; selected variable blocks by (offset, width). */
int orc_schur_covariance(const orc_schur* s, const size_t* offsets, const size_t* widths, size_t nblocks,
                         double tolerance, double* out);

/* BA convenience: LM (or GN) with the Schur backend on a BA problem. The trace receives one
 * record of three doubles per accepted step (cost, attempts so far, damping of the accepted
 * trial), up to trace_capacity records. */
void orc_ba_lm(orc_ba*, size_t max_iterations_cg, double relative_tolerance,
               const orc_optimize_options*, const orc_lm_options*, orc_lm_outcome*,
               double* trace, size_t trace_capacity, size_t* trace_len);
void orc_ba_gn(orc_ba*, size_t max_iterations_cg, double relative_tolerance,
               const orc_optimize_options*, orc_lm_outcome*);

#ifdef __cplusplus
}
#endif
#endif
