// fg_lm.cpp -- Levenberg-Marquardt and Gauss-Newton outer loops over fg_problem callbacks.
//
// Host control only: the numerical work happens behind the callbacks (for bundle adjustment,
// the device evaluator and the GPU Schur backend). Semantics follow
// src/optimization/levenberg_marquardt.rs:330-472 (gain ratio, Nielsen update, doubling retry
// growth reset per linearization, one undamped trial after stagnation, statistics) and
// src/optimization/gauss_newton.rs:101-176. Cost, gain ratio and damping are f64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "../../include/fagra_gpu.h"
#include "fg_internal.h"

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();

// OptimizeOptions::validate, src/optimization.rs:69-84
int validate_optimize(const fg_optimize_options* o) {
  if (o->max_iterations == 0) return FG_INVALID_OPTIONS;
  const double v[3] = {o->gradient_tolerance, o->step_tolerance, o->cost_tolerance};
  for (double x : v)
    if (!std::isfinite(x) || x < 0.0) return FG_INVALID_OPTIONS;
  return FG_OK;
}

// LmOptions::validate, levenberg_marquardt.rs:47-67
int validate_lm(const fg_lm_options* o) {
  if (o->max_trials == 0) return FG_INVALID_OPTIONS;
  const double v[5] = {o->initial_damping, o->min_damping, o->max_damping, o->min_column_norm,
                       o->acceptance_threshold};
  for (double x : v)
    if (!std::isfinite(x) || x <= 0.0) return FG_INVALID_OPTIONS;
  if (o->min_damping > o->initial_damping || o->initial_damping > o->max_damping ||
      o->acceptance_threshold >= 1.0)
    return FG_INVALID_OPTIONS;
  return FG_OK;
}

// checked_cost, src/storage.rs:612-618
int checked(int st, double value) {
  if (st != FG_OK) return st;
  if (!std::isfinite(value) || value < 0.0) return FG_INVALID_EVALUATION;
  return FG_OK;
}

void reset(fg_lm_outcome* out) {
  std::memset(out, 0, sizeof(*out));
  out->status = FG_OK;
  out->termination = FG_TERM_NONE;
  out->initial_cost = kNaN;
  out->final_cost = kNaN;
  out->cost = kNaN;
  out->gradient_norm = kNaN;
  out->step_norm = kNaN;
  out->last_rejection = FG_REJ_NONE;
}

}  // namespace

extern "C" {

fg_status fg_optimize_options_default(fg_optimize_options* o) {
  // src/optimization.rs:57-67 for f64
  o->max_iterations = 50;
  o->gradient_tolerance = 1e-8;
  o->step_tolerance = 1e-10;
  o->cost_tolerance = 1e-12;
  return FG_OK;
}

fg_status fg_lm_options_default(fg_lm_options* o) {
  // levenberg_marquardt.rs:35-46
  o->initial_damping = 1e-3;
  o->min_damping = 1e-12;
  o->max_damping = 1e12;
  o->min_column_norm = 1e-6;
  o->acceptance_threshold = 1e-4;
  o->max_trials = 16;
  return FG_OK;
}

void fg_lm_optimize(const fg_problem* pb, const fg_optimize_options* options, const fg_lm_options* lm,
                    fg_on_accept_fn on_accept, void* user, fg_lm_outcome* out) {
  fg::lm_optimize_hinted(pb, options, lm, on_accept, user, out, nullptr);
}

}  // extern "C"

// The LM loop with an optional hint: right before each linearization, `hint` receives the
// damping the first trial of that iteration will use (0 for an undamped trial), so a backend
// can start that solve while the host still checks the gradient.
void fg::lm_optimize_hinted(const fg_problem* pb, const fg_optimize_options* options, const fg_lm_options* lm,
                            fg_on_accept_fn on_accept, void* user, fg_lm_outcome* out,
                            void (*hint)(void* ctx, double damping)) {
  reset(out);
  auto finish = [&](int status) {
    out->status = status;
    // LmOutcome work: zero linear work when no attempt was made (levenberg_marquardt.rs:281-285)
    if (pb->linear_work && out->attempts > 0) pb->linear_work(pb->ctx, &out->linear_solves, &out->linear_iterations);
  };
  int st = validate_optimize(options);
  if (st == FG_OK) st = validate_lm(lm);
  if (st != FG_OK) return finish(st);
  double lambda = lm->initial_damping;
  out->damping = lambda;
  st = pb->prepare(pb->ctx);
  if (st != FG_OK) return finish(st);
  out->cost_evaluations += 1;
  double initial = 0.0;
  st = checked(pb->cost(pb->ctx, 0, &initial), initial);
  if (st != FG_OK) return finish(st);
  out->cost = initial;
  out->initial_cost = initial;
  double current = initial;
  bool stalled = false, tried_undamped = false;
  for (;;) {
    if (pb->dimension == 0) {
      out->final_cost = current;
      out->iterations = out->accepted_steps;
      out->termination = FG_TERM_NO_VARIABLES;
      return finish(FG_OK);
    }
    out->linearizations += 1;
    if (hint) hint(pb->ctx, stalled ? 0.0 : lambda);
    st = pb->linearize(pb->ctx);
    if (st != FG_OK) return finish(st);
    double gradient = 0.0;
    st = pb->gradient_norm(pb->ctx, &gradient);
    if (st != FG_OK) return finish(st);
    out->gradient_norm = gradient;
    if (gradient <= options->gradient_tolerance) {
      out->final_cost = current;
      out->iterations = out->accepted_steps;
      out->termination = FG_TERM_GRADIENT;
      return finish(FG_OK);
    }
    if (stalled && (tried_undamped || out->accepted_steps == options->max_iterations))
      return finish(FG_NO_PROGRESS);
    if (out->accepted_steps == options->max_iterations) return finish(FG_NO_CONVERGENCE);
    st = pb->prepare_damping(pb->ctx, lm->min_column_norm);
    if (st != FG_OK) return finish(st);
    const bool undamped_trial = stalled;
    tried_undamped = tried_undamped || undamped_trial;
    double growth = 2.0;
    bool accepted = false;
    for (size_t trial = 0; trial < lm->max_trials; ++trial) {
      out->attempts += 1;
      const double damping = undamped_trial ? 0.0 : lambda;
      out->damping = damping;
      double norm = 0.0, predicted = 0.0;
      int rejection;
      st = pb->solve_damped(pb->ctx, damping, &norm, &predicted);
      if (st == FG_LINEAR_SOLVE_FAILED) {
        rejection = FG_REJ_LINEAR_SOLVE;
      } else if (st != FG_OK) {
        return finish(st);
      } else {
        // norm_inf(step).map_err(|_| LinearSolveFailed)? aborts on a nonfinite step.
        if (!std::isfinite(norm)) return finish(FG_LINEAR_SOLVE_FAILED);
        out->step_norm = norm;
        if (!std::isfinite(predicted) || predicted <= 0.0) {
          rejection = FG_REJ_PREDICTION;
        } else {
          st = pb->stage(pb->ctx);
          double value = 0.0;
          if (st == FG_OK) {
            out->cost_evaluations += 1;
            st = checked(pb->cost(pb->ctx, 1, &value), value);
          }
          if (st == FG_INVALID_EVALUATION) {
            pb->reject(pb->ctx);
            rejection = FG_REJ_INVALID_EVALUATION;
          } else if (st != FG_OK) {
            pb->reject(pb->ctx);
            return finish(st);
          } else {
            const double actual = current - value;
            const double ratio = std::min(actual / predicted, 2.0);
            if (actual > 0.0 && ratio >= lm->acceptance_threshold) {
              pb->accept(pb->ctx);
              out->accepted_steps += 1;
              out->cost = value;
              stalled = options->step_tolerance > 0.0 && norm <= options->step_tolerance &&
                        (options->cost_tolerance == 0.0 ||
                         actual / std::max(current, 1.0) <= options->cost_tolerance);
              current = value;
              // Nielsen's update; ratio clamped before cubing (levenberg_marquardt.rs:441-447)
              const double c = 2.0 * std::min(ratio, 1.0) - 1.0;
              const double update = std::max(1.0 - c * c * c, 1.0 / 3.0);
              lambda = std::min(std::max(lambda * update, lm->min_damping), lm->max_damping);
              accepted = true;
              if (on_accept) on_accept(out, user);
              break;
            }
            pb->reject(pb->ctx);
            rejection = FG_REJ_POOR_AGREEMENT;
          }
        }
      }
      out->rejected_steps += 1;
      out->last_rejection = rejection;
      if (undamped_trial || lambda == lm->max_damping) break;
      lambda = std::min(lambda * growth, lm->max_damping);
      growth = std::min(growth * 2.0, std::max(lm->max_damping, 2.0));
    }
    if (!accepted) return finish(FG_NO_PROGRESS);
  }
}

extern "C" {

void fg_gn_optimize(const fg_problem* pb, const fg_optimize_options* options, fg_lm_outcome* out) {
  reset(out);
  auto finish = [&](int status) {
    out->status = status;
    if (pb->linear_work) pb->linear_work(pb->ctx, &out->linear_solves, &out->linear_iterations);
  };
  int st = validate_optimize(options);
  if (st != FG_OK) return finish(st);
  st = pb->prepare(pb->ctx);
  if (st != FG_OK) return finish(st);
  double initial = 0.0;
  st = checked(pb->cost(pb->ctx, 0, &initial), initial);
  if (st != FG_OK) return finish(st);
  out->initial_cost = initial;
  out->cost = initial;
  double current = initial;
  size_t iterations = 0;
  if (pb->dimension == 0) {
    out->final_cost = current;
    out->termination = FG_TERM_NO_VARIABLES;
    return finish(FG_OK);
  }
  for (;;) {
    out->linearizations += 1;
    st = pb->linearize(pb->ctx);
    if (st != FG_OK) return finish(st);
    double gradient = 0.0;
    st = pb->gradient_norm(pb->ctx, &gradient);
    if (st != FG_OK) return finish(st);
    out->gradient_norm = gradient;
    if (gradient <= options->gradient_tolerance) {
      out->final_cost = current;
      out->iterations = iterations;
      out->termination = FG_TERM_GRADIENT;
      return finish(FG_OK);
    }
    if (iterations == options->max_iterations) return finish(FG_NO_CONVERGENCE);
    double norm = 0.0;
    out->attempts += 1;
    st = pb->solve(pb->ctx, &norm);
    if (st != FG_OK) return finish(st);
    if (!std::isfinite(norm)) return finish(FG_LINEAR_SOLVE_FAILED);
    out->step_norm = norm;
    if (norm <= options->step_tolerance) {
      out->final_cost = current;
      out->iterations = iterations;
      out->termination = FG_TERM_STEP;
      return finish(FG_OK);
    }
    st = pb->stage(pb->ctx);
    double trial = 0.0;
    if (st == FG_OK) {
      out->cost_evaluations += 1;
      st = checked(pb->cost(pb->ctx, 1, &trial), trial);
    }
    if (st != FG_OK) {
      pb->reject(pb->ctx);
      return finish(st);
    }
    pb->accept(pb->ctx);
    iterations += 1;
    out->accepted_steps = iterations;
    const double decrease = current - trial;
    const bool converged =
        trial == 0.0 || (decrease >= 0.0 && decrease / std::max(current, 1.0) <= options->cost_tolerance);
    current = trial;
    out->cost = current;
    if (converged) {
      out->final_cost = current;
      out->iterations = iterations;
      out->termination = FG_TERM_COST;
      return finish(FG_OK);
    }
  }
}

const char* fg_status_string(int status) {
  switch (status) {
    case FG_OK: return "ok";
    case FG_INVALID_OPTIONS: return "invalid optimization limits, tolerances, or damping controls";
    case FG_LINEAR_SOLVE_FAILED: return "linear system solve failed";
    case FG_NO_CONVERGENCE: return "optimization did not converge";
    case FG_NO_PROGRESS: return "optimization could not make progress";
    case FG_INVALID_EVALUATION: return "measurement evaluation is undefined or nonfinite";
    case FG_DIMENSION_MISMATCH: return "residual or Jacobian dimensions do not match";
    case FG_INVALID_EMISSION: return "invalid factor scope or emitted variable identity";
    case FG_UNSUPPORTED_STRUCTURE: return "residual structure is unsupported by the selected backend";
    case FG_INVALID_INPUT: return "invalid BAL input";
    case FG_DUPLICATE_COVARIANCE_BLOCK: return "duplicate variable in covariance selection";
    case FG_SINGULAR_INFORMATION: return "undamped information is singular or numerically rank deficient";
    case FG_INVALID_INFORMATION: return "invalid or nonfinite covariance information";
    case FG_INVALID_RANK_TOLERANCE: return "relative rank tolerance must be finite and in [0, 1)";
    case FG_UNKNOWN_KEY: return "unknown handle";
    case FG_DEVICE_ERROR: return "HIP runtime error";
    default: return "unknown status";
  }
}

}  // extern "C"
