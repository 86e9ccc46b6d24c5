"""The oracle's LM loop (orc_lm_optimize) pinned by the reference's own LM fixtures and by
closed forms, on small dense problems; the product's loop (fg_lm_optimize) is held to the same.

Restated reference tests (the factors and expected statistics are theirs; the backends are
dense damped solves written here from src/normal.rs:157-235 and src/qr.rs:226-241):
* tests/levenberg_marquardt/stagnation.rs:57-92, :94-108, :110-150, :153-185, :187-215
* tests/levenberg_marquardt.rs:200-270 (retries on Atan / Log / Square curves), :310-345
Closed forms pinned here:
* Nielsen's update (src/optimization/levenberg_marquardt.rs:441-447): on a linear residual the
  gain ratio is 1, so every accepted step divides the damping by 3 until min_damping.
* Retry growth (:460-466): rejected trials multiply the damping by 2, 4, 8, ... (reset per
  linearization), so the k-th trial of one linearization runs at lambda0 * 2^(k(k+1)/2); the
  damping is capped at max_damping, where the loop stops, and the factor at max(max_damping, 2).
"""
import ctypes
import math

import numpy as np
import pytest

import fagra_amd as fa
from oracle import oracle as O


class Dense:
    """A least-squares problem on x in R^n with a dense damped solve.

    residual(x) returns (r, J) or None when the evaluation is invalid. backend "normal" is
    DenseNormalCholesky (scales sqrt(diag J^T J) floored, Cholesky of J^T J + lambda D^2, its
    failure a LinearSolveFailed, the reference's predicted-reduction sum); "qr" is DenseQr
    (least squares of [J; sqrt(lambda) D] against [-r; 0], predicted -g.d - |J d|^2 / 2)."""

    def __init__(self, x0, residual, backend="qr"):
        self.x = np.array(x0, dtype=np.float64)
        self.dimension = self.x.size
        self.residual = residual
        self.backend = backend
        self.trial = None
        self.calls = {"cost": 0, "jacobian": 0}
        self.lambdas = []
        self.solves = 0

    def prepare(self):
        self.trial = None
        self.solves = 0
        return O.OK

    def cost(self, trial):
        self.calls["cost"] += 1
        m = self.residual(self.trial if trial else self.x)
        if m is None:
            return O.INVALID_EVALUATION, 0.0
        r = np.atleast_1d(m[0])
        return O.OK, float(0.5 * r @ r)

    def linearize(self):
        self.calls["jacobian"] += 1
        m = self.residual(self.x)
        if m is None:
            return O.INVALID_EVALUATION
        self.r = np.atleast_1d(np.asarray(m[0], dtype=np.float64))
        self.J = np.atleast_2d(np.asarray(m[1], dtype=np.float64)).reshape(self.r.size, self.dimension)
        self.N = self.J.T @ self.J
        self.g = self.J.T @ self.r
        return O.OK

    def gradient_norm(self):
        if not np.all(np.isfinite(self.g)):
            return O.INVALID_EVALUATION, 0.0
        return O.OK, float(np.max(np.abs(self.g))) if self.g.size else 0.0

    def prepare_damping(self, floor):
        self.D = np.maximum(np.sqrt(np.diag(self.N)), floor)
        return O.OK

    def solve_damped(self, lam):
        self.lambdas.append(lam)
        self.solves += 1
        n = self.dimension
        if self.backend == "normal":
            H = self.N + np.diag((math.sqrt(lam) * self.D) ** 2)
            try:
                L = np.linalg.cholesky(H)
            except np.linalg.LinAlgError:
                return O.LINEAR_SOLVE_FAILED, None, 0.0
            if not np.all(np.isfinite(L)) or np.any(np.diag(L) <= 0):
                return O.LINEAR_SOLVE_FAILED, None, 0.0
            d = np.linalg.solve(L.T, np.linalg.solve(L, -self.g))
            rhs = -self.g
            pred = 0.0
            for c in range(n):
                pred += rhs[c] * d[c] - 0.5 * d[c] * self.N[c, c] * d[c]
                for row in range(c + 1, n):
                    pred -= d[c] * self.N[row, c] * d[row]
        else:
            A = np.vstack([self.J, np.diag(math.sqrt(lam) * self.D)])
            b = np.concatenate([-self.r, np.zeros(n)])
            d = np.linalg.lstsq(A, b, rcond=None)[0]
            Jd = self.J @ d
            pred = float(-self.g @ d - 0.5 * Jd @ Jd)
        if not np.all(np.isfinite(d)) or not math.isfinite(pred):
            return O.LINEAR_SOLVE_FAILED, None, 0.0
        self.step = d
        return O.OK, d, pred

    def solve(self):
        self.prepare_damping(1e-6)
        st, d, _ = self.solve_damped(0.0)
        return st, d

    def stage(self):
        self.trial = self.x + self.step
        return O.OK

    def accept(self):
        if self.trial is not None:
            self.x = self.trial
        self.trial = None

    def reject(self):
        self.trial = None

    def linear_work(self):
        return self.solves, self.solves


def _put(ptr, status, value=0.0):
    ptr[0] = value
    return status


def run_oracle(prob, options, lm):
    """orc_lm_optimize over the problem's callbacks; returns the LmOutcome and the damping of
    each accepted trial (seen by on_accept)."""
    buf = {}
    accepted = []

    def sd(ctx, lam, delta, pred):
        st, d, p = prob.solve_damped(lam)
        if st == O.OK:
            buf["d"] = np.ascontiguousarray(d)
            delta[0] = buf["d"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            pred[0] = p
        return st

    def s(ctx, delta):
        st, d = prob.solve()
        if st == O.OK:
            buf["d"] = np.ascontiguousarray(d)
            delta[0] = buf["d"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return st

    def work(ctx, a, b):
        a[0], b[0] = prob.linear_work()

    keep = [
        O._FN_PREPARE(lambda ctx: prob.prepare()),
        O._FN_COST(lambda ctx, t, out: _put(out, *prob.cost(bool(t)))),
        O._FN_LIN(lambda ctx: prob.linearize()),
        O._FN_GN(lambda ctx, out: _put(out, *prob.gradient_norm())),
        O._FN_PD(lambda ctx, f: prob.prepare_damping(f)),
        O._FN_SD(sd), O._FN_S(s),
        O._FN_STAGE(lambda ctx, delta: prob.stage()),
        O._FN_VOID(lambda ctx: prob.accept()),
        O._FN_VOID(lambda ctx: prob.reject()),
        O._FN_WORK(work),
        O._FN_ACC(lambda ctx, st: accepted.append(st.contents.damping)),
    ]
    p = O.Problem(None, prob.dimension, *keep)
    out = O.LmOutcome()
    o = O.OptimizeOptions(*options)
    l = O.LmOptions(*lm)
    O.lib().orc_lm_optimize(ctypes.byref(p), ctypes.byref(o), ctypes.byref(l), ctypes.byref(out))
    return out, accepted


class _ProductCallbacks:
    """The same problem through the product's fg_problem protocol (fa.lm_optimize)."""

    def __init__(self, prob):
        self.p = prob
        self.dimension = prob.dimension

    def prepare(self):
        return self.p.prepare()

    def cost(self, trial):
        return self.p.cost(trial)

    def linearize(self):
        return self.p.linearize()

    def gradient_norm(self):
        return self.p.gradient_norm()

    def prepare_damping(self, f):
        return self.p.prepare_damping(f)

    def solve_damped(self, lam):
        st, d, pred = self.p.solve_damped(lam)
        if st != O.OK:
            return st, 0.0, 0.0
        n = float(np.max(np.abs(d))) if np.all(np.isfinite(d)) else float("nan")
        return st, n, pred

    def solve(self):
        st, d = self.p.solve()
        return st, (float(np.max(np.abs(d))) if st == O.OK else 0.0)

    def stage(self):
        return self.p.stage()

    def accept(self):
        self.p.accept()

    def reject(self):
        self.p.reject()

    def linear_work(self):
        return self.p.linear_work()


def run_product(prob, options, lm):
    accepted = []
    out = fa.lm_optimize(_ProductCallbacks(prob), options=fa.OptimizeOptions(*options),
                         lm_options=fa.LmOptions(*lm), on_accept=lambda st: accepted.append(st.damping))
    return out, accepted


RUNNERS = {"oracle": run_oracle, "product": run_product}
DEFAULT_LM = (1e-3, 1e-12, 1e12, 1e-6, 1e-4, 16)


def imu_tolerances(max_iterations):  # tests/levenberg_marquardt/stagnation.rs:47-54
    return (max_iterations, 0.001, 1e-5, 1e-8)


def lm_opts(**kw):
    names = ("initial_damping", "min_damping", "max_damping", "min_column_norm", "acceptance_threshold",
             "max_trials")
    d = dict(zip(names, DEFAULT_LM))
    d.update(kw)
    return tuple(d[n] for n in names)


# --- the reference's factors (tests/levenberg_marquardt.rs:73-116, stagnation.rs:3-41) ---

def curve(kind, target):
    def f(x):
        v = x[0]
        if kind == "linear":
            return np.array([v - target]), np.array([[1.0]])
        if kind == "square":
            return np.array([v * v - target * target]), np.array([[2 * v]])
        if kind == "atan":
            w = v - target
            return np.array([math.atan(w)]), np.array([[1 / (1 + w * w)]])
        if kind == "log":
            if v <= 0:
                return None
            return np.array([math.log(v / target)]), np.array([[1 / v]])
        raise ValueError(kind)
    return f


def with_unused(f):
    """An extra variable that no factor touches (a zero Jacobian column)."""
    def g(x):
        m = f(x[:1])
        if m is None:
            return None
        return m[0], np.hstack([m[1], np.zeros((m[1].shape[0], 1))])
    return g


def coupled(x):
    a, b = x
    return (np.array([1e8 * (a + b), 100.0 * (a - b - 1e-4)]), np.array([[1e8, 1e8], [100.0, -100.0]]))


@pytest.mark.parametrize("who", list(RUNNERS))
def test_full_rank_weak_direction_recovers_at_unchanged_tolerances(who):
    # stagnation.rs:57-92: the damped step stalls in the weak direction, the undamped trial
    # solves it and the gradient tolerance ends the run
    for backend in ("qr", "normal"):
        prob = Dense([0.0, 0.0], coupled, backend)
        out, _ = RUNNERS[who](prob, imu_tolerances(20), lm_opts(max_trials=1))
        assert out.status == O.OK and out.termination == fa.TERM_GRADIENT, backend
        assert out.iterations == 2 and out.attempts == 2 and out.damping == 0.0
        assert out.final_cost < out.initial_cost and out.gradient_norm <= 0.001
        assert abs(prob.x[0] - 5e-5) < 1e-7 and abs(prob.x[1] + 5e-5) < 1e-7


@pytest.mark.parametrize("who", list(RUNNERS))
def test_stagnation_trial_does_not_extend_the_accepted_step_budget(who):
    # stagnation.rs:94-108
    prob = Dense([0.0, 0.0], coupled, "qr")
    out, _ = RUNNERS[who](prob, imu_tolerances(1), DEFAULT_LM)
    assert out.status == O.NO_PROGRESS
    assert out.accepted_steps == 1 and out.attempts == 1
    assert out.damping > 0.0 and out.gradient_norm > 0.001


@pytest.mark.parametrize("who", list(RUNNERS))
@pytest.mark.parametrize("initial,threshold", [(0.1, 1e-4), (0.5, 0.9)])
def test_stagnation_trial_rejects_uphill_or_poor_agreement_and_rolls_back(who, initial, threshold):
    # stagnation.rs:110-150
    prob = Dense([initial], curve("square", 1.0), "qr")
    out, _ = RUNNERS[who](prob, imu_tolerances(20), lm_opts(initial_damping=1e8, acceptance_threshold=threshold))
    assert out.status == O.NO_PROGRESS
    accepted = initial + (1.0 - initial * initial) / (2.0 * initial * (1.0 + 1e8))
    assert abs(prob.x[0] - accepted) < 1e-15
    assert out.accepted_steps == 1 and out.attempts == 2 and out.rejected_steps == 1
    assert out.last_rejection == fa.REJ_POOR_AGREEMENT
    assert out.damping == 0.0
    assert prob.calls["jacobian"] == 2 and prob.calls["cost"] == 3
    assert abs(out.cost - 0.5 * (accepted * accepted - 1.0) ** 2) < 1e-15


@pytest.mark.parametrize("who", list(RUNNERS))
def test_singular_undamped_stagnation_trial_remains_a_bounded_failure(who):
    # stagnation.rs:153-185 (DenseNormalCholesky: the undamped system is singular)
    prob = Dense([0.0, 7.0], with_unused(curve("linear", 1.0)), "normal")
    out, _ = RUNNERS[who](prob, imu_tolerances(20), lm_opts(initial_damping=1e9))
    assert out.status == O.NO_PROGRESS
    assert out.accepted_steps == 1 and out.attempts == 2 and out.rejected_steps == 1
    assert out.last_rejection == fa.REJ_LINEAR_SOLVE
    assert out.cost_evaluations == 2
    assert abs(prob.x[0] - 1.0 / (1.0 + 1e9)) < 1e-20
    assert prob.x[1] == 7.0


@pytest.mark.parametrize("who", list(RUNNERS))
def test_an_accepted_undamped_trial_does_not_restart_stagnation_recovery(who):
    # stagnation.rs:187-215: small damped step, useful undamped step, small damped step
    prob = Dense([0.8], curve("square", 1.0), "qr")
    out, damping = RUNNERS[who](prob, imu_tolerances(20), lm_opts(initial_damping=1e9))
    assert out.status == O.NO_PROGRESS
    assert out.accepted_steps == 3 and out.attempts == 3 and out.rejected_steps == 0
    assert out.damping > 0.0 and out.gradient_norm > 0.001
    assert prob.x[0] > 1.0
    assert damping[0] == 1e9 and damping[1] == 0.0 and damping[2] > 0.0


@pytest.mark.parametrize("who", list(RUNNERS))
@pytest.mark.parametrize("kind,initial,target", [("atan", 2.0, 0.0), ("log", 10.0, 1.0), ("square", 0.1, 1.0)])
@pytest.mark.parametrize("backend", ["qr", "normal"])
def test_retries_reach_the_target_with_consistent_work_counts(who, kind, initial, target, backend):
    # tests/levenberg_marquardt.rs:200-270 (f64)
    prob = Dense([initial], curve(kind, target), backend)
    out, _ = RUNNERS[who](prob, (100, 1e-8, 1e-10, 1e-12), DEFAULT_LM)
    assert out.status == O.OK
    assert abs(prob.x[0] - target) < 2e-5
    assert out.final_cost < out.initial_cost
    assert out.rejected_steps > 0, kind
    assert out.attempts == out.accepted_steps + out.rejected_steps
    assert prob.calls["jacobian"] == out.linearizations == out.accepted_steps + 1
    assert prob.calls["cost"] == out.cost_evaluations
    assert out.gradient_norm <= 1e-8


@pytest.mark.parametrize("who", list(RUNNERS))
@pytest.mark.parametrize("overdamped", [False, True])
def test_rejection_limits_and_overdamping_do_not_report_convergence(who, overdamped):
    # tests/levenberg_marquardt.rs:310-345
    initial = 0.0 if overdamped else 10.0
    prob = Dense([initial], curve("linear" if overdamped else "log", 1.0), "qr")
    lm = lm_opts(max_trials=1, **({"initial_damping": 1e20, "max_damping": 1e20} if overdamped else {}))
    out, _ = RUNNERS[who](prob, (50, 1e-8, 1e-10, 1e-12), lm)
    assert out.status == O.NO_PROGRESS
    assert prob.x[0] == initial
    assert out.accepted_steps == 0 and out.attempts == 1 and prob.calls["jacobian"] == 1
    assert out.last_rejection == (fa.REJ_POOR_AGREEMENT if overdamped else fa.REJ_INVALID_EVALUATION)


@pytest.mark.parametrize("who", list(RUNNERS))
def test_nielsen_update_divides_damping_by_three_on_a_linear_residual(who):
    # ratio = actual / predicted = 1 on a linear residual, so the update factor
    # max(1/3, 1 - (2 min(ratio, 1) - 1)^3) is 1/3 at every accepted step, down to min_damping
    lam0, floor, k = 1e4, 1e-1, 14
    prob = Dense([0.0], curve("linear", 1.0), "qr")
    out, damping = RUNNERS[who](prob, (k, 0.0, 0.0, 0.0), lm_opts(initial_damping=lam0, min_damping=floor))
    assert out.status == O.NO_CONVERGENCE and out.accepted_steps == k and out.rejected_steps == 0
    for i, lam in enumerate(damping):
        expect = max(lam0 / 3.0 ** i, floor)
        assert abs(lam - expect) <= 1e-12 * expect, (i, lam, expect)
    assert damping[-1] == floor
    assert abs(out.damping - floor) <= 1e-15


@pytest.mark.parametrize("who", list(RUNNERS))
def test_rejected_trials_grow_damping_by_doubling_factors(who):
    # A trial that moves x below 2.9 is an invalid evaluation. From x = 3 on r = x^2 - 1 the
    # step is -(4/3) / (1 + lambda), so trials are rejected until lambda >= 12.33: the k-th
    # trial runs at 1e-3 * 2^(k(k+1)/2) (1e-3, 2e-3, 8e-3, 6.4e-2, 1.024, 32.768).
    def f(x):
        if x[0] < 2.9:
            return None
        return curve("square", 1.0)(x)
    prob = Dense([3.0], f, "qr")
    out, damping = RUNNERS[who](prob, (1, 0.0, 0.0, 0.0), DEFAULT_LM)
    expect = [1e-3 * 2.0 ** (k * (k + 1) // 2) for k in range(6)]
    assert len(prob.lambdas) == 6
    for got, want in zip(prob.lambdas, expect):
        assert abs(got - want) <= 1e-15 * want, (prob.lambdas, expect)
    assert out.rejected_steps == 5 and out.accepted_steps == 1
    assert out.last_rejection == fa.REJ_INVALID_EVALUATION and damping == [expect[-1]]
    # capped: the growth factor itself is capped at max(max_damping, 2) = 2 when max_damping is
    # 1, so the damping doubles (1e-3 * 2^k up to 0.512) and the trial at the cap 1 ends the loop
    prob = Dense([3.0], f, "qr")
    out, _ = RUNNERS[who](prob, (1, 0.0, 0.0, 0.0), lm_opts(max_damping=1.0))
    assert prob.lambdas == pytest.approx([1e-3 * 2.0 ** k for k in range(10)] + [1.0], rel=1e-15)
    assert out.status == O.NO_PROGRESS and out.attempts == 11 and out.accepted_steps == 0
