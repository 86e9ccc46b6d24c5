"""CPU tests of the product library's host side: C-ABI exports, and the LM/GN outer loops on
toy problems checked against the oracle's loops and the reference's closed forms."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import fagra_amd as fa
from fagra_amd import _lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "fagra_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fg_[a-z0-9_]+)\s*\(", text)) - {"fg_on_solve_fn", "fg_on_accept_fn"}


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = header_symbols()
    assert declared, "no declarations parsed"
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)


def test_status_codes_mirror_the_oracle_and_reference_errors():
    assert (fa.OK, fa.INVALID_OPTIONS, fa.LINEAR_SOLVE_FAILED, fa.NO_CONVERGENCE, fa.NO_PROGRESS) == \
        (O.OK, O.INVALID_OPTIONS, O.LINEAR_SOLVE_FAILED, O.NO_CONVERGENCE, O.NO_PROGRESS)
    assert (fa.INVALID_EVALUATION, fa.DIMENSION_MISMATCH, fa.INVALID_EMISSION) == \
        (O.INVALID_EVALUATION, O.DIMENSION_MISMATCH, O.INVALID_EMISSION)
    lib = fa.lib()
    assert lib.fg_status_string(fa.NO_PROGRESS).decode().startswith("optimization could not make progress")


def test_defaults_match_reference():
    o = fa.default_optimize_options()
    assert (o.max_iterations, o.gradient_tolerance, o.step_tolerance, o.cost_tolerance) == (50, 1e-8, 1e-10, 1e-12)
    lm = fa.default_lm_options()
    assert (lm.initial_damping, lm.min_damping, lm.max_damping, lm.min_column_norm, lm.acceptance_threshold,
            lm.max_trials) == (1e-3, 1e-12, 1e12, 1e-6, 1e-4, 16)
    cfg = _lib.SchurConfig()
    fa.lib().fg_schur_config_default(ctypes.byref(cfg), fa.FP64)
    assert cfg.max_iterations == 1000 and cfg.relative_tolerance == 1e-10
    fa.lib().fg_schur_config_default(ctypes.byref(cfg), fa.FP32)
    assert abs(cfg.relative_tolerance - 128 * np.finfo(np.float32).eps) < 1e-12


class Curve:
    """Scalar curve factors r_i(x) with one state (tests/levenberg_marquardt.rs:73-154) and an
    exact damped scalar solve, exposed through the fg_problem callback protocol."""

    def __init__(self, kind, x0, targets, invalid_below=None):
        self.kind, self.x, self.targets = kind, float(x0), list(targets)
        self.trial = None
        self.dimension = 1
        self.calls = {"cost": 0, "linearize": 0}
        self.invalid_below = invalid_below
        self.work = [0, 0]

    def model(self, x):
        out = []
        for t in self.targets:
            if self.kind == "square":
                out.append((x * x - t * t, 2 * x))
            elif self.kind == "atan":
                v = x - t
                out.append((math.atan(v), 1 / (1 + v * v)))
            elif self.kind == "log":
                if x <= 0:
                    return None
                out.append((math.log(x / t), 1 / x))
            else:
                out.append((x - t, 1.0))
        return out

    def prepare(self):
        self.trial = None
        self.work = [0, 0]
        return 0

    def cost(self, trial):
        self.calls["cost"] += 1
        x = self.trial if trial else self.x
        if self.invalid_below is not None and x < self.invalid_below:
            return O.INVALID_EVALUATION, 0.0
        m = self.model(x)
        if m is None:
            return O.INVALID_EVALUATION, 0.0
        return 0, sum(0.5 * r * r for r, _ in m)

    def linearize(self):
        self.calls["linearize"] += 1
        m = self.model(self.x)
        if m is None:
            return O.INVALID_EVALUATION
        self.g = sum(j * r for r, j in m)
        self.h = sum(j * j for r, j in m)
        return 0

    def gradient_norm(self):
        return 0, abs(self.g)

    def prepare_damping(self, floor):
        self.d = max(math.sqrt(self.h), floor)
        return 0

    def solve_damped(self, lam):
        self.work[0] += 1
        self.work[1] += 1
        den = self.h + lam * self.d * self.d
        if den == 0:
            return O.LINEAR_SOLVE_FAILED, 0.0, 0.0
        self.step = -self.g / den
        pred = -self.g * self.step - 0.5 * self.h * self.step ** 2
        return 0, abs(self.step), pred

    def solve(self):
        # Schur::solve = prepare_damping(1e-6) + solve_damped(0) (src/schur.rs:486-489)
        self.prepare_damping(1e-6)
        st, n, _ = self.solve_damped(0.0)
        return st, n

    def stage(self):
        self.trial = self.x + self.step
        return 0

    def accept(self):
        self.x = self.trial
        self.trial = None

    def reject(self):
        self.trial = None

    def linear_work(self):
        return tuple(self.work)


def oracle_problem(c):
    """The same callbacks through the oracle's orc_problem struct."""
    keep = [
        O._FN_PREPARE(lambda ctx: c.prepare()),
        O._FN_COST(lambda ctx, t, out: _put(out, *c.cost(bool(t)))),
        O._FN_LIN(lambda ctx: c.linearize()),
        O._FN_GN(lambda ctx, out: _put(out, *c.gradient_norm())),
        O._FN_PD(lambda ctx, f: c.prepare_damping(f)),
        O._FN_SD(lambda ctx, lam, delta, pred: _sd(c, lam, delta, pred)),
        O._FN_S(lambda ctx, delta: _s(c, delta)),
        O._FN_STAGE(lambda ctx, delta: c.stage()),
        O._FN_VOID(lambda ctx: c.accept()),
        O._FN_VOID(lambda ctx: c.reject()),
        O._FN_WORK(lambda ctx, a, b: _work(a, b, c.linear_work())),
        O._FN_ACC(),
    ]
    p = O.Problem(None, c.dimension, *keep)
    return p, keep


def _put(ptr, status, value=0.0):
    ptr[0] = value
    return status


_step_buf = (ctypes.c_double * 1)()


def _sd(c, lam, delta, pred):
    st, _, pr = c.solve_damped(lam)
    _step_buf[0] = getattr(c, "step", 0.0)
    delta[0] = ctypes.cast(_step_buf, ctypes.POINTER(ctypes.c_double))
    pred[0] = pr
    return st


def _s(c, delta):
    st, _ = c.solve()
    _step_buf[0] = getattr(c, "step", 0.0)
    delta[0] = ctypes.cast(_step_buf, ctypes.POINTER(ctypes.c_double))
    return st


def _work(a, b, w):
    a[0], b[0] = w


def run_both(kind, x0, targets, lm_opts=None, options=None, **kw):
    a = Curve(kind, x0, targets, **kw)
    out_a = fa.lm_optimize(a, options=options, lm_options=lm_opts)
    b = Curve(kind, x0, targets, **kw)
    p, keep = oracle_problem(b)
    out_b = O.LmOutcome()
    o_opts = O.OptimizeOptions(*(getattr(options or fa.default_optimize_options(), f)
                                 for f in ("max_iterations", "gradient_tolerance", "step_tolerance",
                                           "cost_tolerance")))
    l = lm_opts or fa.default_lm_options()
    o_lm = O.LmOptions(l.initial_damping, l.min_damping, l.max_damping, l.min_column_norm,
                       l.acceptance_threshold, l.max_trials)
    O.lib().orc_lm_optimize(ctypes.byref(p), ctypes.byref(o_opts), ctypes.byref(o_lm), ctypes.byref(out_b))
    return a, out_a, b, out_b


FIELDS = ["status", "termination", "iterations", "linearizations", "attempts", "rejected_steps", "accepted_steps",
          "cost_evaluations", "last_rejection", "linear_solves", "linear_iterations"]


@pytest.mark.parametrize("kind,x0,targets", [("square", 3.0, [1.0]), ("atan", 4.0, [0.5, 1.5]),
                                             ("log", 0.2, [2.0, 3.0]), ("linear", -2.0, [1.0, 2.0, 4.0])])
def test_product_lm_loop_matches_oracle_loop(kind, x0, targets):
    a, out_a, b, out_b = run_both(kind, x0, targets)
    for f in FIELDS:
        assert getattr(out_a, f) == getattr(out_b, f), f
    for f in ["cost", "damping", "gradient_norm", "step_norm", "initial_cost", "final_cost"]:
        va, vb = getattr(out_a, f), getattr(out_b, f)
        assert (math.isnan(va) and math.isnan(vb)) or va == vb, f
    assert a.x == b.x and a.calls == b.calls
    # LmWork invariants (tests/levenberg_marquardt.rs:253-256)
    if out_a.status == fa.OK:
        assert out_a.linearizations == out_a.accepted_steps + 1
    assert out_a.attempts == out_a.accepted_steps + out_a.rejected_steps


def test_large_initial_damping_gives_closed_form_first_step():
    # tests/levenberg_marquardt/stagnation.rs:133: one step with lambda = 1e8 on r = x^2 - 1
    x0 = 3.0
    lm = fa.default_lm_options()
    lm.initial_damping = 1e8
    c = Curve("square", x0, [1.0])
    out = fa.lm_optimize(c, options=fa.OptimizeOptions(1, 1e-8, 0.0, 0.0), lm_options=lm)
    assert out.status == fa.NO_CONVERGENCE and out.accepted_steps == 1
    expected = x0 + (1 - x0 * x0) / (2 * x0 * (1 + 1e8))
    assert abs(c.x - expected) < 1e-15
    assert abs(out.cost - 0.5 * (expected ** 2 - 1) ** 2) < 1e-12 * out.cost


def test_invalid_trial_evaluation_is_a_rejection_and_invalid_options_fail_early():
    c = Curve("square", 3.0, [1.0], invalid_below=2.9)
    out = fa.lm_optimize(c, options=fa.OptimizeOptions(5, 1e-8, 0.0, 0.0))
    assert out.rejected_steps >= 1
    bad = fa.default_lm_options()
    bad.acceptance_threshold = 1.0
    c2 = Curve("square", 3.0, [1.0])
    out = fa.lm_optimize(c2, lm_options=bad)
    assert out.status == fa.INVALID_OPTIONS and c2.calls["cost"] == 0
    out = fa.lm_optimize(Curve("square", 3.0, [1.0]), options=fa.OptimizeOptions(0, 1e-8, 0, 0))
    assert out.status == fa.INVALID_OPTIONS


def test_no_variables_and_gauss_newton():
    c = Curve("square", 3.0, [1.0])
    c.dimension = 0
    out = fa.lm_optimize(c)
    assert out.status == fa.OK and out.termination == fa.TERM_NO_VARIABLES and out.attempts == 0
    g = Curve("linear", -2.0, [1.0, 2.0, 4.0])
    out = fa.gn_optimize(g, options=fa.OptimizeOptions(10, 1e-8, 1e-10, 1e-12))
    assert out.status == fa.OK
    assert abs(g.x - 7.0 / 3.0) < 1e-12
