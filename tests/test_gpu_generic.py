"""The GPU Schur backend for emissions of any factor family (GenericSchur) against the reference's
own Schur fixtures and the oracle's Schur backend, and the LM loop over a host model that is not
BAL (a planar pose / landmark problem). GPU only.

Tolerances: steps and predicted reductions 1e-7 absolute on the reference fixture (its SVD
golden vectors), 1e-8 relative against the oracle on the planar model; final LM cost 1e-9
relative.
"""
import json
import os

import numpy as np
import pytest

from fagra_amd import generic
import fagra_amd as fa
from oracle import oracle as O
from planar_model import OracleBackend, Planar

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "linear_fixtures.json")))


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def from_golden(g, backend):
    assert backend.prepare(g["dimension"]) == fa.OK
    backend.clear()
    for e in g["emissions"]:
        assert backend.accumulate(e["residual"], [(c, np.array(m)) for c, m in e["blocks"]]) == fa.OK
    return backend


def test_reference_schur_fixture():
    # src/schur.rs:690-789: retries over four dampings, unaries, an unused retained and an
    # unused eliminated block, against the SVD golden vectors
    g = GOLDEN["schur_svd"]
    s = from_golden(g, generic.GenericSchur(g["retained"], g["retained_width"], g["eliminated_width"]))
    assert s.prepare_damping(g["floor"]) == fa.OK
    J, r = np.array(g["J"]), np.array(g["r"])
    st, gn = s.gradient_norm()
    assert st == fa.OK and abs(gn - np.max(np.abs(J.T @ r))) < 1e-12
    for lam, expected, pred_expected in zip(g["lambdas"], g["steps"], g["predicted"]):
        st, d, pred = s.solve_damped(lam)
        assert st == fa.OK, (lam, s.diagnostics().termination)
        expected = np.array(expected)
        assert np.linalg.norm(d - expected) < 1e-7 * max(np.linalg.norm(expected), 1.0), lam
        assert abs(pred - pred_expected) < 1e-7
        assert abs(pred - 0.5 * (r @ r - np.sum((J @ d + r) ** 2))) < 1e-7
        for k in [4, 5, 10, 11]:
            assert d[k] == 0.0
        diag = s.diagnostics()
        assert diag.termination == fa.CG_CONVERGED and diag.damping == lam
        assert diag.reduced_relative_residual < 1e-7 and diag.scaled_relative_residual < 1e-7
    s.max_iterations = 1
    st, _, _ = s.solve_damped(0.001)
    assert st == fa.LINEAR_SOLVE_FAILED
    diag = s.diagnostics()
    assert diag.termination == fa.CG_ITERATION_LIMIT and diag.iterations == 1
    assert np.isnan(diag.estimated_relative_residual)
    assert np.isfinite(diag.reduced_relative_residual) and diag.reduced_relative_residual > s.relative_tolerance
    assert s.statistics()[0] == 5
    assert s.prepare(12) == fa.OK
    assert s.statistics() == (0, 0)


def test_partition_structure_and_singular_elimination_fail_explicitly():
    # src/schur.rs:792-870, the same cases as the oracle's test
    j = np.array([[1.0, 1.0]])
    s = generic.GenericSchur(2, 2, 2)
    assert s.prepare(5) == fa.INVALID_OPTIONS
    assert s.prepare(6) == fa.OK
    s.clear()
    assert s.accumulate([1.0], [(2, j), (4, j)]) == fa.UNSUPPORTED_STRUCTURE
    assert s.accumulate([1.0], [(1, j)]) == fa.UNSUPPORTED_STRUCTURE
    assert s.accumulate([1.0], [(6, j)]) == fa.DIMENSION_MISMATCH
    assert s.solve_damped(0.1)[0] != fa.OK
    assert s.accumulate([1.0], [(0, j), (2, j)]) == fa.OK
    assert s.prepare_damping(0.1) == fa.OK
    assert s.solve_damped(0.0)[0] != fa.OK
    assert s.diagnostics().termination == fa.CG_NOT_POSITIVE_DEFINITE
    assert s.solve_damped(0.1)[0] == fa.OK
    assert s.solve_damped(float("nan"))[0] == fa.INVALID_OPTIONS
    rc = generic.GenericSchur(4, 2, 2)
    assert rc.prepare(6) == fa.OK
    assert rc.accumulate([1.0], [(0, j), (2, j)]) == fa.UNSUPPORTED_STRUCTURE
    u = generic.GenericSchur(2, 2, 2)
    assert u.prepare(4) == fa.OK
    for col in [0, 2]:
        assert u.accumulate([1.0, 2.0], [(col, np.eye(2))]) == fa.OK
    for _ in range(2):
        st, d = u.solve()
        assert st == fa.OK and np.allclose(d, [-1, -2, -1, -2], atol=1e-10)
    for retained, n in [(0, 0), (0, 2), (2, 2)]:
        b = generic.GenericSchur(retained, 2, 2)
        assert b.prepare(n) == fa.OK
        b.clear()
        if n:
            assert b.accumulate([1.0], [(0, j)]) == fa.OK
        assert b.prepare_damping(0.1) == fa.OK
        assert b.solve_damped(0.1)[0] == fa.OK
    with pytest.raises(fa.FagraError):
        generic.GenericSchur(17, 17, 3)  # wider than the generic kernels' blocks


def test_planar_model_solves_match_oracle():
    m = Planar(seed=1)
    rw, ew, ret = 3, 2, 3 * m.np_
    g = generic.GenericSchur(ret, rw, ew, relative_tolerance=1e-12)
    o = OracleBackend(ret, rw, ew)
    o.s.relative_tolerance = 1e-12
    for b in (g, o):
        assert b.prepare(m.dimension) == fa.OK
        b.clear()
        assert m.linearize(b) == fa.OK
    st_g, gn_g = g.gradient_norm()
    st_o, gn_o = o.gradient_norm()
    assert st_g == st_o == fa.OK and abs(gn_g - gn_o) <= 1e-12 * gn_o
    for b in (g, o):
        assert b.prepare_damping(1e-6) == fa.OK
    for lam in [1e-4, 1e-2, 1.0, 100.0]:
        st_g, d_g, p_g = g.solve_damped(lam)
        st_o, d_o, p_o = o.solve_damped(lam)
        assert st_g == st_o == fa.OK, (lam, g.diagnostics().termination)
        assert rel(d_g, d_o) < 1e-8, lam
        assert abs(p_g - p_o) <= 1e-8 * abs(p_o)
        dg, do = g.diagnostics(), o.s.diagnostics()
        assert dg.termination == do.termination == fa.CG_CONVERGED
        assert abs(int(dg.iterations) - int(do.iterations)) <= 2
        assert abs(g.last_step_norm() - np.max(np.abs(d_o))) <= 1e-8 * np.max(np.abs(d_o))
    st_g, d_g = g.solve()
    st_o, d_o = o.solve()
    assert st_g == st_o == fa.OK and rel(d_g, d_o) < 1e-8


def test_lm_over_a_host_model_matches_the_oracle_backend():
    runs = {}
    for name, make in (("gpu", lambda r: generic.GenericSchur(r, 3, 2)), ("oracle", lambda r: OracleBackend(r, 3, 2))):
        m = Planar(seed=2, noise=0.02)
        b = make(3 * m.np_)
        out = generic.optimize(m, b, options=fa.OptimizeOptions(40, 1e-10, 0.0, 0.0))
        runs[name] = (out, m.x.copy())
    (a, xa), (b, xb) = runs["gpu"], runs["oracle"]
    assert a.status == b.status and a.accepted_steps == b.accepted_steps and a.rejected_steps == b.rejected_steps
    assert abs(a.cost - b.cost) <= 1e-9 * b.cost and a.cost < 1e-2 * a.initial_cost
    assert rel(xa, xb) < 1e-8
    assert a.linear_solves == b.linear_solves
    # Gauss-Newton over the same model: the prior fixes the gauge, so the undamped step is unique
    m = Planar(seed=3)
    out = generic.optimize_gn(m, generic.GenericSchur(3 * m.np_, 3, 2), options=fa.OptimizeOptions(10, 1e-10, 0, 0))
    assert out.accepted_steps > 0 and out.cost < out.initial_cost
