"""The oracle's threaded block loops give bit-identical results to its sequential restatement.

The oracle sums every term in the order of the sequential loops it restates (a block's terms in
emission order), whatever the thread count. These tests compare, bit for bit, a single-threaded
run over the per-emission linearization (orc_schur_accumulate, one observation at a time) with
a multi-threaded run over the bulk linearization on the same seeded BA problem: model, scales,
reduced products, damped solves, costs and a short LM trajectory.
"""
import numpy as np
import pytest

from fagra_amd import synthetic
from oracle import oracle as O


@pytest.fixture(scope="module")
def problem():
    return synthetic.generate("ring16", seed=5, n_cameras=12, n_points=700, huber=1.0)


def _linearized(p, threads, sequential):
    O.set_threads(threads)
    ba = O.BA(p, huber=p.huber)
    s = O.Schur(9 * ba.nc, 9, 3)
    s.relative_tolerance = 1e-10
    assert s.prepare(9 * ba.nc + 3 * ba.np) == O.OK
    st = ba.linearize_sequential(s) if sequential else ba.linearize(s)
    assert st == O.OK
    return ba, s


def _snapshot(ba, s):
    out = {"gradient": s.gradient(), "negres": s.negative_residual(), "cost": ba.cost()[1]}
    assert s.prepare_damping(1e-6) == O.OK
    out["scales"] = s.scales()
    x = np.random.default_rng(2).normal(size=9 * ba.nc)
    for lam in (0.0, 1e-4, 10.0):
        out[f"apply{lam}"] = s.reduced_apply(lam, x)[1]
    y = np.random.default_rng(3).normal(size=s.dimension())
    out["Jy"] = s.jacobian_apply(y)
    out["JTr"] = s.jacobian_transpose_apply(np.random.default_rng(4).normal(size=s.rows()))
    for lam in (1e-3, 1.0):
        st, d, pred = s.solve_damped(lam)
        assert st == O.OK
        out[f"step{lam}"] = d
        out[f"pred{lam}"] = pred
    return out


def test_threaded_oracle_is_bit_identical_to_sequential(problem):
    try:
        a = _snapshot(*_linearized(problem, 1, True))
        b = _snapshot(*_linearized(problem, 8, False))
    finally:
        O.set_threads(8)
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_threaded_lm_trajectory_is_bit_identical(problem):
    runs = []
    try:
        for threads in (1, 8):
            O.set_threads(threads)
            out, trace = O.BA(problem, huber=problem.huber).lm(
                options=O.OptimizeOptions(6, 1e-8, 0.0, 0.0), tol=1e-10, trace_capacity=10)
            runs.append((out.cost, out.accepted_steps, out.rejected_steps, out.linear_iterations, trace))
    finally:
        O.set_threads(8)
    assert runs[0][:4] == runs[1][:4]
    assert np.array_equal(runs[0][4], runs[1][4])
