"""The oracle's LSMR backend (src/lsmr.rs over faer 0.24.4's LSMR, restated from Fong &
Saunders 2011) pinned by the reference's own LSMR fixtures, dense least squares and the
minimum-norm solution of the undamped (gauge-singular) BA system."""
import numpy as np
import pytest

from fagra_amd import synthetic
from oracle import oracle as O

MODES = [O.LSMR_DIAGONAL, O.LSMR_BLOCK]


def generic(J, r, floor):
    """One retained block covering all columns (the reference fixtures' single emission)."""
    n = J.shape[1]
    s = O.Schur(n, n, 1)
    assert s.prepare(n) == O.OK
    assert s.accumulate(list(r), [(0, J)]) == O.OK
    assert s.prepare_damping(floor) == O.OK
    return s


def test_prepared_modes_closed_form():
    # src/lsmr.rs:731-778: J = diag(1, 3), r = (-1, -1), lambda = 1 -> step (1/2, 1/6)
    s = generic(np.diag([1.0, 3.0]), [-1.0, -1.0], 0.1)
    for mode in (O.LSMR_IDENTITY, O.LSMR_DIAGONAL, O.LSMR_BLOCK):
        st, d, pred, diag = s.lsmr_solve_damped(1.0, mode)
        assert st == O.OK and diag.termination == O.LSMR_CONVERGED
        assert abs(d[0] - 0.5) < 1e-12 and abs(d[1] - 1.0 / 6.0) < 1e-12
    # identity preconditioning leaves two distinct singular values: one iteration is short
    st, d, pred, diag = s.lsmr_solve_damped(1.0, O.LSMR_IDENTITY, max_iterations=1)
    assert st == O.LINEAR_SOLVE_FAILED and diag.termination == O.LSMR_ITERATION_LIMIT and diag.iterations == 1


@pytest.mark.parametrize("mode", MODES)
def test_mixed_scale_damping_matches_svd(mode):
    # src/lsmr.rs:780-850: columns scaled over 8 decades, an all-zero column, four dampings
    J = np.array([[1e-4, 2.0, 0.0, 0.0], [2e-4, -1.0, 3e4, 0.0], [0.0, 1.0, -1e4, 0.0]])
    r = np.array([1.0, -2.0, 3.0])
    s = generic(J, r, 1e-3)
    D = s.scales()
    assert np.allclose(D, np.maximum(np.linalg.norm(J, axis=0), 1e-3), rtol=1e-15)
    for lam in [0.1, 10.0, 1e-3, 1e6]:
        A = np.vstack([J, np.sqrt(lam) * np.diag(D)])
        expected, *_ = np.linalg.lstsq(A, np.concatenate([-r, np.zeros(4)]), rcond=1e-14)
        st, d, pred, diag = s.lsmr_solve_damped(lam, mode)
        assert st == O.OK and diag.termination == O.LSMR_CONVERGED and diag.finite_step
        assert np.all(np.abs(d - expected) < 1e-7 * np.maximum(np.abs(expected), 1.0)), lam
        assert d[3] == 0.0
        assert abs(pred - 0.5 * (r @ r - np.sum((J @ d + r) ** 2))) < 1e-7
        assert diag.scaled_relative_residual < 1e-7
        assert abs(diag.scaled_relative_residual - diag.estimated_relative_residual) < 1e-7
        assert diag.original_relative_residual < 1e-7
        assert diag.damping == lam and diag.min_scale == D.min() and diag.max_scale == D.max()
    st, d, pred, diag = s.lsmr_solve_damped(1e-3, O.LSMR_DIAGONAL, max_iterations=1)
    assert st == O.LINEAR_SOLVE_FAILED and diag.termination == O.LSMR_ITERATION_LIMIT and diag.iterations == 1
    assert diag.finite_step and np.isfinite(diag.original_relative_residual)
    assert diag.scaled_relative_residual > 1e-10


def ba_dense(seed=3):
    p = synthetic.generate("ring16", seed=seed, n_cameras=4, n_points=60, huber=1.0)
    st, s = O.ba_schur(O.BA(p, huber=1.0), tol=1e-14)
    assert st == O.OK and s.prepare_damping(1e-6) == O.OK
    n = s.dimension()
    J = np.stack([s.jacobian_apply(e) for e in np.eye(n)], axis=1)
    return s, J, -s.negative_residual(), s.scales()


@pytest.mark.parametrize("mode", [O.LSMR_IDENTITY, O.LSMR_DIAGONAL, O.LSMR_BLOCK])
def test_ba_damped_lsmr_matches_dense_and_schur(mode):
    s, J, r, D = ba_dense()
    n = J.shape[1]
    for lam in [1e-3, 1e-1, 10.0]:
        A = np.vstack([J, np.sqrt(lam) * np.diag(D)])
        x, *_ = np.linalg.lstsq(A, np.concatenate([-r, np.zeros(n)]), rcond=None)
        st, d, pred, diag = s.lsmr_solve_damped(lam, mode, max_iterations=5000, tol=1e-12)
        assert st == O.OK, (lam, diag.termination, diag.iterations)
        assert np.linalg.norm(d - x) < 1e-6 * np.linalg.norm(x), lam
        assert abs(pred - 0.5 * (r @ r - np.sum((J @ d + r) ** 2))) < 1e-6 * abs(pred)
        # the Schur backend solves the same damped system exactly
        st, ds, preds = s.solve_damped(lam)
        assert st == O.OK and np.linalg.norm(d - ds) < 1e-6 * np.linalg.norm(ds)
        # explicit diagnostics against the dense normal residual
        nr = J.T @ (J @ d + r) + lam * D**2 * d
        g = J.T @ r
        assert abs(diag.original_absolute_residual - np.linalg.norm(nr)) <= 1e-9 * np.linalg.norm(g)
        assert abs(diag.original_relative_residual - np.linalg.norm(nr) / np.linalg.norm(g)) < 1e-9
        if mode == O.LSMR_DIAGONAL:
            ref = np.linalg.norm(nr / D) / np.linalg.norm(g / D)
            assert abs(diag.scaled_relative_residual - ref) < 1e-9
        if mode == O.LSMR_BLOCK:
            assert diag.block_fallbacks == 0
    stats = s.statistics()
    assert stats[0] >= 3 and stats[1] > 0


def test_undamped_lsmr_is_the_minimum_norm_solution():
    # x0 = 0 keeps LSMR in range(J^T): on the rank-deficient BA Jacobian (similarity gauge)
    # it converges to pinv(J)(-r), the minimum-norm least-squares step
    s, J, r, D = ba_dense(seed=5)
    st, d, diag = s.lsmr_solve(max_iterations=20000, tol=1e-13)
    assert st == O.OK and diag.termination == O.LSMR_CONVERGED
    x = np.linalg.pinv(J, rcond=1e-12) @ (-r)
    assert np.linalg.norm(d - x) < 1e-5 * np.linalg.norm(x)
    assert np.linalg.matrix_rank(J) < J.shape[1]  # the gauge makes J singular
