"""Pins the CPU oracle against the reference tests' golden values and against mathematics.

Sources of truth (none is the oracle's own formula retyped):
* closed forms and literal expectations of the reference's unit tests (file:line cited);
* dense SVD least squares of the same fixtures (tests/golden/linear_fixtures.json, numpy);
* scipy's rotation-vector exponential and central differences for the projection;
* brute-force dense solves of small assembled bundle-adjustment systems.
"""
import json
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from fagra_amd import synthetic
from oracle import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "linear_fixtures.json")))


# ---- projection: examples/bal/model.rs:116-184, tests/bal.rs:63-113 ----------------------

def test_projection_closed_form_and_zero_depth():
    # tests/bal.rs:104-112
    cam = [0, 0, 0, 0, 0, 0, 100.0, 0.1, 0.01]
    st, px = O.project(cam, [1.0, 2.0, -2.0])
    assert st == O.OK
    assert abs(px[0] - 50.0 * (1.0 + 0.1 * 1.25 + 0.01 * 1.25 * 1.25)) < 1e-12
    assert px[1] == 2.0 * px[0]
    st, _ = O.project(cam, [0.0, 0.0, 0.0])
    assert st == O.INVALID_EVALUATION


@pytest.mark.parametrize("rotation", [[0, 0, 0], [1e-10, -2e-10, 3e-10], [0.4, -0.2, 0.1], [3.1, 0.1, -0.1]])
def test_jacobians_match_central_differences(rotation):
    # tests/bal.rs:64-103
    cam = np.array([*rotation, 0.2, -0.1, 0.3, 500.0, 0.01, -0.001])
    pt = np.array([0.8, -0.4, -5.0])
    jc, jp = O.jacobians(cam, pt)
    h = 1e-6
    for k in range(12):
        cp, cm, pp, pm = cam.copy(), cam.copy(), pt.copy(), pt.copy()
        if k < 9:
            cp[k] += h
            cm[k] -= h
        else:
            pp[k - 9] += h
            pm[k - 9] -= h
        num = (O.project(cp, pp)[1] - O.project(cm, pm)[1]) / (2 * h)
        ana = jc[:, k] if k < 9 else jp[:, k - 9]
        assert np.linalg.norm(num - ana) < 1e-5 * max(np.linalg.norm(ana), 1.0), k


@pytest.mark.parametrize("w", [[0, 0, 0], [1e-10, -2e-10, 3e-10], [1e-5, 2e-5, -1e-5], [0.4, -0.2, 0.1],
                               [3.1, 0.1, -0.1], [-1.0, 2.0, 0.5]])
def test_rotation_and_right_jacobian(w):
    R, Jr = O.projection([*w, 0, 0, 0, 1, 0, 0])
    # the quaternion path equals the rotation-vector exponential (scipy), and is orthonormal
    assert np.allclose(R, Rotation.from_rotvec(w).as_matrix(), atol=1e-14, rtol=0)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
    # Jr is the right Jacobian: Exp(w + e) ~= Exp(w) Exp(Jr e)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1e-6
        lhs = Rotation.from_rotvec(np.add(w, e)).as_matrix()
        rhs = R @ Rotation.from_rotvec(Jr @ e).as_matrix()
        assert np.abs(lhs - rhs).max() < 1e-11
    if not np.any(w):
        assert np.array_equal(Jr, np.eye(3))


# ---- linear model and block factors ---------------------------------------------------------

def test_verification_closed_form():
    # src/linear_model.rs:345-367
    s = O.Schur(0, 1, 3)
    assert s.prepare(3) == O.OK
    J = np.array([[1.0, 2.0, 0.0], [3.0, -1.0, 0.0]])
    assert s.accumulate([-2.0, 5.0], [(0, J)]) == O.OK
    assert s.prepare_scales(0.5) == O.OK
    pred, nr, ref, rn, gn = s.verify([0.2, -0.3, 0.4], 0.7)
    assert abs(pred + 5.785) < 1e-12
    assert np.allclose(nr, [16.7, -11.75, 0.07], atol=1e-12, rtol=0)
    assert np.array_equal(ref, [13.0, -9.0, 0.0])
    assert rn > 20.0
    assert np.array_equal(s.gradient(), [13.0, -9.0, 0.0])
    assert np.array_equal(s.negative_residual(), [2.0, -5.0])


def test_block_fallback_counts():
    # src/linear_model/block_factors.rs:212-245 (a 1x2 block [1, 1] on columns 0..2 of 3)
    s = O.Schur(2, 2, 1)
    assert s.prepare(3) == O.OK
    assert s.accumulate([0.0], [(0, np.array([[1.0, 1.0]]))]) == O.OK
    assert s.prepare_blocks([1.0, 1.0, 1.0]) == O.OK
    assert s.factor(0.0) == 1                       # singular Gram [[1,1],[1,1]]
    v = s.apply_factor_inverse([2.0, 3.0, 4.0], False)
    assert np.array_equal(v, [2.0, 3.0, 4.0])       # failed blocks stay identity
    assert s.factor(0.01) == 0
    s.gram_set(0, np.inf)
    assert s.factor(0.01) == 1                      # nonfinite block falls back
    s.clear()
    assert s.prepare_blocks([1.0, 1.0, 1.0]) == O.OK
    assert s.factor(0.1) == 0


def _schur_from_golden(g):
    s = O.Schur(g["retained"], g["retained_width"], g["eliminated_width"])
    assert s.prepare(g["dimension"]) == O.OK
    s.clear()
    for e in g["emissions"]:
        assert s.accumulate(e["residual"], [(c, np.array(m)) for c, m in e["blocks"]]) == O.OK
    return s


def test_schur_matches_augmented_svd_with_retries_unaries_and_unused_blocks():
    # src/schur.rs:690-789 against SVD golden vectors
    g = GOLDEN["schur_svd"]
    s = _schur_from_golden(g)
    assert s.prepare_damping(g["floor"]) == O.OK
    assert np.allclose(s.scales(), g["scales"], rtol=1e-14, atol=0)
    J, r = np.array(g["J"]), np.array(g["r"])
    for lam, expected, pred_expected in zip(g["lambdas"], g["steps"], g["predicted"]):
        st, d, pred = s.solve_damped(lam)
        assert st == O.OK
        expected = np.array(expected)
        assert np.linalg.norm(d - expected) < 1e-7 * max(np.linalg.norm(expected), 1.0), lam
        assert abs(pred - pred_expected) < 1e-7
        assert abs(pred - 0.5 * (r @ r - np.sum((J @ d + r) ** 2))) < 1e-7
        for k in [4, 5, 10, 11]:
            assert d[k] == 0.0
        diag = s.diagnostics()
        assert diag.termination == O.CG_CONVERGED
        assert diag.reduced_relative_residual < 1e-7 and diag.scaled_relative_residual < 1e-7
    s.max_iterations = 1
    st, _, _ = s.solve_damped(0.001)
    assert st == O.LINEAR_SOLVE_FAILED
    diag = s.diagnostics()
    assert diag.termination == O.CG_ITERATION_LIMIT and diag.iterations == 1
    assert np.isnan(diag.estimated_relative_residual)
    assert np.isfinite(diag.reduced_relative_residual) and diag.reduced_relative_residual > s.relative_tolerance
    assert s.prepare(12) == O.OK
    assert s.statistics() == (0, 0)


def test_partition_structure_and_singular_elimination_fail_explicitly():
    # src/schur.rs:792-870
    j = np.array([[1.0, 1.0]])
    s = O.Schur(2, 2, 2)
    assert s.prepare(5) == O.INVALID_OPTIONS
    assert s.prepare(6) == O.OK
    s.clear()
    assert s.accumulate([1.0], [(2, j), (4, j)]) == O.UNSUPPORTED_STRUCTURE
    assert s.accumulate([1.0], [(1, j)]) == O.UNSUPPORTED_STRUCTURE
    assert s.solve_damped(0.1)[0] != O.OK
    assert s.accumulate([1.0], [(0, j), (2, j)]) == O.OK
    assert s.prepare_damping(0.1) == O.OK
    assert s.solve_damped(0.0)[0] != O.OK
    assert s.diagnostics().termination == O.CG_NOT_POSITIVE_DEFINITE
    assert s.solve_damped(0.1)[0] == O.OK
    assert s.solve_damped(float("nan"))[0] != O.OK
    rc = O.Schur(4, 2, 2)
    assert rc.prepare(6) == O.OK
    assert rc.accumulate([1.0], [(0, j), (2, j)]) == O.UNSUPPORTED_STRUCTURE
    u = O.Schur(2, 2, 2)
    assert u.prepare(4) == O.OK
    for col in [0, 2]:
        assert u.accumulate([1.0, 2.0], [(col, np.eye(2))]) == O.OK
    for _ in range(2):
        st, d = u.solve()
        assert st == O.OK and np.allclose(d, [-1, -2, -1, -2], atol=1e-10)
    for retained, n in [(0, 0), (0, 2), (2, 2)]:
        b = O.Schur(retained, 2, 2)
        assert b.prepare(n) == O.OK
        b.clear()
        if n:
            assert b.accumulate([1.0], [(0, j)]) == O.OK
        assert b.prepare_damping(0.1) == O.OK
        assert b.solve_damped(0.1)[0] == O.OK


def test_damped_products_match_dense():
    # src/lsmr.rs:617-728 against dense products (tests/golden)
    g = GOLDEN["lsmr_products"]
    J = np.array(g["J"])
    s = O.Schur(0, 1, 3)
    assert s.prepare(3) == O.OK
    for row in [0, 2]:
        assert s.accumulate([0.0, 0.0], [(0, J[row:row + 2])]) == O.OK
    assert s.prepare_damping(g["floor"]) == O.OK
    assert np.allclose(s.scales(), g["scales"], rtol=1e-15, atol=0)
    x, y = np.array(g["x"]), np.array(g["y"])
    for case in g["cases"]:
        fwd, tr = np.array(case["forward"]), np.array(case["transpose"])
        for k in range(2):
            f = s.damped_apply(case["lambda"], case["scaled"], x[:, k])
            t = s.damped_transpose_apply(case["lambda"], case["scaled"], y[:, k])
            assert np.all(np.abs(f - fwd[:, k]) < 1e-12 * np.maximum(np.abs(fwd[:, k]), 1.0))
            assert np.all(np.abs(t - tr[:, k]) < 1e-12 * np.maximum(np.abs(tr[:, k]), 1.0))


@pytest.mark.parametrize("case", [0, 1])
def test_exact_block_elimination_matches_svd(case):
    # src/lsmr.rs:780-900 fixtures solved through a single eliminated block
    g = GOLDEN["lsmr_svd"][case]
    J, r = np.array(g["J"]), np.array(g["r"])
    n = J.shape[1]
    s = O.Schur(0, 1, n)
    assert s.prepare(n) == O.OK
    assert s.accumulate(r, [(0, J)]) == O.OK
    assert s.prepare_damping(g["floor"]) == O.OK
    assert np.allclose(s.scales(), g["scales"], rtol=1e-15, atol=0)
    for lam, expected in zip(g["lambdas"], g["steps"]):
        st, d, pred = s.solve_damped(lam)
        assert st == O.OK
        assert np.all(np.abs(d - expected) < 1e-7 * np.maximum(np.abs(expected), 1.0)), lam
        assert abs(pred - 0.5 * (r @ r - np.sum((J @ d + r) ** 2))) < 1e-7
        if g["name"] == "mixed_scale":
            assert d[3] == 0.0


# ---- bundle adjustment -------------------------------------------------------------------

def independent_objective(p, cams, pts):
    """The published BAL objective evaluated without the factor code (examples/bal.rs:204-219)."""
    R = Rotation.from_rotvec(cams[p.obs_camera, :3]).as_matrix()
    P = np.einsum("nij,nj->ni", R, pts[p.obs_point]) + cams[p.obs_camera, 3:6]
    u, v = -P[:, 0] / P[:, 2], -P[:, 1] / P[:, 2]
    r2 = u * u + v * v
    c = cams[p.obs_camera]
    sc = c[:, 6] * (1 + c[:, 7] * r2 + c[:, 8] * r2 * r2)
    return 0.5 * np.sum((sc * u - p.obs_pixel[:, 0]) ** 2 + (sc * v - p.obs_pixel[:, 1]) ** 2)


def test_layout_follows_camera_stable_rows():
    p = synthetic.generate("ring16", seed=1, n_cameras=5, n_points=40)
    cam_order, cam_off, pt_perm, pt_off = O.BA(p).layout()
    # stable sort by camera (examples/bal.rs:189-201)
    assert np.array_equal(cam_order, np.argsort(p.obs_camera, kind="stable"))
    assert np.array_equal(cam_off, np.concatenate([[0], np.cumsum(np.bincount(p.obs_camera, minlength=5))]))
    rows_point = p.obs_point[cam_order]
    assert np.array_equal(pt_perm, np.argsort(rows_point, kind="stable"))
    assert np.array_equal(pt_off, np.concatenate([[0], np.cumsum(np.bincount(rows_point, minlength=40))]))


def test_lm_converges_and_matches_independent_objective():
    p = synthetic.generate("ring16", seed=0)
    ba = O.BA(p)
    st, c0 = ba.cost()
    assert abs(c0 - independent_objective(p, p.cameras, p.points)) < 1e-8 * c0
    out, trace = ba.lm(options=O.OptimizeOptions(20, 1e-8, 0.0, 0.0), trace_capacity=100)
    assert out.status in (O.OK, O.NO_PROGRESS, O.NO_CONVERGENCE)
    cams, pts = ba.state()
    obj = independent_objective(p, cams, pts)
    assert abs(out.cost - obj) <= 1e-8 * max(obj, 1.0)          # examples/bal.rs:239-244
    # noise 0.5 px: the optimum sits near 0.5 * sigma^2 * (rows - coordinates)
    dof = 2 * p.n_observations - (9 * p.n_cameras + 3 * p.n_points)
    assert 0.5 * 0.25 * dof * 0.8 < out.cost < 0.5 * 0.25 * 2 * p.n_observations
    assert out.linearizations == out.accepted_steps + 1
    assert out.attempts == out.accepted_steps + out.rejected_steps
    assert np.all(np.diff(trace[:, 0]) < 0)
    assert np.all(np.diff(trace[:, 1]) >= 1) and trace[-1, 1] <= out.attempts
    assert len(trace) == out.accepted_steps


def test_ba_schur_solve_matches_dense_augmented_lstsq():
    # brute force: assemble J column by column and solve [J; sqrt(l) D] d = [-r; 0] densely
    p = synthetic.generate("ring16", seed=2, n_cameras=4, n_points=60)
    ba = O.BA(p)
    st, s = O.ba_schur(ba, tol=1e-14)
    assert st == O.OK and s.prepare_damping(1e-6) == O.OK
    n, rows = s.dimension(), s.rows()
    J = np.stack([s.jacobian_apply(e) for e in np.eye(n)], axis=1)
    r = -s.negative_residual()
    assert np.allclose(J.T @ r, s.gradient(), rtol=1e-12, atol=1e-9)
    D = s.scales()
    assert np.allclose(D, np.maximum(np.linalg.norm(J, axis=0), 1e-6), rtol=1e-14)
    for lam in [1e-4, 1e-2, 1.0]:
        st, d, pred = s.solve_damped(lam)
        assert st == O.OK
        A = np.vstack([J, np.sqrt(lam) * np.diag(D)])
        x, *_ = np.linalg.lstsq(A, np.concatenate([-r, np.zeros(n)]), rcond=None)
        assert np.linalg.norm(d - x) < 1e-7 * np.linalg.norm(x), lam
        assert abs(pred - 0.5 * (r @ r - np.sum((J @ d + r) ** 2))) < 1e-7 * abs(pred)
    # reduced operator = Schur complement of the dense damped normal matrix
    lam = 0.1
    H = J.T @ J / np.outer(D, D) + lam * np.eye(n)
    nr = 9 * p.n_cameras
    S = H[:nr, :nr] - H[:nr, nr:] @ np.linalg.solve(H[nr:, nr:], H[nr:, :nr])
    x = np.random.default_rng(0).normal(size=nr)
    st, y = s.reduced_apply(lam, x)
    assert st == O.OK and np.allclose(y, S @ x, rtol=1e-9, atol=1e-9 * np.abs(S @ x).max())


def test_huber_cost_and_weights():
    # src/factors.rs:42-52 and tests/robust.rs:28-38: r = (3, 4), delta = 1 -> cost 4.5,
    # weight 1/5 per observation block, emitted rows scaled by sqrt(1/5)
    p = synthetic.BalProblem(name="t", cameras=np.array([[0, 0, 0, 0, 0, 0, 1.0, 0, 0]]),
                             points=np.array([[0.0, 0.0, -1.0]]), obs_camera=np.array([0]),
                             obs_point=np.array([0]), obs_pixel=np.array([[-3.0, -4.0]]), huber=1.0)
    ba = O.BA(p, huber=1.0)
    st, c = ba.cost()
    assert st == O.OK and abs(c - 4.5) < 1e-15
    st, s = O.ba_schur(ba)
    assert np.allclose(s.negative_residual(), -np.sqrt(0.2) * np.array([3.0, 4.0]), rtol=1e-15)
    ba2 = O.BA(p, huber=10.0)
    assert abs(ba2.cost()[1] - 12.5) < 1e-15
