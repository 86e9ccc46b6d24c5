"""Selected covariance on the device (fg_dense_covariance, fg_gschur_covariance,
fg_schur_covariance) against the oracle's restatement and the dense inverse. GPU only.

Tolerances: 1e-12 absolute on the reference's 24-coordinate fixture (its own bound); 1e-10
relative to the largest entry elsewhere (blocked vs column Cholesky summation order).
"""
import numpy as np
import pytest

import fagra_amd as fa
from fagra_amd import generic, synthetic
from oracle import oracle as O
from planar_model import Planar

pytestmark = pytest.mark.gpu


def close(a, b, tol=1e-10):
    return np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1e-300)


def test_reference_fixture_and_random_information():
    J = np.array([[4.0 if i == k else ((i * 7 + k * 11) % 13) / 30.0 for k in range(24)] for i in range(24)])
    H = J.T @ J
    idx = list(range(9, 18)) + list(range(0, 9)) + list(range(18, 21))
    st, C = fa.dense_covariance(H, idx)
    assert st == fa.OK
    assert np.max(np.abs(C - np.linalg.inv(H)[np.ix_(idx, idx)])) < 1e-12
    rng = np.random.default_rng(0)
    # sizes around the 64-column panel and the 64 x 64 update tiles
    for n in [1, 5, 63, 64, 65, 130, 257]:
        A = rng.normal(size=(n + 7, n))
        H = A.T @ A + 1e-3 * np.eye(n)
        k = min(n, 9)
        cols = rng.permutation(n)[:k]
        st, C = fa.dense_covariance(H, cols)
        st_o, C_o = O.dense_covariance(H, cols)
        assert st == st_o == fa.OK, n
        assert close(C, C_o), n
        assert close(C, np.linalg.inv(H)[np.ix_(cols, cols)], 1e-8), n


def test_error_cases_match_the_oracle():
    cases = [(np.zeros((1, 1)), [0]), (np.zeros((1, 1)), []), (np.ones((1, 1)), [0]),
             (np.array([[1.0, 1.0], [1.0, 1.0 + 1e-15]]), [0]), (np.zeros((0, 0)), [])]
    bad = np.eye(70)
    bad[66, 3] = np.nan
    cases.append((bad, [1]))
    up = np.eye(3)
    up[0, 2] = np.inf  # upper triangle is not read
    cases.append((up, [2]))
    # rank-deficient beyond the first panel: the failure surfaces in a later tile
    A = np.random.default_rng(1).normal(size=(100, 90))
    A[:, 80] = A[:, 3]
    cases.append((A.T @ A, [0, 80]))
    for H, cols in cases:
        st, C = fa.dense_covariance(H, cols)
        st_o, C_o = O.dense_covariance(H, cols)
        assert st == st_o, (H.shape, cols, st, st_o)
        if st == fa.OK:
            assert close(C, C_o) if C.size else C.shape == C_o.shape
    for t in [-1.0, 1.0, float("nan"), float("inf")]:
        assert fa.dense_covariance(np.ones((1, 1)), [0], tolerance=t)[0] == fa.INVALID_RANK_TOLERANCE
    assert fa.dense_covariance(np.eye(3), [1, 1])[0] == fa.DUPLICATE_COVARIANCE_BLOCK
    assert fa.dense_covariance(np.eye(3), [3])[0] == fa.UNKNOWN_KEY


def test_generic_model_covariance_matches_oracle_and_dense_inverse():
    m = Planar(seed=5, n_poses=20, n_landmarks=120, n_obs=500)
    g = generic.GenericSchur(3 * m.np_, 3, 2)
    assert g.prepare(m.dimension) == fa.OK and m.linearize(g) == fa.OK
    from planar_model import OracleBackend
    o = OracleBackend(3 * m.np_, 3, 2)
    assert o.prepare(m.dimension) == fa.OK and m.linearize(o) == fa.OK
    blocks = [(3 * 7, 3), (3 * m.np_ + 2 * 11, 2), (0, 3), (3 * m.np_ + 2 * 119, 2)]
    st, C = g.covariance(blocks)
    st_o, C_o = o.s.covariance(blocks)
    assert st == st_o == fa.OK
    assert close(C, C_o)
    # the information is assembled with one owner per entry summing in emission order (no
    # atomics), so repeated calls return the same bits
    for _ in range(3):
        st2, C2 = g.covariance(blocks)
        assert st2 == fa.OK and np.array_equal(C2, C)
    J, _ = m.dense_jacobian()
    idx = [c for off, w in blocks for c in range(off, off + w)]
    assert close(C, np.linalg.inv(J.T @ J)[np.ix_(idx, idx)], 1e-8)
    assert g.covariance([(0, 3), (0, 3)])[0] == fa.DUPLICATE_COVARIANCE_BLOCK
    assert g.covariance([(2, 3)])[0] == fa.UNKNOWN_KEY
    assert g.covariance([(0, 3)], tolerance=2.0)[0] == fa.INVALID_RANK_TOLERANCE


@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32])
def test_bundle_adjustment_covariance_reports_the_gauge(precision):
    # the free-gauge BAL information is singular: the reference's screen, the oracle and the
    # device agree; selection errors come first
    p = synthetic.generate("ring16", seed=3, n_cameras=4, n_points=60, huber=0.0)
    st, s_o = O.ba_schur(O.BA(p))
    ba = fa.BalProblem(p, precision=precision)
    s = fa.Schur(precision=precision)
    assert s.prepare(ba) == fa.OK and s.linearize() == fa.OK
    for blocks in ([(0, 9), (36, 3)], []):
        assert s.covariance(blocks)[0] == s_o.covariance(blocks)[0] == fa.SINGULAR_INFORMATION
    assert s.covariance([(9, 9), (9, 9)])[0] == fa.DUPLICATE_COVARIANCE_BLOCK
    assert s.covariance([(37, 3)])[0] == fa.UNKNOWN_KEY
