"""The oracle's selected covariance (src/covariance.rs, src/normal.rs:341-392,
src/linear_model.rs:122-167) pinned by the reference's own covariance fixtures, the dense
inverse, and the error cases of tests/covariance.rs."""
import numpy as np
import pytest

from fagra_amd import synthetic
from oracle import oracle as O
from planar_model import OracleBackend, Planar


def reference_24():
    # src/covariance.rs:330-391: 24 coordinates, 4 on the diagonal, ((7i + 11k) mod 13) / 30 off it
    J = np.array([[4.0 if i == k else ((i * 7 + k * 11) % 13) / 30.0 for k in range(24)] for i in range(24)])
    return J.T @ J, list(range(9, 18)) + list(range(0, 9)) + list(range(18, 21))


def test_reference_fixtures():
    H, idx = reference_24()
    st, C = O.dense_covariance(H, idx)
    assert st == O.OK and C.shape == (21, 21)
    assert np.max(np.abs(C - np.linalg.inv(H)[np.ix_(idx, idx)])) < 1e-12
    assert np.array_equal(C, C.T)
    # tests/covariance.rs:45-153: priors on x and q, differences x-y, y-z, z-q; selection [q, z]
    J = np.array([[1, 0, 0, 0], [-1, 1, 0, 0], [0, -1, 1, 0], [0, 0, -1, 1], [0, 0, 0, 1.0]])
    st, C = O.dense_covariance(J.T @ J, [3, 2])
    assert st == O.OK
    assert np.max(np.abs(C - np.linalg.inv(J.T @ J)[np.ix_([3, 2], [3, 2])])) < 1024 * np.finfo(float).eps


def test_reference_error_cases():
    # tests/covariance.rs:158-210: no information is singular (also for an empty selection), a
    # unit prior gives variance 1, tolerances outside [0, 1) are rejected, empty models are fine
    z = np.zeros((1, 1))
    assert O.dense_covariance(z, [0])[0] == O.SINGULAR_INFORMATION
    assert O.dense_covariance(z, [])[0] == O.SINGULAR_INFORMATION
    st, C = O.dense_covariance(np.ones((1, 1)), [0])
    assert st == O.OK and C[0, 0] == 1.0
    for t in [-1.0, 1.0, float("nan"), float("inf")]:
        assert O.dense_covariance(np.ones((1, 1)), [0], tolerance=t)[0] == O.INVALID_RANK_TOLERANCE
    st, C = O.dense_covariance(np.zeros((0, 0)), [])
    assert st == O.OK and C.shape == (0, 0)
    bad = np.eye(3)
    bad[2, 1] = np.nan
    assert O.dense_covariance(bad, [0])[0] == O.INVALID_INFORMATION
    bad = np.eye(3)
    bad[1, 2] = np.nan  # upper triangle: not read
    assert O.dense_covariance(bad, [0])[0] == O.OK
    # rank screen: a pivot whose square is tiny against its diagonal
    H = np.array([[1.0, 1.0], [1.0, 1.0 + 1e-15]])
    assert O.dense_covariance(H, [0])[0] == O.SINGULAR_INFORMATION
    assert O.dense_covariance(H, [0], tolerance=0.0)[0] == O.OK


def test_model_replay_matches_dense_inverse_and_checks_selection():
    m = Planar(seed=4, n_poses=6, n_landmarks=40, n_obs=150)
    b = OracleBackend(3 * m.np_, 3, 2)
    assert b.prepare(m.dimension) == O.OK and m.linearize(b) == O.OK
    J, r = m.dense_jacobian()
    inv = np.linalg.inv(J.T @ J)
    blocks = [(3 * 4, 3), (3 * m.np_ + 2 * 7, 2), (0, 3)]
    st, C = b.s.covariance(blocks)
    idx = list(range(12, 15)) + list(range(3 * m.np_ + 14, 3 * m.np_ + 16)) + [0, 1, 2]
    assert st == O.OK
    assert np.max(np.abs(C - inv[np.ix_(idx, idx)])) < 1e-10 * np.max(np.abs(inv))
    assert b.s.covariance([(0, 3), (0, 3)])[0] == O.DUPLICATE_COVARIANCE_BLOCK
    assert b.s.covariance([(1, 3)])[0] == O.UNKNOWN_KEY
    assert b.s.covariance([(3 * m.np_ + 1, 2)])[0] == O.UNKNOWN_KEY
    assert b.s.covariance([(0, 3)], tolerance=1.0)[0] == O.INVALID_RANK_TOLERANCE


def test_bundle_adjustment_information_is_gauge_singular():
    # the BAL model is invariant to a similarity of all cameras and points (7 dof), so its
    # undamped information has no inverse and the reference's screen reports it
    p = synthetic.generate("ring16", seed=3, n_cameras=4, n_points=60, huber=0.0)
    st, s = O.ba_schur(O.BA(p))
    assert st == O.OK
    assert s.covariance([(0, 9), (36, 3)])[0] == O.SINGULAR_INFORMATION
    assert s.covariance([])[0] == O.SINGULAR_INFORMATION
