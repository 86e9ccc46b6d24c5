"""fg_dense_covariance_noalloc against fg_dense_covariance, both through the C ABI (GPU only).

The no-alloc entry point runs the same kernels on a file-scope device scratch array instead of
allocated buffers, so on the same inputs its covariance equals the allocating entry point's bit
for bit. Two calls in a row (the second smaller) check that nothing carries over between calls
that share the scratch; the error cases and the size limit are checked too.
"""
import ctypes

import numpy as np
import pytest

import fagra_amd as fa

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)
_szp = ctypes.POINTER(ctypes.c_size_t)


def call(name, H, cols, tolerance=None):
    h = np.asfortranarray(H, dtype=np.float64)
    c = np.ascontiguousarray(cols, dtype=np.uintp)
    k = len(c)
    out = np.full((k, k), np.nan, order="F")
    st = getattr(fa.lib(), name)(h.ctypes.data_as(_dp), h.shape[0], c.ctypes.data_as(_szp), k,
                                 ctypes.byref(fa.covariance_options(tolerance)), out.ctypes.data_as(_dp))
    return st, out


def spd(n, seed):
    a = np.random.default_rng(seed).normal(size=(n + 16, n))
    return a.T @ a


def test_noalloc_matches_allocating_entry_point_on_two_calls():
    for n, k, seed in [(700, 40, 1), (130, 7, 2)]:
        H = spd(n, seed)
        cols = np.random.default_rng(seed + 10).permutation(n)[:k]
        print(f"noalloc: n {n}, k {k}", flush=True)
        st_a, c_a = call("fg_dense_covariance", H, cols)
        st_b, c_b = call("fg_dense_covariance_noalloc", H, cols)
        assert st_a == st_b == fa.OK
        assert np.array_equal(c_a, c_b)
        ref = np.linalg.inv(H)[np.ix_(cols, cols)]
        assert np.linalg.norm(c_b - ref) <= 1e-8 * np.linalg.norm(ref)


def test_noalloc_errors_and_limits():
    H = spd(96, 3)
    sing = H.copy()
    sing[:, 5] = sing[:, 4]
    sing[5, :] = sing[4, :]
    bad = H.copy()
    bad[10, 3] = np.nan
    for M, cols in [(sing, [0, 1]), (bad, [2]), (H, [3, 3]), (H, [0, 96]), (H, [])]:
        st_a, _ = call("fg_dense_covariance", M, cols)
        st_b, _ = call("fg_dense_covariance_noalloc", M, cols)
        print(f"noalloc: error case status {st_a} / {st_b}", flush=True)
        assert st_a == st_b
    assert call("fg_dense_covariance", H, [1], tolerance=1.5)[0] == fa.INVALID_RANK_TOLERANCE
    assert call("fg_dense_covariance_noalloc", H, [1], tolerance=1.5)[0] == fa.INVALID_RANK_TOLERANCE
    big = np.eye(2049)
    assert call("fg_dense_covariance_noalloc", big, [0])[0] == fa.INVALID_OPTIONS
    assert call("fg_dense_covariance_noalloc", spd(300, 4), np.arange(257))[0] == fa.INVALID_OPTIONS
    # a good call after the rejected ones still matches
    cols = [5, 17, 40]
    st_a, c_a = call("fg_dense_covariance", H, cols)
    st_b, c_b = call("fg_dense_covariance_noalloc", H, cols)
    assert st_a == st_b == fa.OK and np.array_equal(c_a, c_b)
