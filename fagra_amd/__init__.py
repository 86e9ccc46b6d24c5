"""fagra_amd -- MI355X (gfx950) bundle-adjustment backend for fagra's LM/GN inner loop.

Python mirror of the reference interfaces over the C ABI in ``include/fagra_gpu.h``:

* :class:`BalProblem` -- the device-resident BAL Frame batch evaluator
  (examples/bal/model.rs:194-237) with trial/accept state buffers (src/storage.rs:282-318).
* :class:`Schur` -- the GPU Schur-PCG ``LeastSquaresBackend`` (src/schur.rs:116-546).
* :class:`Lsmr` -- the GPU LSMR ``LeastSquaresBackend`` (src/lsmr.rs:273-595).
* :class:`LevenbergMarquardt` / :class:`GaussNewton` -- ``solve_batch``
  (src/optimization/levenberg_marquardt.rs:252-298, src/optimization/gauss_newton.rs:63-74).

All numerical work runs in the HIP library; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (CG_CONVERGED, CG_ITERATION_LIMIT, CG_NON_FINITE, CG_NONE, CG_NOT_POSITIVE_DEFINITE,
                   DEVICE_ERROR, DIMENSION_MISMATCH, FP32, FP64, INVALID_EMISSION, INVALID_EVALUATION,
                   INVALID_OPTIONS, LINEAR_SOLVE_FAILED, NO_CONVERGENCE, NO_PROGRESS, OK, PRECOND_BLOCK_JACOBI,
                   PRECOND_LEARNED, REJ_INVALID_EVALUATION, REJ_LINEAR_SOLVE, REJ_NONE, REJ_POOR_AGREEMENT,
                   REJ_PREDICTION, TERM_COST, TERM_GRADIENT, TERM_NO_VARIABLES, TERM_NONE, TERM_STEP,
                   UNSUPPORTED_STRUCTURE, INVALID_INPUT, DUPLICATE_COVARIANCE_BLOCK, SINGULAR_INFORMATION,
                   INVALID_INFORMATION, INVALID_RANK_TOLERANCE, UNKNOWN_KEY, FagraError,
                   LSMR_CONVERGED, LSMR_ITERATION_LIMIT, LSMR_NON_FINITE, LSMR_NONE, LmOptions, LmOutcome,
                   LsmrDiagnostics, LsmrTimings, OptimizeOptions, SchurDiagnostics, SchurTimings, check)

__all__ = ["BalProblem", "Schur", "Lsmr", "LevenbergMarquardt", "GaussNewton", "OptimizeOptions", "LmOptions",
           "LmOutcome", "FagraError", "FP32", "FP64", "device_info", "lm_optimize", "gn_optimize"]


def lib():
    return _lib.load()


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_optimize_options():
    o = OptimizeOptions()
    lib().fg_optimize_options_default(ctypes.byref(o))
    return o


def default_lm_options():
    o = LmOptions()
    lib().fg_lm_options_default(ctypes.byref(o))
    return o


def covariance_options(relative_pivot_tolerance=None):
    o = _lib.CovarianceOptions()
    lib().fg_covariance_options_default(ctypes.byref(o))
    if relative_pivot_tolerance is not None:
        o.relative_pivot_tolerance = relative_pivot_tolerance
    return o


def _selection(blocks):
    offs = np.array([int(o) for o, _ in blocks], dtype=np.uintp)
    wids = np.array([int(w) for _, w in blocks], dtype=np.uintp)
    return offs, wids, int(wids.sum()) if len(wids) else 0


def _covariance(fn, handle, blocks, tolerance):
    """Selected covariance through fn (fg_schur_covariance / fg_gschur_covariance): blocks are
    (offset, width) variable blocks in selection order; returns (status, k x k matrix)."""
    offs, wids, k = _selection(blocks)
    out = np.zeros((k, k), order="F")
    szp = ctypes.POINTER(ctypes.c_size_t)
    st = fn(handle, offs.ctypes.data_as(szp), wids.ctypes.data_as(szp), len(offs),
            ctypes.byref(covariance_options(tolerance)), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return st, (out if st == OK else None)


def dense_covariance(information, columns, tolerance=None):
    """CovarianceBackend for DenseNormalCholesky (src/covariance.rs:275-283) on the device:
    the inverse of a dense information matrix at the selected coordinates."""
    h = np.asfortranarray(information, dtype=np.float64)
    n = h.shape[0]
    cols = np.ascontiguousarray(columns, dtype=np.uintp)
    out = np.zeros((len(cols), len(cols)), order="F")
    st = lib().fg_dense_covariance(h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                   cols.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), len(cols),
                                   ctypes.byref(covariance_options(tolerance)),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return st, (out if st == OK else None)


def device_info():
    name = ctypes.create_string_buffer(64)
    cus = ctypes.c_int()
    lds = ctypes.c_size_t()
    check(lib().fg_device_info(name, 64, ctypes.byref(cus), ctypes.byref(lds)), "fg_device_info")
    return {"arch": name.value.decode(), "compute_units": cus.value, "lds_bytes_per_block": lds.value}


class BalProblem:
    """Device copy of a BAL problem (cameras, points, observations) and its evaluator."""

    def __init__(self, problem, precision=FP64, huber=None):
        self.nc, self.np, self.no = problem.n_cameras, problem.n_points, problem.n_observations
        self.precision = precision
        h = problem.huber if huber is None else huber
        cams, pts = _f64(problem.cameras), _f64(problem.points)
        oc = np.ascontiguousarray(problem.obs_camera, dtype=np.int64)
        op = np.ascontiguousarray(problem.obs_point, dtype=np.int64)
        px = _f64(problem.obs_pixel)
        handle = ctypes.c_void_p()
        i64 = ctypes.POINTER(ctypes.c_int64)
        check(lib().fg_ba_create(self.nc, self.np, self.no, _dptr(cams), _dptr(pts), oc.ctypes.data_as(i64),
                                 op.ctypes.data_as(i64), _dptr(px), float(h), int(precision), ctypes.byref(handle)),
              "fg_ba_create")
        self.h = handle.value

    def close(self):
        if getattr(self, "h", None):
            lib().fg_ba_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dimension(self):
        return 9 * self.nc + 3 * self.np

    def layout(self):
        a = np.zeros(self.no, np.int64)
        b = np.zeros(self.nc + 1, np.int64)
        c = np.zeros(self.no, np.int64)
        d = np.zeros(self.np + 1, np.int64)
        i64 = ctypes.POINTER(ctypes.c_int64)
        check(lib().fg_ba_layout(self.h, *[x.ctypes.data_as(i64) for x in (a, b, c, d)]), "fg_ba_layout")
        return a, b, c, d

    def state(self):
        c = np.zeros((self.nc, 9))
        p = np.zeros((self.np, 3))
        check(lib().fg_ba_get_state(self.h, _dptr(c), _dptr(p)), "fg_ba_get_state")
        return c, p

    def set_state(self, cameras, points):
        c, p = _f64(cameras), _f64(points)
        check(lib().fg_ba_set_state(self.h, _dptr(c), _dptr(p)), "fg_ba_set_state")

    def cost(self, trial=False):
        out = ctypes.c_double()
        st = lib().fg_ba_cost(self.h, int(trial), ctypes.byref(out))
        return st, out.value

    def stage(self, delta):
        d = _f64(delta)
        return lib().fg_ba_stage(self.h, _dptr(d))

    def accept(self):
        lib().fg_ba_accept(self.h)

    def reject(self):
        lib().fg_ba_reject(self.h)


class Schur:
    """GPU Schur backend with the reference's public fields (src/schur.rs:116-125)."""

    def __init__(self, precision=FP64, max_iterations=None, relative_tolerance=None,
                 preconditioner=PRECOND_BLOCK_JACOBI, collect_timings=False, on_solve=None, network=None):
        cfg = _lib.SchurConfig()
        check(lib().fg_schur_config_default(ctypes.byref(cfg), int(precision)), "fg_schur_config_default")
        if max_iterations is not None:
            cfg.max_iterations = max_iterations
        if relative_tolerance is not None:
            cfg.relative_tolerance = relative_tolerance
        cfg.preconditioner = preconditioner
        cfg.collect_timings = int(collect_timings)
        self._on_solve = None
        if on_solve is not None:
            self._on_solve = _lib.ON_SOLVE(lambda d, u: on_solve(d.contents))
            cfg.on_solve = self._on_solve
        self.cfg = cfg
        handle = ctypes.c_void_p()
        check(lib().fg_schur_create(ctypes.byref(cfg), ctypes.byref(handle)), "fg_schur_create")
        self.h = handle.value
        self.problem = None
        self._net = None
        if network is None and preconditioner == PRECOND_LEARNED:
            from . import learned
            if learned.available():
                network = learned.Network.load()
        if network is not None:
            self.set_network(network)

    def set_network(self, network):
        """Attach a learned.Network for PRECOND_LEARNED (kept alive with this backend)."""
        check(lib().fg_schur_set_mlp(self.h, network.h), "fg_schur_set_mlp")
        self._net = network

    # public fields mirrored into the config on every prepare
    @property
    def max_iterations(self):
        return self.cfg.max_iterations

    @max_iterations.setter
    def max_iterations(self, v):
        self.cfg.max_iterations = v
        self._push()

    @property
    def relative_tolerance(self):
        return self.cfg.relative_tolerance

    @relative_tolerance.setter
    def relative_tolerance(self, v):
        self.cfg.relative_tolerance = v
        self._push()

    @property
    def preconditioner(self):
        return self.cfg.preconditioner

    @preconditioner.setter
    def preconditioner(self, v):
        self.cfg.preconditioner = v
        self._push()

    def _push(self):
        check(lib().fg_schur_set_config(self.h, ctypes.byref(self.cfg)), "fg_schur_set_config")

    def close(self):
        if getattr(self, "h", None):
            lib().fg_schur_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # LeastSquaresBackend mirror
    def prepare(self, problem, dimension=None):
        self.problem = problem
        return lib().fg_schur_prepare(self.h, problem.h, problem.dimension if dimension is None else dimension)

    def clear(self):
        lib().fg_schur_clear(self.h)

    def linearize(self):
        return lib().fg_schur_linearize(self.h)

    def gradient_norm(self):
        out = ctypes.c_double()
        st = lib().fg_schur_gradient_norm(self.h, ctypes.byref(out))
        return st, out.value

    def prepare_damping(self, min_column_norm):
        return lib().fg_schur_prepare_damping(self.h, float(min_column_norm))

    def solve_damped(self, lam):
        ptr = ctypes.POINTER(ctypes.c_double)()
        pred = ctypes.c_double()
        st = lib().fg_schur_solve_damped(self.h, float(lam), ctypes.byref(ptr), ctypes.byref(pred))
        if st != OK:
            return st, None, None
        return st, np.ctypeslib.as_array(ptr, shape=(self.problem.dimension,)).copy(), pred.value

    def solve(self):
        ptr = ctypes.POINTER(ctypes.c_double)()
        st = lib().fg_schur_solve(self.h, ctypes.byref(ptr))
        if st != OK:
            return st, None
        return st, np.ctypeslib.as_array(ptr, shape=(self.problem.dimension,)).copy()

    def statistics(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        lib().fg_schur_statistics(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def diagnostics(self):
        d = SchurDiagnostics()
        lib().fg_schur_diagnostics_get(self.h, ctypes.byref(d))
        return d

    def timings(self):
        t = SchurTimings()
        lib().fg_schur_timings_get(self.h, ctypes.byref(t))
        return t

    def covariance(self, blocks, tolerance=None):
        """Selected covariance of the current linearization (src/schur.rs:548-556)."""
        return _covariance(lib().fg_schur_covariance, self.h, blocks, tolerance)

    # parity access
    def model(self):
        n, rows = self.problem.dimension, 2 * self.problem.no
        g, d, r = np.zeros(n), np.zeros(n), np.zeros(rows)
        check(lib().fg_schur_model(self.h, _dptr(g), _dptr(d), _dptr(r)), "fg_schur_model")
        return g, d, r

    def camera_gram(self):
        out = np.zeros((self.problem.nc, 9, 9))
        check(lib().fg_schur_camera_gram(self.h, _dptr(out)), "fg_schur_camera_gram")
        return out

    def time_products(self, lam, x, repeats=50):
        """Mean seconds per launch of the point-major and camera-major product kernels."""
        x = _f64(x)
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().fg_schur_time_products(self.h, float(lam), _dptr(x), int(repeats), ctypes.byref(a),
                                           ctypes.byref(b)), "fg_schur_time_products")
        return a.value, b.value

    def record_mismatches(self):
        """Observations whose two record copies differ (0 after a consistent linearization)."""
        return lib().fg_schur_record_mismatches(self.h)

    def reduced_apply(self, lam, x):
        x = _f64(x)
        y = np.zeros_like(x)
        st = lib().fg_schur_reduced_apply(self.h, float(lam), _dptr(x), _dptr(y))
        return st, y

    def jacobian_apply(self, x):
        x = _f64(x)
        y = np.zeros(2 * self.problem.no)
        check(lib().fg_schur_jacobian_apply(self.h, _dptr(x), _dptr(y)), "fg_schur_jacobian_apply")
        return y

    def jacobian_transpose_apply(self, y):
        y = _f64(y)
        x = np.zeros(self.problem.dimension)
        check(lib().fg_schur_jacobian_transpose_apply(self.h, _dptr(y), _dptr(x)), "fg_schur_jacobian_transpose_apply")
        return x

    def damped_apply(self, lam, scaled, x):
        x = _f64(x)
        y = np.zeros(2 * self.problem.no + self.problem.dimension)
        check(lib().fg_schur_damped_apply(self.h, float(lam), int(scaled), _dptr(x), _dptr(y)), "damped_apply")
        return y

    def damped_transpose_apply(self, lam, scaled, y):
        y = _f64(y)
        x = np.zeros(self.problem.dimension)
        check(lib().fg_schur_damped_transpose_apply(self.h, float(lam), int(scaled), _dptr(y), _dptr(x)),
              "damped_transpose_apply")
        return x


class Lsmr:
    """GPU LSMR backend with the reference's public fields (src/lsmr.rs:273-295).

    Damped solves run right-preconditioned LSMR on [J D^-1; sqrt(lambda) I] (diagonal or
    block-Jacobi preconditioning) or [J; sqrt(lambda) D] (identity); the undamped solve runs
    identity-preconditioned LSMR on J.
    """

    _lm_entry = "fg_lm_solve_batch_lsmr"
    _gn_entry = "fg_gn_solve_batch_lsmr"

    def __init__(self, precision=FP64, max_iterations=None, relative_tolerance=None,
                 diagonal_preconditioning=True, block_preconditioning=False, collect_timings=False,
                 on_solve=None):
        cfg = _lib.LsmrConfig()
        check(lib().fg_lsmr_config_default(ctypes.byref(cfg), int(precision)), "fg_lsmr_config_default")
        if max_iterations is not None:
            cfg.max_iterations = max_iterations
        if relative_tolerance is not None:
            cfg.relative_tolerance = relative_tolerance
        cfg.diagonal_preconditioning = int(diagonal_preconditioning)
        cfg.block_preconditioning = int(block_preconditioning)
        cfg.collect_timings = int(collect_timings)
        self._on_solve = None
        if on_solve is not None:
            self._on_solve = _lib.ON_LSMR_SOLVE(lambda d, u: on_solve(d.contents))
            cfg.on_solve = self._on_solve
        self.cfg = cfg
        handle = ctypes.c_void_p()
        check(lib().fg_lsmr_create(ctypes.byref(cfg), ctypes.byref(handle)), "fg_lsmr_create")
        self.h = handle.value
        self.problem = None

    def _field(name):
        def get(self):
            return getattr(self.cfg, name)

        def put(self, v):
            setattr(self.cfg, name, v)
            check(lib().fg_lsmr_set_config(self.h, ctypes.byref(self.cfg)), "fg_lsmr_set_config")
        return property(get, put)

    max_iterations = _field("max_iterations")
    relative_tolerance = _field("relative_tolerance")
    diagonal_preconditioning = _field("diagonal_preconditioning")
    block_preconditioning = _field("block_preconditioning")
    del _field

    def close(self):
        if getattr(self, "h", None):
            lib().fg_lsmr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # LeastSquaresBackend mirror
    def prepare(self, problem, dimension=None):
        self.problem = problem
        return lib().fg_lsmr_prepare(self.h, problem.h, problem.dimension if dimension is None else dimension)

    def clear(self):
        lib().fg_lsmr_clear(self.h)

    def linearize(self):
        return lib().fg_lsmr_linearize(self.h)

    def gradient_norm(self):
        out = ctypes.c_double()
        st = lib().fg_lsmr_gradient_norm(self.h, ctypes.byref(out))
        return st, out.value

    def prepare_damping(self, min_column_norm):
        return lib().fg_lsmr_prepare_damping(self.h, float(min_column_norm))

    def solve_damped(self, lam):
        ptr = ctypes.POINTER(ctypes.c_double)()
        pred = ctypes.c_double()
        st = lib().fg_lsmr_solve_damped(self.h, float(lam), ctypes.byref(ptr), ctypes.byref(pred))
        if st != OK:
            return st, None, None
        return st, np.ctypeslib.as_array(ptr, shape=(self.problem.dimension,)).copy(), pred.value

    def solve(self):
        ptr = ctypes.POINTER(ctypes.c_double)()
        st = lib().fg_lsmr_solve(self.h, ctypes.byref(ptr))
        if st != OK:
            return st, None
        return st, np.ctypeslib.as_array(ptr, shape=(self.problem.dimension,)).copy()

    def statistics(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        lib().fg_lsmr_statistics(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def diagnostics(self):
        d = LsmrDiagnostics()
        lib().fg_lsmr_diagnostics_get(self.h, ctypes.byref(d))
        return d

    def timings(self):
        t = LsmrTimings()
        lib().fg_lsmr_timings_get(self.h, ctypes.byref(t))
        return t


class LevenbergMarquardt:
    """LevenbergMarquardt<Schur | Lsmr> (src/optimization/levenberg_marquardt.rs:205-317)."""

    def __init__(self, backend):
        self.backend = backend
        self.options = default_lm_options()
        self.on_accept = None
        self._statistics = LmOutcome()

    def solve_batch(self, problem, options=None):
        """Returns the LmOutcome (status, statistics and work of this call)."""
        options = options or default_optimize_options()
        out = LmOutcome()
        cb = _lib.ON_ACCEPT(lambda st, u: self.on_accept(st.contents)) if self.on_accept else _lib.ON_ACCEPT()
        self.backend.problem = problem
        entry = getattr(lib(), getattr(self.backend, "_lm_entry", "fg_lm_solve_batch"))
        entry(self.backend.h, problem.h, ctypes.byref(options), ctypes.byref(self.options), cb, None,
              ctypes.byref(out))
        self._statistics = out
        return out

    def statistics(self):
        return self._statistics


class GaussNewton:
    """GaussNewton<Schur | Lsmr> (src/optimization/gauss_newton.rs:21-177)."""

    def __init__(self, backend):
        self.backend = backend

    def solve_batch(self, problem, options=None):
        options = options or default_optimize_options()
        out = LmOutcome()
        self.backend.problem = problem
        entry = getattr(lib(), getattr(self.backend, "_gn_entry", "fg_gn_solve_batch"))
        entry(self.backend.h, problem.h, ctypes.byref(options), ctypes.byref(out))
        return out


def _problem_struct(cbs):
    """Wrap a Python callback object into the fg_problem struct (kept alive by the caller)."""
    p = _lib.Problem()
    p.ctx = None
    p.dimension = cbs.dimension
    keep = [
        _lib.FN_PREPARE(lambda c: cbs.prepare()),
        _lib.FN_COST(lambda c, trial, out: _store(out, *cbs.cost(bool(trial)))),
        _lib.FN_LIN(lambda c: cbs.linearize()),
        _lib.FN_GN(lambda c, out: _store(out, *cbs.gradient_norm())),
        _lib.FN_PD(lambda c, f: cbs.prepare_damping(f)),
        _lib.FN_SD(lambda c, lam, norm, pred: _store2(norm, pred, *cbs.solve_damped(lam))),
        _lib.FN_S(lambda c, norm: _store(norm, *cbs.solve())),
        _lib.FN_STAGE(lambda c: cbs.stage()),
        _lib.FN_VOID(lambda c: cbs.accept()),
        _lib.FN_VOID(lambda c: cbs.reject()),
        _lib.FN_WORK(lambda c, a, b: _store_work(a, b, cbs.linear_work())),
    ]
    (p.prepare, p.cost, p.linearize, p.gradient_norm, p.prepare_damping, p.solve_damped, p.solve, p.stage,
     p.accept, p.reject, p.linear_work) = keep
    return p, keep


def _store(ptr, status, value=0.0):
    ptr[0] = value
    return status


def _store2(p1, p2, status, v1=0.0, v2=0.0):
    p1[0] = v1
    p2[0] = v2
    return status


def _store_work(a, b, work):
    a[0], b[0] = work


def lm_optimize(callbacks, options=None, lm_options=None, on_accept=None):
    """Run the library's LM loop over Python callbacks (see tests for the protocol)."""
    options = options or default_optimize_options()
    lm_options = lm_options or default_lm_options()
    p, keep = _problem_struct(callbacks)
    out = LmOutcome()
    cb = _lib.ON_ACCEPT(lambda st, u: on_accept(st.contents)) if on_accept else _lib.ON_ACCEPT()
    lib().fg_lm_optimize(ctypes.byref(p), ctypes.byref(options), ctypes.byref(lm_options), cb, None,
                         ctypes.byref(out))
    del keep
    return out


def gn_optimize(callbacks, options=None):
    options = options or default_optimize_options()
    p, keep = _problem_struct(callbacks)
    out = LmOutcome()
    lib().fg_gn_optimize(ctypes.byref(p), ctypes.byref(options), ctypes.byref(out))
    del keep
    return out
