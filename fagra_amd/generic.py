"""The GPU Schur backend for models other than the BAL evaluator.

:class:`GenericSchur` mirrors ``Schur::new(retained, retained_width, eliminated_width)`` and its
``LeastSquaresBackend`` methods (src/schur.rs:116-656) over the C ABI ``fg_gschur_*``: a host
model emits its residuals and Jacobian blocks (``accumulate``, one emission as in the reference
or a batch of same-shape emissions), and the damped systems are solved on the device.

:func:`optimize` / :func:`optimize_gn` run the library's LM / GN loops (``fg_lm_optimize`` /
``fg_gn_optimize``) over such a host model with this backend. The model protocol:

* ``dimension``: number of scalar coordinates;
* ``cost(trial) -> (status, value)``: total cost of the accepted or the staged trial state;
* ``linearize(backend) -> status``: emit every residual block into ``backend.accumulate*``;
* ``stage(delta) -> status``: form the trial state from the accepted one and the step;
* ``accept()`` / ``reject()``: keep or drop the staged trial.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import OK, SchurDiagnostics, check


def _lib_handle():
    return _lib.load()


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class GenericSchur:
    """Schur backend over explicit emissions (FP64, block-Jacobi PCG, block widths <= 16)."""

    def __init__(self, retained, retained_width, eliminated_width, max_iterations=None, relative_tolerance=None,
                 on_solve=None):
        lib = _lib_handle()
        cfg = _lib.SchurConfig()
        check(lib.fg_schur_config_default(ctypes.byref(cfg), _lib.FP64), "fg_schur_config_default")
        if max_iterations is not None:
            cfg.max_iterations = max_iterations
        if relative_tolerance is not None:
            cfg.relative_tolerance = relative_tolerance
        self._on_solve = None
        if on_solve is not None:
            self._on_solve = _lib.ON_SOLVE(lambda d, u: on_solve(d.contents))
            cfg.on_solve = self._on_solve
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(lib.fg_gschur_create(ctypes.byref(cfg), retained, retained_width, eliminated_width, ctypes.byref(h)),
              "fg_gschur_create")
        self.h = h.value
        self.retained, self.retained_width, self.eliminated_width = retained, retained_width, eliminated_width
        self.dimension = 0

    def _field(name):
        def get(self):
            return getattr(self.cfg, name)

        def put(self, v):
            setattr(self.cfg, name, v)
            check(_lib_handle().fg_gschur_set_config(self.h, ctypes.byref(self.cfg)), "fg_gschur_set_config")
        return property(get, put)

    max_iterations = _field("max_iterations")
    relative_tolerance = _field("relative_tolerance")
    del _field

    def close(self):
        if getattr(self, "h", None):
            _lib_handle().fg_gschur_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # LeastSquaresBackend mirror
    def prepare(self, dimension):
        self.dimension = dimension
        return _lib_handle().fg_gschur_prepare(self.h, dimension)

    def clear(self):
        _lib_handle().fg_gschur_clear(self.h)

    def accumulate(self, residual, blocks):
        """One emission: residual (rows,), blocks [(column offset, rows x width matrix), ...]."""
        r = np.atleast_1d(np.asarray(residual, dtype=np.float64))
        mats = [np.asarray(m, dtype=np.float64).reshape(len(r), -1) for _, m in blocks]
        cols = np.array([c for c, _ in blocks], dtype=np.uintp).reshape(1, len(blocks))
        return self.accumulate_batch(r[None, :], cols, [m[None] for m in mats])

    def accumulate_batch(self, residuals, cols, values):
        """n emissions of the same shape: residuals (n, rows), cols (n, k) column offsets,
        values: k arrays of shape (n, rows, width_k)."""
        res = np.ascontiguousarray(residuals, dtype=np.float64)
        n, rows = res.shape
        cols = np.ascontiguousarray(np.asarray(cols, dtype=np.uintp).reshape(n, -1))
        k = cols.shape[1]
        widths = np.array([np.asarray(v).shape[2] for v in values], dtype=np.uintp)
        if k != len(values):
            raise ValueError("one value array per block column")
        # per emission: block after block, each column-major rows x width
        flat = np.concatenate([np.asarray(v, dtype=np.float64).transpose(0, 2, 1).reshape(n, -1) for v in values],
                              axis=1) if k else np.zeros((n, 0))
        flat = np.ascontiguousarray(flat)
        szp = ctypes.POINTER(ctypes.c_size_t)
        return _lib_handle().fg_gschur_accumulate(self.h, n, rows, _dp(res), k, cols.ctypes.data_as(szp),
                                                  widths.ctypes.data_as(szp), _dp(flat))

    def gradient_norm(self):
        out = ctypes.c_double()
        st = _lib_handle().fg_gschur_gradient_norm(self.h, ctypes.byref(out))
        return st, out.value

    def prepare_damping(self, min_column_norm):
        return _lib_handle().fg_gschur_prepare_damping(self.h, float(min_column_norm))

    def solve_damped(self, lam):
        ptr = ctypes.POINTER(ctypes.c_double)()
        pred = ctypes.c_double()
        st = _lib_handle().fg_gschur_solve_damped(self.h, float(lam), ctypes.byref(ptr), ctypes.byref(pred))
        if st != OK:
            return st, None, None
        return st, np.ctypeslib.as_array(ptr, shape=(self.dimension,)).copy() if self.dimension else np.zeros(0), \
            pred.value

    def solve(self):
        ptr = ctypes.POINTER(ctypes.c_double)()
        st = _lib_handle().fg_gschur_solve(self.h, ctypes.byref(ptr))
        if st != OK:
            return st, None
        return st, np.ctypeslib.as_array(ptr, shape=(self.dimension,)).copy() if self.dimension else np.zeros(0)

    def diagnostics(self):
        d = SchurDiagnostics()
        _lib_handle().fg_gschur_diagnostics_get(self.h, ctypes.byref(d))
        return d

    def statistics(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        _lib_handle().fg_gschur_statistics(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def last_step_norm(self):
        return _lib_handle().fg_gschur_last_step_norm(self.h)

    def covariance(self, blocks, tolerance=None):
        """Selected covariance (offset, width blocks in selection order) of the accumulated model."""
        from . import _covariance
        return _covariance(_lib_handle().fg_gschur_covariance, self.h, blocks, tolerance)


class _Callbacks:
    """fg_problem callbacks joining a host model and a GenericSchur backend."""

    def __init__(self, model, backend):
        self.model, self.backend = model, backend
        self.dimension = model.dimension
        self.delta = None

    def prepare(self):
        return self.backend.prepare(self.model.dimension)

    def cost(self, trial):
        return self.model.cost(trial)

    def linearize(self):
        self.backend.clear()
        return self.model.linearize(self.backend)

    def gradient_norm(self):
        return self.backend.gradient_norm()

    def prepare_damping(self, floor):
        return self.backend.prepare_damping(floor)

    def solve_damped(self, lam):
        st, d, pred = self.backend.solve_damped(lam)
        if st != OK:
            return st, 0.0, 0.0
        self.delta = d
        return st, self.backend.last_step_norm(), pred

    def solve(self):
        st, d = self.backend.solve()
        if st != OK:
            return st, 0.0
        self.delta = d
        return st, self.backend.last_step_norm()

    def stage(self):
        return self.model.stage(self.delta)

    def accept(self):
        self.model.accept()

    def reject(self):
        self.model.reject()

    def linear_work(self):
        return self.backend.statistics()


def optimize(model, backend, options=None, lm_options=None, on_accept=None):
    """LevenbergMarquardt::optimize (src/optimization/levenberg_marquardt.rs:330-472) over a
    host model with the GPU generic Schur backend; returns the LmOutcome."""
    from . import lm_optimize
    return lm_optimize(_Callbacks(model, backend), options=options, lm_options=lm_options, on_accept=on_accept)


def optimize_gn(model, backend, options=None):
    """GaussNewton::optimize (src/optimization/gauss_newton.rs:101-176) with the same protocol."""
    from . import gn_optimize
    return gn_optimize(_Callbacks(model, backend), options=options)
