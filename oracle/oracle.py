"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (``fg_oracle.cpp``) is an f64 restatement of the reference's bundle-adjustment
path. Its loops over independent blocks run on OpenMP threads with every sum taken in the
sequential order, so results are bit-identical for any thread count (``set_threads``).
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OK, INVALID_OPTIONS, LINEAR_SOLVE_FAILED, NO_CONVERGENCE, NO_PROGRESS = 0, 1, 2, 3, 4
INVALID_EVALUATION, DIMENSION_MISMATCH, INVALID_EMISSION, UNSUPPORTED_STRUCTURE = 5, 6, 7, 8
CG_NONE, CG_CONVERGED, CG_ITERATION_LIMIT, CG_NOT_POSITIVE_DEFINITE, CG_NON_FINITE = -1, 0, 1, 2, 3
DUPLICATE_COVARIANCE_BLOCK, SINGULAR_INFORMATION, INVALID_INFORMATION, INVALID_RANK_TOLERANCE = 10, 11, 12, 13
UNKNOWN_KEY = 14
COVARIANCE_TOLERANCE = 64 * np.finfo(np.float64).eps  # CovarianceOptions default (src/covariance.rs:145-151)

_size_t = ctypes.c_size_t
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_szp = ctypes.POINTER(ctypes.c_size_t)


class SchurDiagnostics(ctypes.Structure):
    _fields_ = [("termination", ctypes.c_int), ("iterations", _size_t), ("damping", ctypes.c_double),
                ("estimated_relative_residual", ctypes.c_double),
                ("reduced_relative_residual", ctypes.c_double),
                ("scaled_relative_residual", ctypes.c_double),
                ("original_relative_residual", ctypes.c_double)]


class LsmrDiagnostics(ctypes.Structure):
    """LsmrDiagnostics (src/lsmr.rs:72-105); NaN = None."""
    _fields_ = [("termination", ctypes.c_int), ("iterations", _size_t), ("finite_step", ctypes.c_int),
                ("estimated_absolute_residual", ctypes.c_double), ("estimated_relative_residual", ctypes.c_double),
                ("original_absolute_residual", ctypes.c_double), ("original_relative_residual", ctypes.c_double),
                ("scaled_absolute_residual", ctypes.c_double), ("scaled_relative_residual", ctypes.c_double),
                ("min_scale", ctypes.c_double), ("max_scale", ctypes.c_double), ("damping", ctypes.c_double),
                ("block_fallbacks", _size_t)]


LSMR_IDENTITY, LSMR_DIAGONAL, LSMR_BLOCK = 0, 1, 2
LSMR_CONVERGED, LSMR_ITERATION_LIMIT, LSMR_NON_FINITE = 0, 1, 2


class OptimizeOptions(ctypes.Structure):
    _fields_ = [("max_iterations", _size_t), ("gradient_tolerance", ctypes.c_double),
                ("step_tolerance", ctypes.c_double), ("cost_tolerance", ctypes.c_double)]

    @classmethod
    def default(cls):
        # src/optimization.rs:57-67 for f64
        return cls(50, 1e-8, 1e-10, 1e-12)


class LmOptions(ctypes.Structure):
    _fields_ = [("initial_damping", ctypes.c_double), ("min_damping", ctypes.c_double),
                ("max_damping", ctypes.c_double), ("min_column_norm", ctypes.c_double),
                ("acceptance_threshold", ctypes.c_double), ("max_trials", _size_t)]

    @classmethod
    def default(cls):
        # src/optimization/levenberg_marquardt.rs:35-46
        return cls(1e-3, 1e-12, 1e12, 1e-6, 1e-4, 16)


class LmOutcome(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("termination", ctypes.c_int), ("iterations", _size_t),
                ("initial_cost", ctypes.c_double), ("final_cost", ctypes.c_double),
                ("linearizations", _size_t), ("attempts", _size_t), ("rejected_steps", _size_t),
                ("accepted_steps", _size_t), ("cost_evaluations", _size_t),
                ("damping", ctypes.c_double), ("cost", ctypes.c_double),
                ("gradient_norm", ctypes.c_double), ("step_norm", ctypes.c_double),
                ("last_rejection", ctypes.c_int), ("linear_solves", _size_t),
                ("linear_iterations", _size_t)]


_VP = ctypes.c_void_p
_FN_PREPARE = ctypes.CFUNCTYPE(ctypes.c_int, _VP)
_FN_COST = ctypes.CFUNCTYPE(ctypes.c_int, _VP, ctypes.c_int, _dp)
_FN_LIN = ctypes.CFUNCTYPE(ctypes.c_int, _VP)
_FN_GN = ctypes.CFUNCTYPE(ctypes.c_int, _VP, _dp)
_FN_PD = ctypes.CFUNCTYPE(ctypes.c_int, _VP, ctypes.c_double)
_FN_SD = ctypes.CFUNCTYPE(ctypes.c_int, _VP, ctypes.c_double, ctypes.POINTER(_dp), _dp)
_FN_S = ctypes.CFUNCTYPE(ctypes.c_int, _VP, ctypes.POINTER(_dp))
_FN_STAGE = ctypes.CFUNCTYPE(ctypes.c_int, _VP, _dp)
_FN_VOID = ctypes.CFUNCTYPE(None, _VP)
_FN_WORK = ctypes.CFUNCTYPE(None, _VP, _szp, _szp)
_FN_ACC = ctypes.CFUNCTYPE(None, _VP, ctypes.POINTER(LmOutcome))


class Problem(ctypes.Structure):
    _fields_ = [("ctx", _VP), ("dimension", _size_t), ("prepare", _FN_PREPARE), ("cost", _FN_COST),
                ("linearize", _FN_LIN), ("gradient_norm", _FN_GN), ("prepare_damping", _FN_PD),
                ("solve_damped", _FN_SD), ("solve", _FN_S), ("stage", _FN_STAGE),
                ("accept", _FN_VOID), ("reject", _FN_VOID), ("linear_work", _FN_WORK),
                ("on_accept", _FN_ACC)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_projection": (ctypes.c_int, [_dp, _dp, _dp]),
        "orc_project": (ctypes.c_int, [_dp, _dp, _dp]),
        "orc_jacobians": (None, [_dp, _dp, _dp, _dp]),
        "orc_schur_new": (_VP, [_size_t, _size_t, _size_t]),
        "orc_schur_free": (None, [_VP]),
        "orc_schur_set_options": (None, [_VP, _size_t, ctypes.c_double]),
        "orc_schur_prepare": (ctypes.c_int, [_VP, _size_t]),
        "orc_schur_clear": (None, [_VP]),
        "orc_schur_accumulate": (ctypes.c_int, [_VP, _dp, _size_t, _size_t, _szp, _szp, _dp]),
        "orc_schur_gradient_norm": (ctypes.c_int, [_VP, _dp]),
        "orc_schur_prepare_damping": (ctypes.c_int, [_VP, ctypes.c_double]),
        "orc_schur_solve_damped": (ctypes.c_int, [_VP, ctypes.c_double, _dp, _dp]),
        "orc_schur_solve": (ctypes.c_int, [_VP, _dp]),
        "orc_schur_diagnostics_get": (None, [_VP, ctypes.POINTER(SchurDiagnostics)]),
        "orc_schur_statistics": (None, [_VP, _szp, _szp]),
        "orc_schur_dimension": (_size_t, [_VP]),
        "orc_schur_rows": (_size_t, [_VP]),
        "orc_schur_gradient": (None, [_VP, _dp]),
        "orc_schur_scales": (None, [_VP, _dp]),
        "orc_schur_negative_residual": (None, [_VP, _dp]),
        "orc_schur_reduced_apply": (ctypes.c_int, [_VP, ctypes.c_double, _dp, _dp]),
        "orc_schur_reduced_rhs": (None, [_VP, _dp]),
        "orc_schur_gram_block": (ctypes.c_int, [_VP, _size_t, _dp, _szp]),
        "orc_schur_jacobian_apply": (None, [_VP, _dp, _dp]),
        "orc_schur_jacobian_transpose_apply": (None, [_VP, _dp, _dp]),
        "orc_schur_damped_apply": (None, [_VP, ctypes.c_double, ctypes.c_int, _dp, _dp]),
        "orc_schur_damped_transpose_apply": (None, [_VP, ctypes.c_double, ctypes.c_int, _dp, _dp]),
        "orc_schur_factor": (_size_t, [_VP, ctypes.c_double]),
        "orc_schur_apply_factor_inverse": (None, [_VP, _dp, ctypes.c_int]),
        "orc_schur_prepare_blocks": (ctypes.c_int, [_VP, _dp]),
        "orc_schur_gram_set": (None, [_VP, _size_t, ctypes.c_double]),
        "orc_schur_prepare_scales": (ctypes.c_int, [_VP, ctypes.c_double]),
        "orc_schur_verify": (ctypes.c_double, [_VP, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]),
        "orc_lsmr_solve_damped": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, _size_t, ctypes.c_double,
                                                 _dp, _dp, ctypes.POINTER(LsmrDiagnostics)]),
        "orc_lsmr_solve": (ctypes.c_int, [_VP, _size_t, ctypes.c_double, _dp, ctypes.POINTER(LsmrDiagnostics)]),
        "orc_dense_covariance": (ctypes.c_int, [_dp, _size_t, _szp, _size_t, ctypes.c_double, _dp]),
        "orc_schur_covariance": (ctypes.c_int, [_VP, _szp, _szp, _size_t, ctypes.c_double, _dp]),
        "orc_ba_new": (_VP, [_size_t, _size_t, _size_t, _dp, _dp, _i64p, _i64p, _dp, ctypes.c_double]),
        "orc_ba_free": (None, [_VP]),
        "orc_ba_layout": (None, [_VP, _i64p, _i64p, _i64p, _i64p]),
        "orc_ba_cost": (ctypes.c_int, [_VP, _dp]),
        "orc_ba_get_state": (None, [_VP, _dp, _dp]),
        "orc_ba_set_state": (None, [_VP, _dp, _dp]),
        "orc_ba_linearize": (ctypes.c_int, [_VP, _VP]),
        "orc_ba_linearize_sequential": (ctypes.c_int, [_VP, _VP]),
        "orc_set_threads": (None, [ctypes.c_int]),
        "orc_get_threads": (ctypes.c_int, []),
        "orc_ba_retract": (None, [_VP, _dp]),
        "orc_lm_optimize": (None, [ctypes.POINTER(Problem), ctypes.POINTER(OptimizeOptions),
                                   ctypes.POINTER(LmOptions), ctypes.POINTER(LmOutcome)]),
        "orc_gn_optimize": (None, [ctypes.POINTER(Problem), ctypes.POINTER(OptimizeOptions),
                                   ctypes.POINTER(LmOutcome)]),
        "orc_ba_lm": (None, [_VP, _size_t, ctypes.c_double, ctypes.POINTER(OptimizeOptions),
                             ctypes.POINTER(LmOptions), ctypes.POINTER(LmOutcome), _dp, _size_t, _szp]),
        "orc_ba_gn": (None, [_VP, _size_t, ctypes.c_double, ctypes.POINTER(OptimizeOptions),
                             ctypes.POINTER(LmOutcome)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def set_threads(n):
    """OpenMP threads of the oracle's block loops (1 = the single-threaded restatement)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _out(n):
    a = np.zeros(n, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def projection(cam):
    cam, pc = _d(cam)
    r, pr = _out(9)
    j, pj = _out(9)
    lib().orc_projection(pc, pr, pj)
    return r.reshape(3, 3), j.reshape(3, 3)


def project(cam, point):
    cam, pc = _d(cam)
    pt, pp = _d(point)
    px, ppx = _out(2)
    st = lib().orc_project(pc, pp, ppx)
    return st, px


def jacobians(cam, point):
    """(Jc 2x9, Jp 2x3) as dense arrays."""
    cam, pc = _d(cam)
    pt, pp = _d(point)
    jc, pjc = _out(18)
    jp, pjp = _out(6)
    lib().orc_jacobians(pc, pp, pjc, pjp)
    return jc.reshape(9, 2).T.copy(), jp.reshape(3, 2).T.copy()


def dense_covariance(information, columns, tolerance=COVARIANCE_TOLERANCE):
    """SelectedCovariance on a dense information matrix (src/covariance.rs:163-262)."""
    h = np.asfortranarray(information, dtype=np.float64)
    cols = (ctypes.c_size_t * max(1, len(columns)))(*[int(c) for c in columns])
    out = np.zeros((len(columns), len(columns)), order="F")
    st = lib().orc_dense_covariance(h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), h.shape[0], cols,
                                    len(columns), tolerance, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return st, (out if st == OK else None)


class Schur:
    """Generic Schur backend (src/schur.rs) driven through its trait methods."""

    def __init__(self, retained, retained_width, eliminated_width):
        self.h = lib().orc_schur_new(retained, retained_width, eliminated_width)
        self.max_iterations = 1000
        self.relative_tolerance = 1e-10

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_schur_free(self.h)
            self.h = None

    def _opts(self):
        lib().orc_schur_set_options(self.h, self.max_iterations, self.relative_tolerance)

    def prepare(self, dimension):
        self._opts()
        return lib().orc_schur_prepare(self.h, dimension)

    def clear(self):
        lib().orc_schur_clear(self.h)

    def accumulate(self, residual, blocks):
        """blocks: list of (col, dense rows x width matrix)."""
        residual = np.atleast_1d(np.asarray(residual, dtype=np.float64))
        cols = (ctypes.c_size_t * len(blocks))(*[int(c) for c, _ in blocks])
        widths = (ctypes.c_size_t * len(blocks))(*[np.asarray(m).reshape(len(residual), -1).shape[1]
                                                     for _, m in blocks])
        vals = np.concatenate([np.asarray(m, dtype=np.float64).reshape(len(residual), -1).ravel(order="F")
                               for _, m in blocks]) if blocks else np.zeros(0)
        r, pr = _d(residual)
        v, pv = _d(vals)
        return lib().orc_schur_accumulate(self.h, pr, len(residual), len(blocks), cols, widths, pv)

    def dimension(self):
        return lib().orc_schur_dimension(self.h)

    def rows(self):
        return lib().orc_schur_rows(self.h)

    def gradient_norm(self):
        out = ctypes.c_double()
        st = lib().orc_schur_gradient_norm(self.h, ctypes.byref(out))
        return st, out.value

    def prepare_damping(self, floor):
        self._opts()
        return lib().orc_schur_prepare_damping(self.h, floor)

    def solve_damped(self, lam):
        self._opts()
        d, pd = _out(self.dimension())
        pred = ctypes.c_double()
        st = lib().orc_schur_solve_damped(self.h, lam, pd, ctypes.byref(pred))
        return st, d, pred.value

    def solve(self):
        self._opts()
        d, pd = _out(self.dimension())
        st = lib().orc_schur_solve(self.h, pd)
        return st, d

    # LSMR backend over this model (src/lsmr.rs); blocks come from prepare_damping
    def lsmr_solve_damped(self, lam, mode=LSMR_DIAGONAL, max_iterations=1000, tol=1e-10):
        d, pd = _out(self.dimension())
        pred = ctypes.c_double(float("nan"))
        diag = LsmrDiagnostics()
        st = lib().orc_lsmr_solve_damped(self.h, lam, mode, max_iterations, tol, pd, ctypes.byref(pred),
                                         ctypes.byref(diag))
        return st, d, pred.value, diag

    def covariance(self, blocks, tolerance=COVARIANCE_TOLERANCE):
        """LinearizedModel::covariance (src/linear_model.rs:122-167) for (offset, width) blocks."""
        offs = (ctypes.c_size_t * max(1, len(blocks)))(*[int(o) for o, _ in blocks])
        wids = (ctypes.c_size_t * max(1, len(blocks)))(*[int(w) for _, w in blocks])
        k = sum(int(w) for _, w in blocks)
        out = np.zeros((k, k), order="F")
        st = lib().orc_schur_covariance(self.h, offs, wids, len(blocks), tolerance,
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return st, (out if st == OK else None)

    def lsmr_solve(self, max_iterations=1000, tol=1e-10):
        d, pd = _out(self.dimension())
        diag = LsmrDiagnostics()
        st = lib().orc_lsmr_solve(self.h, max_iterations, tol, pd, ctypes.byref(diag))
        return st, d, diag

    def diagnostics(self):
        d = SchurDiagnostics()
        lib().orc_schur_diagnostics_get(self.h, ctypes.byref(d))
        return d

    def statistics(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        lib().orc_schur_statistics(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def gradient(self):
        g, p = _out(self.dimension())
        lib().orc_schur_gradient(self.h, p)
        return g

    def scales(self):
        g, p = _out(self.dimension())
        lib().orc_schur_scales(self.h, p)
        return g

    def negative_residual(self):
        g, p = _out(self.rows())
        lib().orc_schur_negative_residual(self.h, p)
        return g

    def reduced_apply(self, lam, x):
        x, px = _d(x)
        y, py = _out(len(x))
        st = lib().orc_schur_reduced_apply(self.h, lam, px, py)
        return st, y

    def reduced_rhs(self, retained):
        y, py = _out(retained)
        lib().orc_schur_reduced_rhs(self.h, py)
        return y

    def gram_block(self, col, width):
        y, py = _out(width * width)
        w = ctypes.c_size_t()
        st = lib().orc_schur_gram_block(self.h, col, py, ctypes.byref(w))
        return st, y.reshape(width, width).T.copy()   # column-major -> (row, col)

    def jacobian_apply(self, x):
        x, px = _d(x)
        y, py = _out(self.rows())
        lib().orc_schur_jacobian_apply(self.h, px, py)
        return y

    def jacobian_transpose_apply(self, y):
        y, py = _d(y)
        x, px = _out(self.dimension())
        lib().orc_schur_jacobian_transpose_apply(self.h, py, px)
        return x

    def damped_apply(self, lam, scaled, x):
        x, px = _d(x)
        y, py = _out(self.rows() + self.dimension())
        lib().orc_schur_damped_apply(self.h, lam, int(scaled), px, py)
        return y

    def damped_transpose_apply(self, lam, scaled, y):
        y, py = _d(y)
        x, px = _out(self.dimension())
        lib().orc_schur_damped_transpose_apply(self.h, lam, int(scaled), py, px)
        return x

    def factor(self, lam):
        return lib().orc_schur_factor(self.h, lam)

    def apply_factor_inverse(self, rhs, transpose):
        r, pr = _d(np.array(rhs, dtype=np.float64))
        lib().orc_schur_apply_factor_inverse(self.h, pr, int(transpose))
        return r

    def prepare_blocks(self, scales):
        s, ps = _d(scales)
        return lib().orc_schur_prepare_blocks(self.h, ps)

    def gram_set(self, index, value):
        lib().orc_schur_gram_set(self.h, index, value)

    def prepare_scales(self, floor):
        return lib().orc_schur_prepare_scales(self.h, floor)

    def verify(self, step, lam):
        s, ps = _d(step)
        n = self.dimension()
        nr, pnr = _out(n)
        ref, pref = _out(n)
        a, b = ctypes.c_double(), ctypes.c_double()
        pred = lib().orc_schur_verify(self.h, ps, lam, pnr, pref, ctypes.byref(a), ctypes.byref(b))
        return pred, nr, ref, a.value, b.value


class BA:
    """BA problem with the BAL model and one Frame batch per camera."""

    def __init__(self, problem, huber=0.0):
        self.nc, self.np, self.no = problem.n_cameras, problem.n_points, problem.n_observations
        self._keep = [_d(problem.cameras), _d(problem.points),
                      np.ascontiguousarray(problem.obs_camera, dtype=np.int64),
                      np.ascontiguousarray(problem.obs_point, dtype=np.int64), _d(problem.obs_pixel)]
        cams, pts, oc, op, px = self._keep
        self.h = lib().orc_ba_new(self.nc, self.np, self.no, cams[1], pts[1],
                                  oc.ctypes.data_as(_i64p), op.ctypes.data_as(_i64p), px[1], float(huber))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ba_free(self.h)
            self.h = None

    def layout(self):
        a = np.zeros(self.no, np.int64)
        b = np.zeros(self.nc + 1, np.int64)
        c = np.zeros(self.no, np.int64)
        d = np.zeros(self.np + 1, np.int64)
        lib().orc_ba_layout(self.h, *[x.ctypes.data_as(_i64p) for x in (a, b, c, d)])
        return a, b, c, d

    def cost(self):
        out = ctypes.c_double()
        st = lib().orc_ba_cost(self.h, ctypes.byref(out))
        return st, out.value

    def state(self):
        c, pc = _out(9 * self.nc)
        p, pp = _out(3 * self.np)
        lib().orc_ba_get_state(self.h, pc, pp)
        return c.reshape(self.nc, 9), p.reshape(self.np, 3)

    def set_state(self, cams, pts):
        c, pc = _d(cams)
        p, pp = _d(pts)
        lib().orc_ba_set_state(self.h, pc, pp)

    def linearize(self, schur):
        return lib().orc_ba_linearize(self.h, schur.h)

    def linearize_sequential(self, schur):
        """The same linearization one emission at a time through Schur.accumulate."""
        return lib().orc_ba_linearize_sequential(self.h, schur.h)

    def retract(self, delta):
        d, pd = _d(delta)
        lib().orc_ba_retract(self.h, pd)

    def lm(self, options=None, lm_options=None, max_cg=1000, tol=1e-10, trace_capacity=0):
        options = options or OptimizeOptions.default()
        lm_options = lm_options or LmOptions.default()
        out = LmOutcome()
        tr, ptr = _out(3 * max(trace_capacity, 1))
        n = ctypes.c_size_t()
        lib().orc_ba_lm(self.h, max_cg, tol, ctypes.byref(options), ctypes.byref(lm_options),
                        ctypes.byref(out), ptr, trace_capacity, ctypes.byref(n))
        # per accepted step: cost, attempts so far, damping of the accepted trial
        return out, tr[:3 * n.value].reshape(n.value, 3).copy()

    def gn(self, options=None, max_cg=1000, tol=1e-10):
        options = options or OptimizeOptions.default()
        out = LmOutcome()
        lib().orc_ba_gn(self.h, max_cg, tol, ctypes.byref(options), ctypes.byref(out))
        return out


def ba_schur(ba, max_cg=1000, tol=1e-10):
    """A Schur backend prepared for ``ba``'s coordinates and linearized at its state."""
    s = Schur(9 * ba.nc, 9, 3)
    s.max_iterations = max_cg
    s.relative_tolerance = tol
    assert s.prepare(9 * ba.nc + 3 * ba.np) == OK
    st = ba.linearize(s)
    return st, s
