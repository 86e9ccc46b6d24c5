"""The learned block-Jacobi correction: network handle, training data capture and training.

Per camera the network sees 47 features of its Gram block (log diagonal, correlations, log
observation count, log damping) and returns 45 values forming G = I + s tril(y); the PCG then
applies M = L^-T G G^T L^-1 instead of (B + lambda I)^-1 = L^-T L^-1. Training fits G G^T to the
inverse of A = L^-1 S_cc L^-T, S_cc the camera's diagonal block of the reduced system, on Schur
systems that the LM loop itself produces on seeded 256-camera problems. The output layer starts
at zero, so an untrained network is exactly block-Jacobi.
"""
import ctypes
import os

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "learned_precond.f32")
RANGE = os.path.join(HERE, "learned_precond.range.f32")  # 47 lower then 47 upper feature bounds
N_IN, N_OUT = 47, 45


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def available():
    return os.path.exists(WEIGHTS)


class Network:
    """Device parameters of the preconditioner network (fg_mlp)."""

    def __init__(self, seed=0, params=None):
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.fg_mlp_create(seed, ctypes.byref(h)), "fg_mlp_create")
        self.h = h.value
        self.n_params = lib.fg_mlp_param_count()
        if params is not None:
            self.set_params(params)

    @classmethod
    def load(cls, path=WEIGHTS, range_path=RANGE):
        net = cls(params=np.fromfile(path, dtype=np.float32))
        if range_path and os.path.exists(range_path):
            r = np.fromfile(range_path, dtype=np.float32)
            net.set_input_range(r[:N_IN], r[N_IN:])
        return net

    def save(self, path=WEIGHTS, range_path=RANGE):
        self.params().astype(np.float32).tofile(path)
        if range_path and getattr(self, "input_range", None) is not None:
            np.concatenate(self.input_range).astype(np.float32).tofile(range_path)

    def set_input_range(self, lo, hi):
        """Cameras whose features leave [lo, hi] keep block-Jacobi (None clears the range)."""
        if lo is None:
            _lib.check(_lib.load().fg_mlp_set_input_range(self.h, None, None), "fg_mlp_set_input_range")
            self.input_range = None
            return
        a, pa = _f(lo)
        b, pb = _f(hi)
        _lib.check(_lib.load().fg_mlp_set_input_range(self.h, pa, pb), "fg_mlp_set_input_range")
        self.input_range = (a.copy(), b.copy())

    def close(self):
        if getattr(self, "h", None):
            _lib.load().fg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out, p = _f(np.zeros(self.n_params))
        _lib.check(_lib.load().fg_mlp_get_params(self.h, p), "fg_mlp_get_params")
        return out

    def set_params(self, params):
        params = np.asarray(params, dtype=np.float32)
        assert params.size == self.n_params
        a, p = _f(params)
        _lib.check(_lib.load().fg_mlp_set_params(self.h, p), "fg_mlp_set_params")

    def forward(self, features):
        x, px = _f(features)
        n = x.shape[0]
        y, py = _f(np.zeros((n, N_OUT)))
        _lib.check(_lib.load().fg_mlp_forward(self.h, n, px, py), "fg_mlp_forward")
        return y

    def loss_grad(self, features, targets):
        x, px = _f(features)
        t, pt = _f(targets)
        g, pg = _f(np.zeros(self.n_params))
        loss = ctypes.c_double()
        _lib.check(_lib.load().fg_mlp_loss_grad(self.h, x.shape[0], px, pt, ctypes.byref(loss), pg), "loss_grad")
        return loss.value, g

    def train(self, features, targets, epochs=20, batch=1024, lr=1e-3, seed=0):
        x, px = _f(features)
        t, pt = _f(targets)
        losses = np.zeros(epochs)
        _lib.check(_lib.load().fg_mlp_train(self.h, x.shape[0], px, pt, epochs, batch, lr, seed,
                                            losses.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "fg_mlp_train")
        return losses


def capture(schur, lam):
    """Features (nc x 47), targets (nc x 45) and observation counts of the prepared model."""
    nc = schur.problem.nc
    f, pf = _f(np.zeros((nc, N_IN)))
    t, pt = _f(np.zeros((nc, N_OUT)))
    c = np.zeros(nc, dtype=np.int32)
    _lib.check(_lib.load().fg_schur_capture(schur.h, float(lam), pf, pt,
                                            c.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "fg_schur_capture")
    return f, t, c


def feature_range(features, margin=0.02):
    """Per-feature bounds of a training set, widened by margin times their span; the damping
    feature (the last) keeps the exact span of the dampings trained on."""
    lo, hi = features.min(axis=0), features.max(axis=0)
    pad = margin * (hi - lo)
    lo, hi = lo - pad, hi + pad
    lo[-1], hi[-1] = features[:, -1].min(), features[:, -1].max()
    return lo.astype(np.float32), hi.astype(np.float32)


def collect(seeds, workload="ring256", precision=None, cg_tol=1e-3, max_iter=15, log=None, **overrides):
    """Run LM on seeded problems and capture every attempted Schur system (one sample per
    observed camera and attempt); overrides go to the workload generator."""
    import fagra_amd as fa
    from . import synthetic
    precision = fa.FP32 if precision is None else precision
    feats, targs = [], []
    for seed in seeds:
        p = synthetic.generate(workload, seed=seed, **overrides)
        ba = fa.BalProblem(p, precision=precision)
        holder = {}

        def on_solve(d):
            if d.termination in (fa.CG_CONVERGED, fa.CG_ITERATION_LIMIT):
                f, t, c = capture(holder["s"], d.damping)
                keep = c > 0
                feats.append(f[keep])
                targs.append(t[keep])

        s = fa.Schur(precision=precision, relative_tolerance=cg_tol, on_solve=on_solve)
        holder["s"] = s
        out = fa.LevenbergMarquardt(s).solve_batch(ba, fa.OptimizeOptions(max_iter, 0.0, 0.0, 0.0))
        if log:
            log(f"collect: seed {seed}: {out.accepted_steps} steps, cost {out.cost:.6e}, "
                f"{sum(len(x) for x in feats)} samples")
        s.close()
        ba.close()
    return np.concatenate(feats), np.concatenate(targs)
