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
    def load(cls, path=WEIGHTS):
        return cls(params=np.fromfile(path, dtype=np.float32))

    def save(self, path=WEIGHTS):
        self.params().astype(np.float32).tofile(path)

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


def collect(seeds, workload="ring256", precision=None, cg_tol=1e-3, max_iter=15, log=None):
    """Run LM on seeded problems and capture every attempted Schur system (one sample per
    observed camera and attempt)."""
    import fagra_amd as fa
    from . import synthetic
    precision = fa.FP32 if precision is None else precision
    feats, targs = [], []
    for seed in seeds:
        p = synthetic.generate(workload, seed=seed)
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
