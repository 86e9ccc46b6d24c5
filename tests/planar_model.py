"""A host model that is not BAL, for the generic-emission backend's tests: planar poses
(x, y, theta) observing planar landmarks in their own frame, plus a prior on the first pose,
and the oracle's Schur backend behind the same methods. Test infrastructure."""
import numpy as np

import fagra_amd as fa
from oracle import oracle as O


class Planar:
    """Planar poses (x, y, theta; retained) observing landmarks (x, y; eliminated) in their own
    frame, with a prior fixing the first pose; two emission families (observations: pose +
    landmark blocks, 2 rows; the prior: one pose block, 3 rows)."""

    def __init__(self, n_poses=12, n_landmarks=150, n_obs=600, seed=0, noise=0.01):
        rng = np.random.default_rng(seed)
        self.np_, self.nl = n_poses, n_landmarks
        self.poses_true = np.column_stack([rng.uniform(-5, 5, (n_poses, 2)), rng.uniform(-np.pi, np.pi, n_poses)])
        self.lm_true = rng.uniform(-8, 8, (n_landmarks, 2))
        pi = np.concatenate([np.arange(n_poses), rng.integers(0, n_poses, n_obs - n_poses)])
        li = np.concatenate([rng.permutation(n_landmarks)[:n_poses], rng.integers(0, n_landmarks, n_obs - n_poses)])
        # every landmark seen at least twice
        extra_l = np.repeat(np.arange(n_landmarks), 2)
        extra_p = rng.integers(0, n_poses, extra_l.size)
        self.pi = np.concatenate([pi, extra_p])
        self.li = np.concatenate([li, extra_l])
        self.z = self.measure(self.poses_true, self.lm_true) + rng.normal(0, noise, (self.pi.size, 2))
        self.prior = self.poses_true[0].copy()
        self.x = np.concatenate([(self.poses_true + rng.normal(0, [0.1, 0.1, 0.05], self.poses_true.shape)).ravel(),
                                 (self.lm_true + rng.normal(0, 0.2, self.lm_true.shape)).ravel()])
        self.trial = None
        self.dimension = 3 * n_poses + 2 * n_landmarks

    def split(self, x):
        return x[:3 * self.np_].reshape(-1, 3), x[3 * self.np_:].reshape(-1, 2)

    def measure(self, poses, lms):
        t, th = poses[self.pi, :2], poses[self.pi, 2]
        d = lms[self.li] - t
        c, s = np.cos(th), np.sin(th)
        return np.column_stack([c * d[:, 0] + s * d[:, 1], -s * d[:, 0] + c * d[:, 1]])

    def residuals(self, x):
        poses, lms = self.split(x)
        return self.measure(poses, lms) - self.z, poses[0] - self.prior

    def cost(self, trial):
        r, p = self.residuals(self.trial if trial else self.x)
        return fa.OK, 0.5 * (np.sum(r * r) + np.sum(p * p))

    def jacobians(self, x):
        poses, lms = self.split(x)
        t, th = poses[self.pi, :2], poses[self.pi, 2]
        d = lms[self.li] - t
        c, s = np.cos(th), np.sin(th)
        n = self.pi.size
        jl = np.zeros((n, 2, 2))
        jl[:, 0, 0], jl[:, 0, 1], jl[:, 1, 0], jl[:, 1, 1] = c, s, -s, c
        jp = np.zeros((n, 2, 3))
        jp[:, :, :2] = -jl
        jp[:, 0, 2] = -s * d[:, 0] + c * d[:, 1]
        jp[:, 1, 2] = -c * d[:, 0] - s * d[:, 1]
        return jp, jl

    def linearize(self, backend):
        r, p = self.residuals(self.x)
        jp, jl = self.jacobians(self.x)
        cols = np.column_stack([3 * self.pi, 3 * self.np_ + 2 * self.li])
        st = backend.accumulate_batch(r, cols, [jp, jl])
        if st != fa.OK:
            return st
        return backend.accumulate(p, [(0, np.eye(3))])

    def stage(self, delta):
        self.trial = self.x + delta
        return fa.OK

    def accept(self):
        self.x = self.trial

    def reject(self):
        self.trial = None


    def dense_jacobian(self):
        """J (rows in emission order: observations, then the prior) and r at the accepted state."""
        r, p = self.residuals(self.x)
        jp, jl = self.jacobians(self.x)
        n_obs = self.pi.size
        J = np.zeros((2 * n_obs + 3, self.dimension))
        for e in range(n_obs):
            J[2 * e:2 * e + 2, 3 * self.pi[e]:3 * self.pi[e] + 3] = jp[e]
            c = 3 * self.np_ + 2 * self.li[e]
            J[2 * e:2 * e + 2, c:c + 2] = jl[e]
        J[2 * n_obs:, 0:3] = np.eye(3)
        return J, np.concatenate([r.ravel(), p])


class OracleBackend:
    """The oracle's Schur backend behind the same methods (one accumulate per emission)."""

    def __init__(self, retained, rw, ew):
        self.s = O.Schur(retained, rw, ew)
        self.d = None

    def prepare(self, n):
        return self.s.prepare(n)

    def clear(self):
        self.s.clear()

    def accumulate(self, residual, blocks):
        return self.s.accumulate(residual, blocks)

    def accumulate_batch(self, residuals, cols, values):
        for e in range(residuals.shape[0]):
            st = self.s.accumulate(residuals[e], [(int(cols[e, k]), values[k][e]) for k in range(len(values))])
            if st != O.OK:
                return st
        return O.OK

    def gradient_norm(self):
        return self.s.gradient_norm()

    def prepare_damping(self, f):
        return self.s.prepare_damping(f)

    def solve_damped(self, lam):
        st, d, pred = self.s.solve_damped(lam)
        self.d = d
        return st, d, pred

    def solve(self):
        st, d = self.s.solve()
        self.d = d
        return st, d

    def last_step_norm(self):
        return float(np.max(np.abs(self.d))) if self.d is not None and self.d.size else 0.0

    def statistics(self):
        return self.s.statistics()
