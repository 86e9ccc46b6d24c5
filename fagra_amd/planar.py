"""A planar pose / landmark problem: a host model that is not BAL, fed to the generic-emission
Schur backend (fagra_amd.generic) for its tests and its benchmark.

Poses (x, y, theta) are the retained variables, landmarks (x, y) the eliminated ones; each
observation is a landmark's position in the observing pose's frame (2 rows, one pose block and
one landmark block), and a prior on the first pose (3 rows) fixes the gauge. Linearization is
vectorized numpy and emits both families in bulk (accumulate_batch).
"""
import numpy as np

from ._lib import OK


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
        return OK, 0.5 * (np.sum(r * r) + np.sum(p * p))

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
        if st != OK:
            return st
        return backend.accumulate(p, [(0, np.eye(3))])

    def stage(self, delta):
        self.trial = self.x + delta
        return OK

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
