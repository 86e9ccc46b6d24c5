"""A host model that is not BAL, for the generic-emission backend's tests: planar poses
(x, y, theta) observing planar landmarks in their own frame, plus a prior on the first pose,
and the oracle's Schur backend behind the same methods. Test infrastructure."""
import numpy as np

from fagra_amd.planar import Planar  # noqa: F401 (re-exported for the tests)
from oracle import oracle as O


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
