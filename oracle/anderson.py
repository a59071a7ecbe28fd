"""Anderson acceleration, restated for checking (test infrastructure: only
the tests import it; the product is bs_anderson_* in the HIP library).

AndersonMixing follows src/solver/anderson.ts:20-123 line by line: the
differences of the last `depth` iterates and corrections, gamma by
modified Gram-Schmidt with the oldest columns dropped while a column's
orthogonal part is not above 1e-10 of its norm. `accelerated_steps`
follows the inner loop of solveTransonic (src/analysis/transonic-flow.ts:
184-208) on the FP64 oracle solver.
"""
from __future__ import annotations

import math

import numpy as np

from pyoracle import NonPhysicalStateError


class AndersonMixing:
    """anderson.ts:20-123."""

    def __init__(self, size: int, depth: int, dependence: float = 1e-10):
        if not (isinstance(depth, int) and depth >= 1):
            raise ValueError("depth must be at least 1")
        self.size, self.depth, self.dependence = size, depth, dependence
        self.dx: list = []
        self.df: list = []
        self.last_x = None
        self.last_f = None
        self.gamma = np.zeros(0)

    def reset(self):
        """anderson.ts:34-40."""
        self.dx.clear()
        self.df.clear()
        self.last_x = None
        self.last_f = None

    def next(self, x, f) -> np.ndarray:
        """anderson.ts:43-70."""
        x = np.asarray(x, dtype=np.float64)
        f = np.asarray(f, dtype=np.float64)
        if x.size != self.size or f.size != self.size:
            raise ValueError("wrong vector size")
        if self.last_x is not None and self.last_f is not None:
            self.dx.append(x - self.last_x)
            self.df.append(f - self.last_f)
            if len(self.dx) > self.depth:
                self.dx.pop(0)
                self.df.pop(0)
        self.last_x = x.copy()
        self.last_f = f.copy()
        out = x + f
        gamma = self._least_squares(f)
        self.gamma = gamma
        for c, g in enumerate(gamma):
            if g != 0:
                out = out - g * (self.dx[c] + self.df[c])
        return out

    def _least_squares(self, f) -> np.ndarray:
        """anderson.ts:73-116: modified Gram-Schmidt, dropping the oldest
        columns until the newest are well conditioned."""
        m = len(self.df)
        gamma = np.zeros(m)
        for first in range(m):
            cols = m - first
            q = []
            r = np.zeros((cols, cols))
            ok = True
            for c in range(cols):
                v = self.df[first + c].copy()
                norm0 = math.sqrt(float(v @ v))
                for p in range(c):
                    proj = float(q[p] @ v)
                    r[p, c] = proj
                    v = v - proj * q[p]
                norm = math.sqrt(float(v @ v))
                if not norm > self.dependence * norm0 or norm0 == 0:
                    ok = False
                    break
                r[c, c] = norm
                q.append(v / norm)
            if not ok:
                continue
            rhs = [float(qc @ f) for qc in q]
            for c in range(cols - 1, -1, -1):
                s = rhs[c]
                for p in range(c + 1, cols):
                    s -= r[c, p] * gamma[first + p]
                gamma[first + c] = s / r[c, c]
            return gamma
        return gamma


def accelerated_steps(orc, anderson: AndersonMixing, count: int, cfl: float, cfl_max: float, omega: float,
                      relaxation: str, sweeps: int, best: float = math.inf, restarts: int = 0):
    """transonic-flow.ts:184-208 on the oracle solver: `count` steps, the
    CFL number ramped by 1.2 up to cfl_max after each. Returns (cfl, best,
    restarts)."""
    for _ in range(count):
        x = orc.conserved().reshape(-1)
        f = orc.implicit_correction(cfl, omega, relaxation, sweeps).reshape(-1)
        norm = float(f @ f)
        if norm > 100 * best:
            anderson.reset()
            restarts += 1
            best = norm
        best = min(best, norm)
        nxt = anderson.next(x, f)
        try:
            orc.set_conserved(nxt)
        except NonPhysicalStateError:
            orc.set_conserved(x)
            orc.iterate_implicit(1, cfl, omega, relaxation, sweeps)
            anderson.reset()
            restarts += 1
        cfl = min(cfl_max, cfl * 1.2)
    return cfl, best, restarts
