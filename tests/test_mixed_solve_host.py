"""The host logic of batch.solve_mixed on a scripted stand-in for the two
batches (no GPU): argument checks, the hand-over of every case's state and
CFL number, the restart from the free stream of a case that went
non-physical in FP32, and the steps and residual history of both phases."""
import numpy as np
import pytest

from bowshock_amd.batch import STATUS_NON_PHYSICAL, SolvePlan, solve_mixed
from bowshock_amd.engine import RangeError


class ScriptedBatch:
    """Residuals fall tenfold per chunk down to the precision's floor;
    cases in `bad` report a non-physical state at their first check."""

    def __init__(self, precision, n, bad=()):
        self.precision = precision
        self.count = n
        self.shapes = [(8, 4)] * n
        self.res = np.ones(n)
        self.steps = np.zeros(n, dtype=np.int64)
        self.word = np.zeros(n, dtype=np.uint32)
        self.active = np.ones(n, dtype=np.uint8)
        self.floor = 1e-5 if precision == "fp32" else 0.0
        self.bad = set(bad)
        self.cfl = None
        self.restarts = []

    def setActive(self, mask):
        self.active = np.array(mask, dtype=np.uint8)

    def setRelaxation(self, form):
        pass

    def setImplicitCfl(self, cfl):
        self.cfl = np.full(self.count, cfl)

    def setImplicitCfls(self, cfl):
        self.cfl = np.array(cfl, dtype=np.float64)

    def iterateImplicit(self, steps, cfl, sweeps, omega=1.5, growth=1.2):
        for k in np.nonzero(self.active)[0]:
            self.steps[k] += steps
            self.res[k] = max(self.floor, 0.1 * self.res[k])
            if k in self.bad:
                self.word[k] |= STATUS_NON_PHYSICAL

    def status(self):
        return self.word.copy(), self.steps.copy()

    def densityResidualNorms(self):
        return self.res.copy()

    def forebodyDrags(self, p_inf, q_inf, base_radius):
        return np.full(self.count, 0.25)

    def setFromCoarser(self, other):
        self.res = other.res.copy()
        self.steps[:] = 0
        self.word[:] = 0

    def setCellPrimitives(self, index, primitives):
        self.restarts.append((index, tuple(primitives)))
        self.res[index] = 1.0


PLAN32 = SolvePlan(implicit=True, chunk=60, max_steps=6000, cfl=10.0, cfl_start=1.0)
PLAN64 = SolvePlan(implicit=True, chunk=60, max_steps=6000, cfl=10.0, cfl_start=1.0, residual=1e-6, drop=0.0,
                   stall=10 ** 9)


def test_the_batches_must_match():
    frees = [(1.0, 2.0, 0.0, 0.7)] * 2
    with pytest.raises(RangeError):
        solve_mixed(ScriptedBatch("fp64", 2), ScriptedBatch("fp32", 2), frees, 1.0, 1.0, 0.5, PLAN32, PLAN64)
    b64 = ScriptedBatch("fp64", 2)
    b64.shapes = [(8, 4), (16, 8)]
    with pytest.raises(RangeError):
        solve_mixed(ScriptedBatch("fp32", 2), b64, frees, 1.0, 1.0, 0.5, PLAN32, PLAN64)


def test_hand_over_restart_and_step_count():
    frees = [(1.0, 2.0 + k, 0.0, 0.7) for k in range(3)]
    b32, b64 = ScriptedBatch("fp32", 3, bad={1}), ScriptedBatch("fp64", 3)
    res = solve_mixed(b32, b64, frees, 1.0, 1.0, 0.5, PLAN32, PLAN64)
    # Case 1 failed in FP32 and restarts from its own free stream with the
    # CFL ramp; the others continue from their FP32 states at the final CFL.
    assert b64.restarts == [(1, frees[1])]
    assert list(b64.cfl) == [10.0, 1.0, 10.0]
    for k, r in enumerate(res):
        assert r.converged and r.residual <= 1e-6, (k, r.reason)
        assert r.steps == 60 * len(r.history)
    # FP32: 1 -> 1e-5 takes 5 chunks (the drop of 1e-4 from the first
    # reading, 0.1), then FP64: 1e-5 -> 1e-6 and the drag settled.
    assert res[0].steps == res[2].steps
    assert res[0].history[:5] == pytest.approx([0.1, 1e-2, 1e-3, 1e-4, 1e-5])
    assert res[0].history[-1] <= 1e-6
    # The restarted case's FP32 phase ended at its first check; its FP64
    # phase starts again from a residual of 1.
    assert res[1].history[0] == pytest.approx(0.1)
    assert res[1].history[1] == pytest.approx(0.1)
