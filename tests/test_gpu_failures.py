"""Failure detection on the device (euler2d.wgsl is_physical and the status
word's atomicOr, 265-269 and 488-490; StepLedger.account,
time-loop.ts:125-150; the reference's tests/engine-gpu/time-loop.test.ts).

A state with one non-physical cell is uploaded with bs_set_state, which
does not check it (bs_set_primitives refuses such a state on the host), so
only the kernels can notice it: the stencil's physical check (explicit
steps), the time-step reduction (global steps) or the implicit update's.
Every kernel exits normally (NaN only flows through arithmetic; the
update's halving loop is bounded); the batch of steps ends with the status
bit set and the call raises NonPhysicalStateError naming the steps, and so
does a later readState. A healthy state next to it is unaffected."""
import numpy as np
import pytest

import cases
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D, NonPhysicalStateError

pytestmark = pytest.mark.gpu


def _poisoned(ctx, precision, stepping="local"):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(32, 16, mach=2.0)
    cfg = Euler2DConfig(grid=grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4, flux="hllc",
                        limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5,
                        timeStepping=stepping)
    gpu = GpuEuler2D.create(ctx, cfg, precision)
    gpu.setCellPrimitives(np.array(free))
    st = gpu.readState()
    q = [np.array(a, dtype=np.float64) for a in (st.rho, st.mu, st.mv, st.energy)]
    c = 7 + 5 * 32
    q[0][c] = -q[0][c]  # negative density in one cell
    gpu.setState(*q)
    return gpu, free


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("stepping", ["local", "global"])
def test_explicit_steps_report_a_non_physical_cell_from_the_device(ctx, precision, stepping):
    gpu, _ = _poisoned(ctx, precision, stepping)
    with pytest.raises(NonPhysicalStateError, match="between steps 0 and"):
        gpu.iterate(24)
    with pytest.raises(NonPhysicalStateError):
        gpu.readState()
    gpu.destroy()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_implicit_steps_report_a_non_physical_cell_from_the_device(ctx, precision):
    gpu, _ = _poisoned(ctx, precision)
    gpu.setImplicitCfl(1.0)
    with pytest.raises(NonPhysicalStateError, match="between steps 0 and"):
        gpu.iterateImplicit(3, cfl=10.0, sweeps=2)
    gpu.destroy()


def test_a_new_start_clears_the_flag_and_a_healthy_state_runs(ctx):
    gpu, free = _poisoned(ctx, "fp32")
    with pytest.raises(NonPhysicalStateError):
        gpu.iterate(3)
    gpu.setCellPrimitives(np.array(free))
    gpu.iterate(24)
    st = gpu.readState()
    assert np.isfinite(st.rho).all() and (st.rho > 0).all()
    assert st.steps == 24
    gpu.destroy()
