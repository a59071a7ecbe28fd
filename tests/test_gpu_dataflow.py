"""Dataflow launches (bs_set_dataflow) against one launch per stage: the
tile code is the same, so the states must agree bit for bit, whatever
order the (stage, tile) work items run in. Three integrators (the
three-buffer rotation of forward Euler, the in-place last stage of SSP-RK2
and SSP-RK3), runs crossing the 256-step status batches, a periodic grid
(whose tiles' neighbours wrap around), far-field and wall sides, and the
1280 x 640 cone the benchmark times. FP64 solvers refuse; a fitted shock
keeps the per-stage launches."""
import numpy as np
import pytest

import cases
from bowshock_amd import grids
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D, RangeError
from test_validation_pins import GAMMA, PERIODIC, vortex_state, wavy_grid

pytestmark = pytest.mark.gpu


def _state(gpu):
    s = gpu.readState()
    return np.stack([s.rho, s.mu, s.mv, s.energy])


def _pair(ctx, cfg, start):
    ref = GpuEuler2D.create(ctx, cfg)
    df = GpuEuler2D.create(ctx, cfg)
    df.setDataflow(True)
    for g in (ref, df):
        g.setCellPrimitives(start)
    return ref, df


@pytest.mark.parametrize("integrator", ["ssprk3", "ssprk2", "forwardEuler"])
def test_dataflow_equals_per_stage_launches(ctx, integrator):
    grid, free, _, _, bcs = cases.cone_case(130, 64, mach=2.5)
    cfg = Euler2DConfig(grid=grid, boundaries=bcs, geometry="axisymmetric", gamma=cases.GAMMA, flux="hllc",
                        limiter="minmod", reconstruction="perVariable", timeIntegrator=integrator, cfl=0.5,
                        timeStepping="local")
    ref, df = _pair(ctx, cfg, np.array(free))
    for count in (37, 300):
        ref.iterate(count)
        df.iterate(count)
        assert np.array_equal(_state(df), _state(ref)), count
        assert df.last_report.steps == ref.last_report.steps
    ref.destroy()
    df.destroy()


def test_dataflow_on_a_periodic_grid(ctx):
    """The isentropic vortex on a distorted grid periodic in both
    directions, local steps: the corner tiles wait on the tiles across the
    grid."""
    from bowshock_amd.engine import gpu_primitives  # noqa: F401
    import pyoracle as O

    g = wavy_grid(40)
    cfg = Euler2DConfig(grid=g, geometry="planar", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", timeIntegrator="ssprk3", cfl=0.4, timeStepping="local",
                        boundaries=PERIODIC)
    m = O.compute_metrics(g, False)
    ref, df = _pair(ctx, cfg, vortex_state(m["cx"], m["cy"], 0.0))
    ref.iterate(120)
    df.iterate(120)
    assert np.array_equal(_state(df), _state(ref))
    ref.destroy()
    df.destroy()


def test_dataflow_on_the_benchmark_grid(ctx):
    body = grids.cone_cylinder(15.0, 4.0)
    bg = grids.body_grid_fixed(body, 1280, 640, 3.0, 1.4)
    free, _, _ = grids.free_stream_at(3.0, 1.4)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
           "jMax": {"kind": "inflow", "state": free}}
    cfg = Euler2DConfig(grid=bg.grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4, flux="hllc",
                        limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5,
                        timeStepping="local")
    ref, df = _pair(ctx, cfg, np.array(free))
    ref.iterate(60)
    df.iterate(60)
    assert np.array_equal(_state(df), _state(ref))
    ref.destroy()
    df.destroy()


def test_dataflow_refusals(ctx):
    grid, free, _, _, bcs = cases.cone_case(40, 20)
    cfg = Euler2DConfig(grid=grid, boundaries=bcs, geometry="axisymmetric", gamma=cases.GAMMA, flux="hllc",
                        limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5,
                        timeStepping="local")
    g64 = GpuEuler2D.create(ctx, cfg, "fp64")
    with pytest.raises(RangeError):
        g64.setDataflow(True)
    g64.destroy()
    # Global time steps keep the per-stage launches (their time step is a
    # reduction over the grid between steps): the same states either way.
    gcfg = Euler2DConfig(grid=grid, boundaries=bcs, geometry="axisymmetric", gamma=cases.GAMMA, flux="hllc",
                         limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5,
                         timeStepping="global")
    ref, df = _pair(ctx, gcfg, np.array(free))
    ref.iterate(25)
    df.iterate(25)
    assert np.array_equal(_state(df), _state(ref))
    ref.destroy()
    df.destroy()
