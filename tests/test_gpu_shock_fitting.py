"""The fitted-shock boundary on the GPU (bs_set_fitted_shock,
bs_read_shock; euler2d.wgsl:402-454) against the oracle's restatement of
src/solver/shock-fitting.ts (compatibleShock, shockBoundary): explicit
steps with the jMax faces set from the flow at the start of every step,
the grid held.

The case: the 15 deg cone at Mach 2 on a grid whose jMax boundary runs
at 25 deg from the tip, inside the conical shock (about 32 deg), so the
fitted boundary carries a shock of normal Mach number above 1 once the
cone has compressed the flow below it. Bars: the FP32 solver's transient
bar (L1 5e-6, Linf 2e-4; shock speeds 1e-4 relative to the sound speed),
the FP64 solver's (L1 1e-12, Linf 1e-10; speeds 1e-10).
"""
import math

import numpy as np
import pytest

import cases
from bowshock_amd import grids
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D, RangeError, UnsupportedError
from cases import GAMMA
from shock_fitting import fitted_steps

pytestmark = pytest.mark.gpu


def fitted_case(ni=48, nj=16):
    length = 0.5 / math.tan(math.radians(15.0))
    tan_wall, tan_top = math.tan(math.radians(15.0)), math.tan(math.radians(25.0))

    def node(i, j):
        x = length * i / ni
        wall = x * tan_wall
        top = 0.02 + x * tan_top
        return x, wall + (top - wall) * j / nj

    grid = grids.grid_from_function(ni, nj, node)
    free, p_inf, q_inf = grids.free_stream_at(2.0, GAMMA)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "faceStates", "states": np.tile(np.array(free), (ni, 1))}}
    return grid, free, bcs


def _pair(ctx, grid, bcs, precision):
    orc, _ = cases.make_pair(None, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled",
                             integrator="ssprk2", cfl=0.5)
    cfg = Euler2DConfig(grid=grid, geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", timeIntegrator="ssprk2", cfl=0.5, timeStepping="local",
                        boundaries=bcs)
    return orc, GpuEuler2D.create(ctx, cfg, precision)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_fitted_steps_follow_the_oracle(ctx, precision):
    grid, free, bcs = fitted_case()
    orc, gpu = _pair(ctx, grid, bcs, precision)
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    gpu.setFittedShock(free)
    steps = 120
    gpu.iterate(steps)
    last = fitted_steps(orc, grid, free, GAMMA, steps)
    s = gpu.readState()
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1).astype(np.float64)
    l1, linf = cases.field_errors(q, orc.conserved())
    shock = gpu.readShock()
    a_free = math.sqrt(GAMMA * free[3] / free[0])
    speed_err = np.abs(shock["speed"] - last[2]).max() / a_free
    print(precision, "L1", l1, "Linf", linf, "speed", speed_err, "mach max", last[3].max())
    assert last[3].max() > 1.05            # the boundary carries a shock
    if precision == "fp64":
        assert l1.max() < 1e-12 and linf.max() < 1e-10
        assert speed_err < 1e-10
        assert np.abs(shock["mach"] - last[3]).max() < 1e-10
    else:
        assert l1.max() < 5e-6 and linf.max() < 2e-4
        assert speed_err < 1e-4
    assert np.abs(shock["signal"] - last[4]).max() < (1e-10 if precision == "fp64" else 1e-5) * last[4].max()


def test_fitted_shock_rules(ctx):
    grid, free, bcs = fitted_case()
    _, gpu = _pair(ctx, grid, bcs, "fp32")
    gpu.setCellPrimitives(np.array(free))
    with pytest.raises(RangeError):
        gpu.readShock()
    gpu.setFittedShock(free)
    with pytest.raises(UnsupportedError):
        gpu.iterateImplicit(1, cfl=10.0, sweeps=2)
    gpu.setFittedShock(None)
    gpu.iterateImplicit(1, cfl=10.0, sweeps=2)
    plain = dict(bcs, jMax={"kind": "inflow", "state": free})
    _, g2 = _pair(ctx, grid, plain, "fp32")
    with pytest.raises(RangeError):
        g2.setFittedShock(free)
