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


def _line_cone(ni=40, nj=16):
    """Vertical lines from the 15 deg cone's surface, uniform fractions, the
    top at 25 deg from the tip (inside the conical shock): the first node
    held at the tip."""
    length = 0.5 / math.tan(math.radians(15.0))
    xs = length * np.arange(ni + 1) / ni
    lines = np.stack([xs, xs * math.tan(math.radians(15.0)), np.zeros(ni + 1), np.ones(ni + 1)], axis=1)
    eta = np.tile((np.arange(nj + 1) / nj)[:, None], (1, ni + 1))
    h0 = 0.02 + xs * (math.tan(math.radians(25.0)) - math.tan(math.radians(15.0)))
    return lines, eta, h0, False


def _line_sphere(ni=24, nj=12):
    """Lines normal to a sphere of radius 0.5 from the stagnation point
    (line 0 along the axis, upstream) to 60 deg, stretched fractions: the
    first node normal to the axis."""
    theta = np.radians(60.0) * np.arange(ni + 1) / ni
    lines = np.stack([0.5 - 0.5 * np.cos(theta), 0.5 * np.sin(theta), -np.cos(theta), np.sin(theta)], axis=1)
    b = 1.5
    eta = np.tile((np.expm1(b * np.arange(nj + 1) / nj) / np.expm1(b))[:, None], (1, ni + 1))
    h0 = 0.25 + 0.1 * theta
    return lines, eta, h0, True


def _fitted_pair(ctx, lines, eta, h0, precision, mach):
    from shock_fitting import grid_on_lines
    import pyoracle as O

    grid = grid_on_lines(lines, eta, h0)
    ni = grid.ni
    free, p_inf, q_inf = grids.free_stream_at(mach, GAMMA)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "faceStates", "states": np.tile(np.array(free), (ni, 1))}}
    if lines[0, 1] == 0.0 and lines[0, 3] == 0.0:
        bcs["iMin"] = {"kind": "wall"}   # the axis

    def make(g, states):
        b = dict(bcs, jMax={"kind": "faceStates", "states": np.asarray(states)})
        return O.Euler2D(g, geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                         reconstruction="sharedScaled", time_integrator="ssprk2", cfl=0.5, time_stepping="local",
                         boundaries=b)

    cfg = Euler2DConfig(grid=grid, geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", timeIntegrator="ssprk2", cfl=0.5, timeStepping="local",
                        boundaries=bcs)
    return grid, free, make, GpuEuler2D.create(ctx, cfg, precision)


@pytest.mark.parametrize("shape", ["cone", "sphere"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_fitted_moves_follow_the_oracle(ctx, shape, precision):
    """60 moves of one step each (iterateFitted): heights and flow against
    the oracle's fitShock loop on grids rebuilt and remapped on the CPU."""
    from shock_fitting import fitted_moves

    lines, eta, h0, symmetry = _line_cone() if shape == "cone" else _line_sphere()
    mach = 2.0 if shape == "cone" else 3.0
    grid, free, make, gpu = _fitted_pair(ctx, lines, eta, h0, precision, mach)
    gpu.setLineGrid(lines, eta, "symmetry" if symmetry else "fixed")
    gpu.setHeights(h0)
    gpu.setCellPrimitives(np.array(free))
    gpu.setFittedShock(free)
    moves = 60
    gpu.iterateFitted(moves, 1)
    h_gpu = gpu.readHeights()
    prim0 = np.tile(np.array(free), (grid.ni * grid.nj, 1))
    h_orc, orc, _ = fitted_moves(make, lines, eta, h0, free, GAMMA, prim0, moves, symmetry)
    s = gpu.readState()
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1).astype(np.float64)
    l1, linf = cases.field_errors(q, orc.conserved())
    dh = np.abs(h_gpu - h_orc).max() / np.abs(h_orc).max()
    print(shape, precision, "height moved", np.abs(h_orc - h0).max(), "dh", dh, "L1", l1.max(), "Linf", linf.max())
    assert np.abs(h_orc - h0).max() > 1e-3           # the shock moved
    assert gpu.last_report.steps == moves
    if precision == "fp64":
        assert dh < 1e-11 and l1.max() < 1e-11 and linf.max() < 1e-9
    else:
        assert dh < 1e-5 and l1.max() < 5e-6 and linf.max() < 2e-4


def test_move_grid_remaps_along_the_lines(ctx):
    """moveGrid on the FP64 solver against remapAlongLines of the same
    flow, and the moved metrics against the oracle's for the new nodes."""
    from shock_fitting import grid_on_lines, remap_along_lines
    from bowshock_amd.engine import gpu_primitives
    import pyoracle as O

    lines, eta, h0, _ = _line_sphere()
    grid, free, make, gpu = _fitted_pair(ctx, lines, eta, h0, "fp64", 3.0)
    gpu.setLineGrid(lines, eta, "symmetry")
    gpu.setHeights(h0)
    gpu.setCellPrimitives(np.array(free))
    gpu.iterate(40)
    before = gpu_primitives(gpu.readState(), GAMMA)
    top_cell = h0 * (eta[-1] - eta[-2])
    h1 = h0 + 0.3 * top_cell * np.sin(np.arange(len(h0)))
    gpu.moveGrid(h1)
    after = gpu_primitives(gpu.readState(), GAMMA)
    expect = remap_along_lines(lines, grid.nj, before, grid, grid_on_lines(lines, eta, h1))
    assert np.abs(after - expect).max() < 1e-12 * np.abs(expect).max()
    assert np.allclose(gpu.readHeights(), h1, rtol=0, atol=0)
    # The moved metrics: the residual of a state on the moved grid.
    orc = make(grid_on_lines(lines, eta, h1), np.tile(np.array(free), (grid.ni, 1)))
    orc.set_conserved(np.stack([gpu.readState().rho, gpu.readState().mu, gpu.readState().mv,
                                gpu.readState().energy], axis=1))
    r_gpu, r_orc = gpu.residualVector(), orc.residual_vector()
    scale = np.abs(r_orc).max(axis=0)
    assert (np.abs(r_gpu - r_orc).max(axis=0) / scale).max() < 1e-11
