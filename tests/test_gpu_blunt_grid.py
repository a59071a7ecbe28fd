"""The HIP solver on the blunt-body C grid (bowshock_amd/blunt_grid.py):
the captured hemisphere-cylinder of the reference's
tests/engine-cpu/blunt-body.test.ts (Mach 3, 20 cells across, the axis
ahead of the nose a wall) and the fitted shock moving on the generator's
lines with fixed node fractions (the fitted flow's grids,
src/analysis/fitted-flow.ts bluntFitFamily).

Bars: the FP64 solver converges to round-off (density residual 1e-13 of
its start, as the reference's CPU engine does) onto the oracle's converged
field within 1e-10, and stands the shock off within 10 % of Billig's
correlation and within 1e-4 of the 0.1078 the reference recorded; the
FP32 solver follows the oracle's transient within its bar (L1 5e-6, Linf
2e-4); fitted moves as in the cone and sphere cases (FP64 heights 1e-11,
fields L1 1e-11, Linf 1e-9).
"""
import math

import numpy as np
import pytest

import cases
from bowshock_amd import grids
from bowshock_amd.blunt_grid import billig_standoff, blunt_body_grid, with_fixed_fractions
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D
from cases import GAMMA

pytestmark = pytest.mark.gpu

MACH = 3.0
HEMI = grids.Body(grids.Nose("ellipsoid", 0.5, 0.5), 4.0)
FREE = (1.0, MACH * math.sqrt(GAMMA), 0.0, 1.0)


def _captured(ctx, precision):
    import pyoracle as O

    grid = blunt_body_grid(HEMI, 20, MACH).grid
    bcs = {"iMin": {"kind": "wall"}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
           "jMax": {"kind": "inflow", "state": FREE}}
    scheme = dict(geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                  reconstruction="sharedScaled")
    orc = O.Euler2D(grid, time_integrator="ssprk2", cfl=0.8, time_stepping="local", boundaries=bcs, **scheme)
    cfg = Euler2DConfig(grid=grid, timeIntegrator="ssprk2", cfl=0.8, timeStepping="local", boundaries=bcs, **scheme)
    return grid, orc, GpuEuler2D.create(ctx, cfg, precision)


def _standoff(grid, prim, cx):
    """Where the pressure on the stagnation line crosses halfway through
    the normal shock (blunt-body.test.ts:56-72)."""
    ratio = 1 + 2 * GAMMA / (GAMMA + 1) * (MACH * MACH - 1)
    half = 0.5 * (1 + ratio)
    ni = grid.ni
    for j in range(grid.nj - 1):
        p0, p1 = prim[j * ni, 3], prim[(j + 1) * ni, 3]
        if p0 > half >= p1:
            x0, x1 = cx[j * ni], cx[(j + 1) * ni]
            return -(x0 + (x1 - x0) * (p0 - half) / (p0 - p1))
    return math.nan


def test_captured_hemisphere_converges_onto_the_oracle(ctx):
    from bowshock_amd.engine import gpu_primitives

    grid, orc, gpu = _captured(ctx, "fp64")
    gpu.setCellPrimitives(np.array(FREE))
    r0 = gpu.densityResidualNorm()
    gpu.iterate(6000)
    ratio = gpu.densityResidualNorm() / r0
    orc.set_primitives(np.array(FREE))
    orc.iterate(6000)
    s = gpu.readState()
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1)
    l1, linf = cases.field_errors(q, orc.conserved())
    prim = gpu_primitives(s, GAMMA)
    standoff = _standoff(grid, prim, orc.metrics["cx"])
    # blunt-body.test.ts:51-53: the bow shock stays inside the outer
    # boundary, the free stream along it (verified-drag.ts outerDeviation).
    from bowshock_amd.verified_drag import outer_deviation

    assert outer_deviation(prim.reshape(grid.nj, grid.ni, 4), FREE, GAMMA) < 1e-5
    print("residual ratio", ratio, "L1", l1.max(), "Linf", linf.max(), "standoff", standoff,
          "Billig", billig_standoff(MACH, 0.5))
    assert ratio < 1e-13
    assert linf.max() < 1e-10
    assert abs(standoff / billig_standoff(MACH, 0.5) - 1) < 0.1
    # The reference's CPU engine gave 0.1078 on its grid (blunt-body.test.ts:58).
    assert abs(standoff - 0.1078) < 1e-4


def test_captured_hemisphere_fp32_transient(ctx):
    grid, orc, gpu = _captured(ctx, "fp32")
    gpu.setCellPrimitives(np.array(FREE))
    orc.set_primitives(np.array(FREE))
    gpu.iterate(300)
    orc.iterate(300)
    s = gpu.readState()
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1).astype(np.float64)
    l1, linf = cases.field_errors(q, orc.conserved())
    print("L1", l1, "Linf", linf)
    assert l1.max() < 5e-6 and linf.max() < 2e-4


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_fitted_moves_on_the_generator_lines(ctx, precision):
    """The generator's lines with their node fractions fixed at 0.55 of
    their heights (withFixedFractions), first node normal to the axis:
    60 moves of one step from the free stream against the oracle's fitShock
    loop on grids rebuilt and remapped on the CPU."""
    import pyoracle as O
    from shock_fitting import fitted_moves, grid_on_lines

    nj = 16
    b = blunt_body_grid(HEMI, nj, MACH)
    h0 = 0.55 * b.lines.h
    fixed = with_fixed_fractions(b.lines.with_heights(h0), nj, 0.5)
    lines, eta = fixed.origins_and_directions(), fixed.fractions
    grid = grid_on_lines(lines, eta, h0)
    ni = grid.ni
    free = grids.free_stream_at(MACH, GAMMA)[0]
    bcs = {"iMin": {"kind": "wall"}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
           "jMax": {"kind": "faceStates", "states": np.tile(np.array(free), (ni, 1))}}
    scheme = dict(geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                  reconstruction="sharedScaled", cfl=0.5)

    def make(g, states):
        return O.Euler2D(g, time_integrator="ssprk2", time_stepping="local",
                         boundaries=dict(bcs, jMax={"kind": "faceStates", "states": np.asarray(states)}), **scheme)

    cfg = Euler2DConfig(grid=grid, timeIntegrator="ssprk2", timeStepping="local", boundaries=bcs, **scheme)
    gpu = GpuEuler2D.create(ctx, cfg, precision)
    gpu.setLineGrid(lines, eta, "symmetry")
    gpu.setHeights(h0)
    gpu.setCellPrimitives(np.array(free))
    gpu.setFittedShock(free)
    moves = 60
    gpu.iterateFitted(moves, 1)
    h_gpu = gpu.readHeights()
    prim0 = np.tile(np.array(free), (ni * nj, 1))
    h_orc, orc, _ = fitted_moves(make, lines, eta, h0, free, GAMMA, prim0, moves, True)
    s = gpu.readState()
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1).astype(np.float64)
    l1, linf = cases.field_errors(q, orc.conserved())
    dh = np.abs(h_gpu - h_orc).max() / np.abs(h_orc).max()
    print(precision, "ni", ni, "moved", np.abs(h_orc - h0).max(), "dh", dh, "L1", l1.max(), "Linf", linf.max())
    assert np.abs(h_orc - h0).max() > 1e-2 and h_orc.min() > 0
    if precision == "fp64":
        assert dh < 1e-11 and l1.max() < 1e-11 and linf.max() < 1e-9
    else:
        assert dh < 1e-5 and l1.max() < 5e-6 and linf.max() < 2e-4


def _line_solver(ctx, lines, eta, heights, precision, free):
    from shock_fitting import grid_on_lines

    grid = grid_on_lines(lines, eta, heights)
    ni = grid.ni
    bcs = {"iMin": {"kind": "wall"}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
           "jMax": {"kind": "inflow", "state": free}}
    cfg = Euler2DConfig(grid=grid, geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", timeIntegrator="ssprk2", cfl=0.5, timeStepping="local",
                        boundaries=bcs)
    return grid, bcs, GpuEuler2D.create(ctx, cfg, precision)


def _residual_error(gpu, grid, bcs, heights_grid):
    """The solver's residual of its own state against the oracle's on the
    grid the metrics should describe: equal metrics give equal residuals."""
    import pyoracle as O

    s = gpu.readState()
    orc = O.Euler2D(heights_grid, geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                    reconstruction="sharedScaled", time_integrator="ssprk2", cfl=0.5, time_stepping="local",
                    boundaries=bcs)
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1).astype(np.float64)
    if s.rho.dtype == np.float32:
        from bowshock_amd.engine import gpu_primitives

        orc.set_primitives(gpu_primitives(s, GAMMA))
    else:
        orc.set_conserved(q)
    r_gpu, r_orc = gpu.residualVector(), orc.residual_vector()
    return (np.abs(r_gpu - r_orc).max(axis=0) / np.abs(r_orc).max(axis=0)).max()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_line_metrics_on_the_c_grid_heights_moved(ctx, precision):
    """line-grid-gpu.test.ts:113-127: the hemisphere-cylinder's C grid with
    fixed fractions, heights moved 5 %: the metrics the GPU computes from
    the heights give the oracle's residual on the moved nodes (FP64 1e-11;
    FP32, whose metrics are rounded to FP32, 2e-4)."""
    from shock_fitting import grid_on_lines

    nj = 20
    b = blunt_body_grid(HEMI, nj, MACH)
    fixed = with_fixed_fractions(b.lines, nj, 0.5)
    lines, eta = fixed.origins_and_directions(), fixed.fractions
    free = grids.free_stream_at(MACH, GAMMA)[0]
    grid, bcs, gpu = _line_solver(ctx, lines, eta, fixed.h, precision, free)
    gpu.setLineGrid(lines, eta, "symmetry")
    h1 = 1.05 * fixed.h
    gpu.setHeights(h1)
    rng = np.random.default_rng(4)
    prim = np.array(free) * (1 + 0.05 * rng.standard_normal((grid.ni * grid.nj, 4)))
    prim[:, 2] = 0.1 * rng.standard_normal(grid.ni * grid.nj)
    gpu.setCellPrimitives(prim)
    err = _residual_error(gpu, grid, bcs, grid_on_lines(lines, eta, h1))
    print(precision, "residual error on the moved grid", err)
    assert err < (1e-11 if precision == "fp64" else 2e-4)


def test_line_metrics_behind_a_sharp_tip(ctx):
    """line-grid-gpu.test.ts:128-141: vertical lines behind a 15 deg cone's
    tip, spaced geometrically from 1e-3 of the nose (cells 1e3 times
    smaller at the tip than at the base): FP64 metrics from the heights
    give the oracle's residual within 1e-11."""
    from shock_fitting import grid_on_lines

    length = 0.5 / math.tan(math.radians(15.0))
    ni, nj = 40, 12
    xs = np.concatenate([[0.0], length * 1e-3 * (1e3 ** (np.arange(ni) / (ni - 1)))])
    lines = np.stack([xs, xs * math.tan(math.radians(15.0)), np.zeros(ni + 1), np.ones(ni + 1)], axis=1)
    eta = np.tile((np.arange(nj + 1) / nj)[:, None], (1, ni + 1))
    h0 = 0.02 + xs * 0.3
    free = grids.free_stream_at(2.0, GAMMA)[0]
    grid, bcs, gpu = _line_solver(ctx, lines, eta, h0, "fp64", free)
    gpu.setLineGrid(lines, eta, "fixed")
    h1 = h0 * 1.03
    gpu.setHeights(h1)
    rng = np.random.default_rng(6)
    prim = np.array(free) * (1 + 0.05 * rng.standard_normal((grid.ni * grid.nj, 4)))
    gpu.setCellPrimitives(prim)
    err = _residual_error(gpu, grid, bcs, grid_on_lines(lines, eta, h1))
    print("residual error behind the sharp tip", err)
    assert err < 1e-11


def test_remap_on_the_c_grid(ctx):
    """line-grid-gpu.test.ts:144-: the flow remapped as the C grid's lines
    move 2 %, against the oracle's remapAlongLines of the same flow."""
    from bowshock_amd.engine import gpu_primitives
    from shock_fitting import grid_on_lines, remap_along_lines

    nj = 20
    b = blunt_body_grid(HEMI, nj, MACH)
    fixed = with_fixed_fractions(b.lines, nj, 0.5)
    lines, eta = fixed.origins_and_directions(), fixed.fractions
    free = grids.free_stream_at(MACH, GAMMA)[0]
    grid, bcs, gpu = _line_solver(ctx, lines, eta, fixed.h, "fp64", free)
    gpu.setLineGrid(lines, eta, "symmetry")
    gpu.setHeights(fixed.h)
    gpu.setCellPrimitives(np.array(free))
    gpu.iterate(200)
    before = gpu_primitives(gpu.readState(), GAMMA)
    h1 = fixed.h * (1 + 0.02 * np.cos(np.arange(len(fixed.h))))
    gpu.moveGrid(h1)
    after = gpu_primitives(gpu.readState(), GAMMA)
    expect = remap_along_lines(lines, nj, before, grid, grid_on_lines(lines, eta, h1))
    assert np.abs(after - expect).max() < 1e-12 * np.abs(expect).max()


def test_fitted_hemisphere_from_the_captured_start(ctx):
    """The fitted bow shock on the hemisphere-cylinder at Mach 3, end to
    end on the GPU as the reference's fitted flow runs it
    (src/analysis/fitted-flow.ts bluntStart, bluntFitFamily; shock-
    fitting.ts fitShock): the captured FP64 solution on the generator's C
    grid, the fitted start on its lines (bowshock_amd/fitted_start.py:
    shock heights at half the pressure rise, smoothed; fixed fractions; the
    flow at the same distance along each line), then 6,000 moves of two
    steps with the shock fitted and the grid moving on the GPU. Bars: the
    fitted grid meets the fitted grids' gates; the shock settles (speeds
    below 1e-9 of the free stream's sound speed, the residual 1e-10 of its
    first reading); the standoff within 1 % of the captured one and 10 % of
    Billig's; the drag within 1 % of the captured grid's."""
    import pyoracle as O
    from bowshock_amd.engine import gpu_primitives
    from bowshock_amd.fitted_start import fitted_start

    nj = 20
    bg = blunt_body_grid(HEMI, nj, MACH)
    grid, ni = bg.grid, bg.grid.ni
    free, p_inf, q_inf = grids.free_stream_at(MACH, GAMMA)
    scheme = dict(geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                  reconstruction="sharedScaled", timeIntegrator="ssprk2", timeStepping="local")
    bcs = {"iMin": {"kind": "wall"}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
           "jMax": {"kind": "inflow", "state": free}}
    cap = GpuEuler2D.create(ctx, Euler2DConfig(grid=grid, boundaries=bcs, cfl=0.8, **scheme), "fp64")
    cap.setCellPrimitives(np.array(free))
    cap.iterate(6000)
    cd_cap = cap.forebodyDrag(p_inf, q_inf, 0.5)
    prim = gpu_primitives(cap.readState(), GAMMA)
    st = fitted_start(bg.lines, grid, prim, 0.5)
    fixed, heights, fg = st.lines, st.heights, st.grid
    assert grids.quality_failures(grids.grid_quality(fg), grids.FITTED_QUALITY_LIMITS) == []
    m = O.compute_metrics(grid, True)
    fbcs = dict(bcs, jMax={"kind": "faceStates", "states": np.tile(np.array(free), (ni, 1))})
    fit = GpuEuler2D.create(ctx, Euler2DConfig(grid=fg, boundaries=fbcs, cfl=0.5, **scheme), "fp64")
    fit.setLineGrid(fixed.origins_and_directions(), fixed.fractions, "symmetry")
    fit.setHeights(heights)
    fit.setCellPrimitives(st.primitives)
    fit.setFittedShock(free)
    fit.iterateFitted(500, 2)
    r_first = fit.densityResidualNorm()
    fit.iterateFitted(5500, 2)
    h = fit.readHeights()
    speed = np.abs(fit.readShock()["speed"]).max() / math.sqrt(GAMMA)
    res = fit.densityResidualNorm() / r_first
    cd = fit.forebodyDrag(p_inf, q_inf, 0.5)
    standoff_cap = _standoff(grid, prim, m["cx"])
    print("standoff", h[0], "captured", standoff_cap, "Billig", billig_standoff(MACH, 0.5), "cd", cd, "captured",
          cd_cap, "residual", res, "speed", speed)
    assert speed < 1e-9 and res < 1e-10
    assert abs(h[0] / standoff_cap - 1) < 0.01 and abs(h[0] / billig_standoff(MACH, 0.5) - 1) < 0.1
    assert abs(cd / cd_cap - 1) < 0.01


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("kind", ["H", "C"])
def test_free_stream_preserved_on_generated_grids(ctx, precision, kind):
    """euler2d-gpu.test.ts:117-165: every side imposing the free stream
    (the body too), the uniform stream is the exact solution on the
    generated grids, skewed cells included: 50 steps keep it to FP32
    round-off (1e-5, the reference's bar) or FP64 round-off (1e-12). The H
    grid of a 3 deg cone at Mach 1.1 and the C grid of the
    hemisphere-cylinder at Mach 3 (its iMin on the axis)."""
    from bowshock_amd.engine import gpu_primitives

    if kind == "H":
        mach = 1.1
        body = grids.Body(grids.Nose("conical", 0.5 / math.tan(math.radians(3.0)), 0.5), 4.0)
        grid = grids.body_grid(body, 20, mach, GAMMA).grid
    else:
        mach = MACH
        grid = blunt_body_grid(HEMI, 20, mach).grid
    free = (1.0, mach * math.sqrt(GAMMA), 0.0, 1.0)
    inflow = {"kind": "inflow", "state": free}
    bcs = {"iMin": inflow, "iMax": {"kind": "outflow"}, "jMin": inflow, "jMax": inflow}
    cfg = Euler2DConfig(grid=grid, geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", timeIntegrator="ssprk2", cfl=0.8, timeStepping="local",
                        boundaries=bcs)
    gpu = GpuEuler2D.create(ctx, cfg, precision)
    gpu.setCellPrimitives(np.array(free))
    gpu.iterate(50)
    w = gpu_primitives(gpu.readState(), GAMMA)
    bar = 1e-12 if precision == "fp64" else 1e-5
    err = max(np.abs(w[:, 0] - 1).max(), np.abs(w[:, 3] - 1).max(), np.abs(w[:, 2]).max() / free[1])
    print(kind, precision, "largest deviation", err)
    assert err < bar
