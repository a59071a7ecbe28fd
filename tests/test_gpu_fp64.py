"""The FP64 solver (bs_solver_create_fp64) against the FP64 oracle.

Both sides compute in FP64 from the same grids and inputs; they differ in
operation order, fused multiply-adds and the GPU's reductions, so per
conserved component (scaled by its largest magnitude) transients agree
within L1 1e-12 and Linf 1e-10, residuals within 1e-11 relative, and the
FP64 wall evaluation within 1e-12. The reference's two-float engine, which
this replaces, holds its residual within about 1e-12 of the CPU's
(src/engine-gpu/euler2d-two-float.ts:1-14).
"""
import ctypes as C
import math

import numpy as np
import pytest

import cases
from bowshock_amd import _native as N
from bowshock_amd import grids
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D, gpu_primitives
from cases import GAMMA

pytestmark = pytest.mark.gpu

L1_TOL, LINF_TOL = 1e-12, 1e-10


def make_pair64(ctx, grid, bcs, **kw):
    orc, _ = cases.make_pair(None, grid, bcs, **kw)
    cfg = Euler2DConfig(grid=grid, geometry=kw.get("geometry", "axisymmetric"), gamma=GAMMA,
                        flux=kw.get("flux", "hllc"), limiter=kw.get("limiter", "minmod"),
                        reconstruction=kw.get("reconstruction", "perVariable"),
                        timeIntegrator=kw.get("integrator", "ssprk3"), cfl=kw.get("cfl", 0.5),
                        timeStepping=kw.get("stepping", "local"), boundaries=bcs)
    return orc, GpuEuler2D.create(ctx, cfg, precision="fp64")


def conserved64(gpu):
    s = gpu.readState()
    assert s.rho.dtype == np.float64
    return np.stack([s.rho, s.mu, s.mv, s.energy], axis=1)


def load_oracle(orc, gpu):
    """The GPU's FP64 state into the oracle, exactly."""
    orc.set_conserved(conserved64(gpu))


def test_precision_is_reported_and_state_round_trips(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(40, 20)
    _, gpu = make_pair64(ctx, grid, bcs)
    p = C.c_int32()
    assert N.lib().bs_solver_precision(gpu._h, C.byref(p)) == 0 and p.value == N.BS_PRECISION_FP64
    rng = np.random.default_rng(5)
    q = [1 + 0.1 * rng.random(gpu.n), rng.random(gpu.n), rng.random(gpu.n), 5 + rng.random(gpu.n)]
    gpu.setState(*q)
    s = gpu.readState()
    for a, b in zip((s.rho, s.mu, s.mv, s.energy), q):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("geometry,w", [("planar", (1.2, 0.7, -0.4, 0.9)), ("axisymmetric", (1.2, 0.7, 0.0, 0.9))])
def test_uniform_flow_preserved_on_distorted_grid(ctx, geometry, w):
    """tests/engine-cpu/euler2d.test.ts:48-71: 1e-13 in FP64, as the
    reference's CPU engine."""
    grid = cases.distorted_grid(12, 10)
    bcs = {s: {"kind": "inflow", "state": w} for s in ("iMin", "iMax", "jMin", "jMax")}
    _, gpu = make_pair64(ctx, grid, bcs, geometry=geometry, stepping="global", cfl=0.8)
    gpu.setInitialCondition(lambda x, y: w)
    gpu.advanceTo(0.5)
    prim = gpu_primitives(gpu.readState(), GAMMA)
    assert np.abs(prim - np.array(w)).max() < 1e-13
    assert abs(gpu.last_report.time - 0.5) < 1e-14


@pytest.mark.parametrize("limiter,reconstruction,flux", [
    ("minmod", "perVariable", "hllc"),
    ("vanLeer", "perVariable", "hllc"),
    ("vanAlbada", "perVariable", "hllc"),
    ("monotonizedCentral", "perVariable", "hllc"),
    ("unlimitedCentral", "perVariable", "hllc"),
    ("firstOrder", "perVariable", "rusanov"),
    ("vanAlbada", "shared", "hllc"),
    ("vanAlbada", "sharedScaled", "hllc"),
    ("vanAlbada", "sharedScaled", "rusanov"),
])
def test_cone_steps_match_oracle(ctx, limiter, reconstruction, flux):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24)
    orc, gpu = make_pair64(ctx, grid, bcs, limiter=limiter, reconstruction=reconstruction, flux=flux)
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    orc.iterate(20)
    gpu.iterate(20)
    l1, linf = cases.field_errors(conserved64(gpu), orc.conserved())
    print("L1", l1, "Linf", linf)
    assert l1.max() < L1_TOL
    assert linf.max() < LINF_TOL


@pytest.mark.parametrize("integrator", ["forwardEuler", "ssprk2", "ssprk3"])
def test_integrators_and_global_steps(ctx, integrator):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(40, 20, mach=3.0)
    kw = dict(limiter="vanAlbada", reconstruction="sharedScaled", integrator=integrator, stepping="global", cfl=0.4)
    orc, gpu = make_pair64(ctx, grid, bcs, **kw)
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    orc.iterate(10)
    gpu.iterate(10)
    l1, linf = cases.field_errors(conserved64(gpu), orc.conserved())
    assert l1.max() < L1_TOL
    assert linf.max() < LINF_TOL
    assert gpu.last_report.steps == 10
    assert abs(gpu.last_report.time / orc.time - 1) < 1e-13


def test_advance_to_lands_on_t_end(ctx):
    grid = cases.distorted_grid(16, 12)
    bcs = {s: {"kind": "periodic"} for s in ("iMin", "iMax", "jMin", "jMax")}
    orc, gpu = make_pair64(ctx, grid, bcs, geometry="planar", limiter="vanLeer", stepping="global", cfl=0.8)

    def field(x, y):
        return (1.0 + 0.2 * math.sin(math.pi * x) * math.sin(2 * math.pi * y), 0.5, 0.2, 1.0)

    orc.set_initial_condition(field)
    gpu.setInitialCondition(field)
    orc.advance_to(0.3)
    gpu.advanceTo(0.3)
    assert abs(gpu.last_report.time - 0.3) < 1e-14
    l1, linf = cases.field_errors(conserved64(gpu), orc.conserved())
    assert l1.max() < L1_TOL
    assert linf.max() < LINF_TOL


def test_far_field_and_face_states(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(40, 20, mach=2.5)
    states = np.tile(np.array(free), (20, 1))
    bcs = {"iMin": {"kind": "faceStates", "states": states}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "farField", "state": free}}
    orc, gpu = make_pair64(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    rho, u, v, p = free
    e = p / (GAMMA - 1) + 0.5 * rho * u * u
    flux = np.tile(np.array([rho * u, rho * u * u + p, 0.0, u * (e + p)]), (20, 1))
    orc.set_face_fluxes("iMin", flux)
    gpu.setFaceFluxes("iMin", flux)
    orc.iterate(15)
    gpu.iterate(15)
    l1, linf = cases.field_errors(conserved64(gpu), orc.conserved())
    assert l1.max() < L1_TOL
    assert linf.max() < LINF_TOL


def test_residual_norm_and_vector(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24)
    orc, gpu = make_pair64(ctx, grid, bcs)
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    assert abs(gpu.densityResidualNorm() / orc.density_residual_norm() - 1) < 1e-11
    gpu.iterate(10)
    load_oracle(orc, gpu)
    r_gpu = gpu.residualVector()
    r_orc = orc.residual_vector()
    scale = np.abs(r_orc).max(axis=0)
    assert (np.abs(r_gpu - r_orc).max(axis=0) / scale).max() < 1e-11
    assert abs(gpu.densityResidualNorm() / orc.density_residual_norm() - 1) < 1e-11


def test_wall_fluxes_and_drag_on_same_state(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24)
    orc, gpu = make_pair64(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    gpu.setCellPrimitives(np.array(free))
    gpu.iterate(40)
    load_oracle(orc, gpu)
    w_gpu = gpu.wallFluxes("jMin")
    w_orc = orc.boundary_fluxes("jMin")
    assert np.abs(w_gpu - w_orc).max() < 1e-12 * max(1.0, np.abs(w_orc).max())
    cd_gpu = gpu.forebodyDrag(p_inf, q_inf, 0.5)
    cd_orc = orc.forebody_drag(p_inf, q_inf, 0.5)
    assert abs(cd_gpu - cd_orc) < 1e-13 * max(1.0, abs(cd_orc))


def test_converged_drag_matches_oracle(ctx):
    """3000 steps on both sides from the free stream: the FP64 iterations
    track each other to the end (the FP32 engine's bar here is 1e-5)."""
    grid, free, p_inf, q_inf, bcs = cases.cone_case(64, 32, mach=3.0)
    orc, gpu = make_pair64(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    orc.iterate(3000)
    gpu.iterate(3000)
    cd_o = orc.forebody_drag(p_inf, q_inf, 0.5)
    cd_g = gpu.forebodyDrag(p_inf, q_inf, 0.5)
    print("cd", cd_o, cd_g, cd_g / cd_o - 1)
    assert abs(cd_g / cd_o - 1) < 1e-11
    l1, linf = cases.field_errors(conserved64(gpu), orc.conserved())
    assert l1.max() < 1e-11 and linf.max() < 1e-9


def test_converges_below_the_fp32_floor(ctx):
    """The single-solve cone (15 deg + 4 calibers, M2, 160 x 80, minmod,
    HLLC, SSP-RK3 at CFL 0.5, local steps): the FP64 solver takes the
    density residual (RMS of R_rho / V, free stream rho = p = 1) to 1e-6
    and its drag to within 1e-5 of the FP32 engine's."""
    grid, free, p_inf, q_inf, bcs = cases.cone_case(160, 80)
    _, gpu = make_pair64(ctx, grid, bcs)
    gpu.setCellPrimitives(np.array(free))
    steps = 0
    res = gpu.densityResidualNorm()
    while res > 1e-6 and steps < 60000:
        gpu.iterate(2000)
        steps += 2000
        res = gpu.densityResidualNorm()
    print("fp64 steps", steps, "residual", res)
    assert res <= 1e-6
    cd64 = gpu.forebodyDrag(p_inf, q_inf, 0.5)
    _, g32 = cases.make_pair(ctx, grid, bcs)
    g32.setCellPrimitives(np.array(free))
    g32.iterate(steps)
    cd32 = g32.forebodyDrag(p_inf, q_inf, 0.5)
    print("fp32 residual", g32.densityResidualNorm(), "cd", cd32, cd64, cd32 / cd64 - 1)
    assert abs(cd32 / cd64 - 1) < 1e-5


@pytest.mark.parametrize("sweeps", [1, 2])
def test_implicit_steps_follow_the_oracle(ctx, sweeps):
    """Red-black steps with the CFL number ramped from 1 against the
    oracle's red-black relaxation with the same linearized neighbour term
    (euler2d-two-float.wgsl:574-598): the FP32 relaxation sets the
    step-by-step bar, the FP32 solver's (L1 5e-6, Linf 2e-4)."""
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=2.0)
    orc, gpu = make_pair64(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    cfl = 1.0
    for _ in range(8):
        orc.iterate_implicit(1, cfl, 1.5, "redBlackLinear", sweeps)
        cfl = min(10.0, cfl * 1.2)
    gpu.setImplicitCfl(1.0)
    gpu.iterateImplicit(8, cfl=10.0, sweeps=sweeps)
    assert gpu.last_report.steps == 8
    l1, linf = cases.field_errors(conserved64(gpu), orc.conserved())
    print("L1", l1, "Linf", linf)
    assert l1.max() < 5e-6
    assert linf.max() < 2e-4


def _oracle_converged(grid, free, bcs, tol=1e-12):
    orc, _ = cases.make_pair(None, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc.set_primitives(np.array(free))
    for _ in range(20):
        orc.iterate(1000)
        if orc.density_residual_norm() < tol:
            break
    assert orc.density_residual_norm() < tol
    return orc


def test_implicit_steps_hold_the_converged_solution(ctx):
    """transonic-gpu.test.ts:172-197 (the two-float engine holds the CPU's
    converged solution, drag within 1e-10, no wander): from the oracle's
    converged state, 4 x 250 red-black steps at CFL 10 keep the drag within
    1e-10 of the oracle's and the density residual below 1e-11; the FP32
    solver's drag wanders by up to ~1e-6 here."""
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=2.0)
    orc = _oracle_converged(grid, free, bcs)
    cd_o = orc.forebody_drag(p_inf, q_inf, 0.5)
    _, gpu = make_pair64(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    q = orc.conserved()
    gpu.setState(q[:, 0], q[:, 1], q[:, 2], q[:, 3])
    gpu.setImplicitCfl(10.0)
    for _ in range(4):
        gpu.iterateImplicit(250, cfl=10.0, sweeps=2)
        cd = gpu.forebodyDrag(p_inf, q_inf, 0.5)
        res = gpu.densityResidualNorm()
        print("drag", cd / cd_o - 1, "residual", res)
        assert abs(cd / cd_o - 1) < 1e-10
        assert res < 1e-11


def test_implicit_converges_from_the_free_stream(ctx):
    """The mixed-precision implicit step's fixed point is the FP64
    residual's: from the free stream the density residual falls below
    1e-11 and the drag reaches the oracle's converged one within 1e-10."""
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=2.0)
    orc = _oracle_converged(grid, free, bcs)
    cd_o = orc.forebody_drag(p_inf, q_inf, 0.5)
    _, gpu = make_pair64(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    gpu.setCellPrimitives(np.array(free))
    gpu.setImplicitCfl(1.0)
    steps, res = 0, math.inf
    while res > 1e-11 and steps < 6000:
        gpu.iterateImplicit(500, cfl=10.0, sweeps=2)
        steps += 500
        res = gpu.densityResidualNorm()
    cd = gpu.forebodyDrag(p_inf, q_inf, 0.5)
    print("steps", steps, "residual", res, "drag", cd / cd_o - 1)
    assert res < 1e-11
    assert abs(cd / cd_o - 1) < 1e-10


def test_verified_drag_levels_equal_the_oracle_converged_drags(ctx):
    """The reference's verified-drag GPU case (20 deg cone, 4 calibers, M3,
    refinements 1 and 2; tests/gpu/verified-drag-gpu.test.ts:15-47) at the
    CPU engine's noise, 1e-12: each settled FP64 level equals the oracle's
    converged drag on the same grid (residual below 2e-13) within 1e-12.
    Against the reference CPU engine's pinned drags
    (verified-drag-gpu.test.ts:26, "measured when written") both the oracle
    and this solver sit 4.2e-9 and 4.4e-9 above; the bar there is 1e-8."""
    import pyoracle as O
    from bowshock_amd.verified_drag import BASE_NJ, CPU_NOISE, verified_forebody_drag

    cone = grids.Body(grids.Nose("conical", 0.5 / math.tan(math.radians(20)), 0.5), 4.0)
    pinned = [0.283146976102, 0.283533123407]
    res = verified_forebody_drag(ctx, cone, mach=3.0, gamma=1.4, tolerance=1e-4, max_refinement=2,
                                 noise=CPU_NOISE, floor=2e-4, order_tolerance=0.5, precision="fp64")
    assert res.kind == "solved", res.reason
    assert [r.refinement for r in res.reports] == [1, 2]
    free, p_inf, q_inf = grids.free_stream_at(3.0, 1.4)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
    for r, c in zip(res.reports, pinned):
        bg = grids.body_grid(cone, BASE_NJ, 3.0, 1.4, refinement=r.refinement)
        orc = O.Euler2D(bg.grid, geometry="axisymmetric", gamma=1.4, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", time_integrator="ssprk2", cfl=0.8, time_stepping="local",
                        boundaries=bcs)
        orc.set_primitives(np.array(free))
        orc.iterate(500 * r.refinement)
        assert orc.density_residual_norm() < 2e-13
        cd_o = orc.forebody_drag(p_inf, q_inf, 0.5)
        print("level", r.refinement, r.drag, cd_o, r.drag / cd_o - 1, r.drag / c - 1, r.iterations)
        assert abs(r.drag / cd_o - 1) < 1e-12
        assert abs(r.drag / c - 1) < 1e-8
