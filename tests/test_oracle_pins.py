"""Pins of the FP64 oracle (CPU): the reference's golden values, closed forms
and invariants that a plausible mistake (a dropped term, a wrong sign or
index, a transposed operand) would break.

The reference's TypeScript engine cannot run here (no TypeScript toolchain,
Node 12); these checks are what ties the oracle to it.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import cases
import pyoracle as O
from bowshock_amd import grids
from cases import GAMMA

DEG = math.pi / 180


def _verified_case(body, mach, nj=20, refinement=1, jmax="inflow"):
    """The case verifiedForebodyDrag solves (verified-drag.ts:241-257)."""
    bg = grids.body_grid(body, nj, mach, GAMMA, refinement=refinement)
    free, p_inf, q_inf = grids.free_stream_at(mach, GAMMA)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": jmax, "state": free}}
    s = O.Euler2D(bg.grid, flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled",
                  time_integrator="ssprk2", cfl=0.8, time_stepping="local", boundaries=bcs)
    s.set_primitives(np.array(free))
    return s, bg, free, p_inf, q_inf


@pytest.mark.parametrize("refinement,golden", [(1, 0.283146976102), (2, 0.283533123407)])
def test_golden_cone_drags(refinement, golden):
    """tests/gpu/verified-drag-gpu.test.ts:17-44: the reference's CPU engine
    on bodyGrid(nj = 20, r) of a 20 deg cone with 4 calibers of cylinder at
    M3 gave these drags once settled to 1e-6 noise. The oracle, fully
    converged, on the restated bodyGrid, must match well inside that noise."""
    body = grids.cone_cylinder(20.0, 4.0)
    s, bg, free, p_inf, q_inf = _verified_case(body, 3.0, refinement=refinement)
    assert (bg.grid.ni, bg.grid.nj) == (100 * refinement, 20 * refinement)
    s.iterate(1000 * refinement)
    cd = s.forebody_drag(p_inf, q_inf, 0.5)
    assert abs(cd / golden - 1) < 2e-8


@pytest.mark.parametrize("geometry,w", [("planar", (1.2, 0.7, -0.4, 0.9)), ("axisymmetric", (1.2, 0.7, 0.0, 0.9))])
def test_uniform_flow_preserved_on_distorted_grid(geometry, w):
    """tests/engine-cpu/euler2d.test.ts:48-71."""
    grid = cases.distorted_grid(12, 10)
    bcs = {s: {"kind": "inflow", "state": w} for s in ("iMin", "iMax", "jMin", "jMax")}
    s = O.Euler2D(grid, geometry=geometry, limiter="vanLeer", time_integrator="ssprk2", cfl=0.8,
                  time_stepping="global", boundaries=bcs)
    s.set_initial_condition(lambda x, y: w)
    s.advance_to(0.5)
    assert abs(s.time - 0.5) < 1e-15
    assert np.abs(s.primitives() - np.array(w)).max() < 1e-13


def test_metrics_of_a_rectangle_in_closed_form():
    ni, nj = 5, 4
    g = grids.rectangular_grid(ni, nj, 0.0, 2.0, 1.0, 3.0)
    dx, dy = 2.0 / ni, 2.0 / nj
    m = O.compute_metrics(g, False)
    assert np.allclose(m["inx"], 1.0) and np.allclose(m["iny"], 0.0)
    assert np.allclose(m["jnx"], 0.0) and np.allclose(m["jny"], 1.0)
    assert np.allclose(m["iarea"], dy) and np.allclose(m["jarea"], dx)
    assert np.allclose(m["volume"], dx * dy)
    ma = O.compute_metrics(g, True)
    yc = 1.0 + dy * (np.arange(nj) + 0.5)
    assert np.allclose(ma["cy"].reshape(nj, ni), yc[:, None])
    assert np.allclose(ma["volume"].reshape(nj, ni), dx * dy * yc[:, None])
    # i-face (i, j) from node (i, j) to (i, j + 1): area = dy * mean radius.
    assert np.allclose(ma["iarea"].reshape(nj, ni + 1), (dy * yc)[:, None])
    yj = 1.0 + dy * np.arange(nj + 1)
    assert np.allclose(ma["jarea"].reshape(nj + 1, ni), dx * yj[:, None])


def test_closed_cells_and_well_balanced_axisymmetric_metrics():
    """Planar faces of every cell close (sum of n A = 0); axisymmetric ones
    leave exactly the planar area in y (euler2d.ts:955-960), which the
    source p A_p balances."""
    g = cases.distorted_grid(9, 7)
    for axisym in (False, True):
        m = O.compute_metrics(g, axisym)
        ni, nj = 9, 7
        inx, iny, ia = (m[k].reshape(nj, ni + 1) for k in ("inx", "iny", "iarea"))
        jnx, jny, ja = (m[k].reshape(nj + 1, ni) for k in ("jnx", "jny", "jarea"))
        sx = inx[:, :-1] * ia[:, :-1] - inx[:, 1:] * ia[:, 1:] + jnx[:-1] * ja[:-1] - jnx[1:] * ja[1:]
        sy = iny[:, :-1] * ia[:, :-1] - iny[:, 1:] * ia[:, 1:] + jny[:-1] * ja[:-1] - jny[1:] * ja[1:]
        assert np.abs(sx).max() < 1e-14
        expect = m["planar"].reshape(nj, ni) if axisym else 0.0
        assert np.abs(-sy - expect).max() < 1e-14


def test_time_step_closed_form():
    """tests/engine-cpu/euler2d.test.ts:141-153."""
    w = (1.3, 0.4, -0.7, 0.9)
    g = grids.rectangular_grid(10, 100, 0.0, 1.0, 0.0, 1.0)
    bcs = {s: {"kind": "inflow", "state": w} for s in ("iMin", "iMax", "jMin", "jMax")}
    s = O.Euler2D(g, geometry="planar", cfl=0.8, time_stepping="global", boundaries=bcs)
    s.set_primitives(np.array(w))
    a = math.sqrt(GAMMA * w[3] / w[0])
    dx, dy = 0.1, 0.01
    expected = (0.8 * dx * dy) / ((abs(w[1]) + a) * dy + (abs(w[2]) + a) * dx)
    assert abs(s.stable_time_step() - expected) < 1e-15


def _physical_face_flux(w, g=GAMMA):
    rho, un, ut, p = w
    e = p / (g - 1) + 0.5 * rho * (un * un + ut * ut)
    return np.array([rho * un, rho * un * un + p, rho * un * ut, un * (e + p)])


@pytest.mark.parametrize("flux", ["hllc", "rusanov", "hll"])
def test_flux_consistency(flux):
    """F(w, w) is the physical flux (tests/scheme/flux.test.ts:41-60)."""
    rng = np.random.default_rng(31)
    for _ in range(200):
        w = np.array([rng.uniform(0.1, 3), rng.uniform(-3, 3), rng.uniform(-2, 2), rng.uniform(0.1, 3)])
        assert np.allclose(O.face_frame_flux(flux, w, w, GAMMA), _physical_face_flux(w), rtol=1e-13, atol=1e-13)


def test_hllc_is_upwind_for_supersonic_flow_and_exact_on_contacts():
    rng = np.random.default_rng(33)
    for _ in range(200):
        l = np.array([rng.uniform(0.5, 2), rng.uniform(4, 6), rng.uniform(-1, 1), rng.uniform(0.5, 2)])
        r = np.array([rng.uniform(0.5, 2), rng.uniform(4, 6), rng.uniform(-1, 1), rng.uniform(0.5, 2)])
        # Both states supersonic to the right: the left state's flux.
        if l[1] - math.sqrt(GAMMA * l[3] / l[0]) > 0.5:
            assert np.allclose(O.face_frame_flux("hllc", l, r, GAMMA), _physical_face_flux(l), rtol=1e-13)
    # A stationary contact with shear: pressure only, no mass flux
    # (tests/scheme/flux.test.ts contact case).
    f = O.face_frame_flux("hllc", [1.0, 0.0, 0.3, 1.0], [0.2, 0.0, -0.5, 1.0], GAMMA)
    assert np.allclose(f, [0.0, 1.0, 0.0, 0.0], atol=1e-14)


def test_hllc_rotational_invariance_and_galilean_frame():
    """faceFlux2D of rotated states equals the rotated flux
    (tests/scheme/flux2d.test.ts:57-152)."""
    rng = np.random.default_rng(37)
    for _ in range(100):
        l = np.array([rng.uniform(0.2, 2), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.2, 2)])
        r = np.array([rng.uniform(0.2, 2), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.2, 2)])
        t = rng.uniform(0, 2 * math.pi)
        c, s = math.cos(t), math.sin(t)
        f0 = O.face_flux_2d("hllc", l, r, 1.0, 0.0, GAMMA)
        rot = lambda w: np.array([w[0], c * w[1] - s * w[2], s * w[1] + c * w[2], w[3]])
        f1 = O.face_flux_2d("hllc", rot(l), rot(r), c, s, GAMMA)
        assert np.allclose(f1, rot(f0), rtol=1e-12, atol=1e-12)


def test_hllc_close_to_exact_godunov_for_weak_jumps():
    """Agreement with the exact Riemann solver's flux to second order in
    the jump (tests/scheme/flux.test.ts:99-113), the exact solution from
    Toro's star-state iteration written here with scipy."""
    from scipy.optimize import brentq

    def exact_flux(l, r, g=GAMMA):
        def f(p, w):
            rho, u, pk = w
            a = math.sqrt(g * pk / rho)
            if p > pk:
                A, B = 2 / ((g + 1) * rho), (g - 1) / (g + 1) * pk
                return (p - pk) * math.sqrt(A / (p + B))
            return 2 * a / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

        L, R = (l[0], l[1], l[3]), (r[0], r[1], r[3])
        ps = brentq(lambda p: f(p, L) + f(p, R) + R[1] - L[1], 1e-10, 100)
        us = 0.5 * (L[1] + R[1]) + 0.5 * (f(ps, R) - f(ps, L))
        # Subsonic states here: the face lies in a star region.
        side = L if us >= 0 else R
        ut = l[2] if us >= 0 else r[2]
        rho, u, pk = side
        if ps > pk:
            rs = rho * ((ps / pk + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pk + 1))
        else:
            rs = rho * (ps / pk) ** (1 / g)
        return _physical_face_flux((rs, us, ut, ps))

    base = np.array([1.0, 0.3, 0.1, 1.0])
    errs = []
    for h in (1e-2, 5e-3):
        l = base
        r = base + h * np.array([0.7, -0.5, 0.2, 0.9])
        errs.append(np.abs(O.face_frame_flux("hllc", l, r, GAMMA) - exact_flux(l, r)).max())
    assert errs[0] < 5e-3 * 1e-2
    assert errs[0] / errs[1] > 3.0  # second order: 4 when halving the jump


@pytest.mark.parametrize("name", ["minmod", "vanLeer", "monotonizedCentral"])
def test_tvd_limiters(name):
    """Zero at extrema, symmetric, exact on linear data, inside the TVD
    region (tests/scheme/limiters.test.ts)."""
    rng = np.random.default_rng(41)
    for _ in range(500):
        a, b = rng.uniform(-2, 2, size=2)
        phi = O.limiter(name, a, b)
        if a * b <= 0:
            assert phi == 0.0
            continue
        assert abs(phi - O.limiter(name, b, a)) < 1e-15
        assert abs(O.limiter(name, a, a) - a) < 1e-15
        s = abs(phi)
        assert s <= 2 * min(abs(a), abs(b)) + 1e-15 and np.sign(phi) == np.sign(a)
    assert O.limiter("minmod", 1.0, 3.0) == 1.0 and O.limiter("minmod", -3.0, -1.0) == -1.0


def test_shared_factor_bounds_and_unit_invariance():
    """reconstruction.ts:70-151: the shared factor scales the central slope
    by f in [0, 1]; the scaled form is independent of the units of rho and
    p (tests/scheme/reconstruction.test.ts:29-173)."""
    rng = np.random.default_rng(43)
    for _ in range(200):
        p, c, n = (np.array([rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2)])
                   for _ in range(3))
        for rec in ("shared", "sharedScaled"):
            face = O.reconstruct(rec, "vanAlbada", p, c, n, 1, GAMMA)
            k = (face - c) / (0.25 * (n - p))
            f = k[np.abs(n - p) > 1e-6]
            assert np.all(f >= -1e-12) and np.all(f <= 1 + 1e-12)
            assert np.allclose(f, f[0])
        # Units: rho by 7, p by 7 * 3^2, velocities by 3 leaves the factor.
        sc = np.array([7.0, 3.0, 3.0, 63.0])
        f1 = O.reconstruct("sharedScaled", "vanAlbada", p, c, n, -1, GAMMA)
        f2 = O.reconstruct("sharedScaled", "vanAlbada", p * sc, c * sc, n * sc, -1, GAMMA)
        assert np.allclose(f2, f1 * sc, rtol=1e-12)


def test_far_field_state_invariants():
    """euler2d.ts:130-138: a free stream inside gives it back; supersonic
    inflow takes the free stream, supersonic outflow the inside state
    (tests/engine-cpu/implicit.test.ts:21-111)."""
    free = np.array([1.0, 2.0, 0.0, 1.0])
    for nx, ny in ((1.0, 0.0), (0.6, 0.8), (0.0, -1.0)):
        assert np.allclose(O.far_field_state(free, free, nx, ny, GAMMA), free, rtol=1e-13, atol=1e-13)
    inside = np.array([1.3, 3.5, 0.2, 1.4])  # supersonic, leaving through +x
    assert np.allclose(O.far_field_state(inside, free, 1.0, 0.0, GAMMA), inside, rtol=1e-13)
    # Subsonic outflow: pressure from the free stream's invariant, entropy
    # and tangential velocity from inside.
    sub = np.array([1.0, 0.5, 0.3, 1.0])
    w = O.far_field_state(sub, free, 1.0, 0.0, GAMMA)
    assert abs(w[2] - sub[2]) < 1e-14
    assert abs(w[3] / w[0] ** GAMMA - sub[3] / sub[0] ** GAMMA) < 1e-13
    sup = np.array([1.0, 3.0, 0.0, 1.0])
    assert np.allclose(O.far_field_state(free, sup, -1.0, 0.0, GAMMA), sup, rtol=1e-13)


def test_residual_conserves_across_faces():
    """Telescoping: the residuals of all cells sum to the net inflow
    through the boundary faces (plus the axisymmetric source), so every
    interior face is added to one cell and taken from the other."""
    g = cases.distorted_grid(10, 8)
    rng = np.random.default_rng(7)
    w = np.column_stack([rng.uniform(0.8, 1.2, 80), rng.uniform(0.5, 1, 80), rng.uniform(-0.2, 0.2, 80),
                         rng.uniform(0.8, 1.2, 80)])
    bcs = {"iMin": {"kind": "inflow", "state": (1, 0.7, 0, 1)}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "farField", "state": (1, 0.7, 0, 1)}}
    for geometry in ("planar", "axisymmetric"):
        s = O.Euler2D(g, geometry=geometry, limiter="vanLeer", boundaries=bcs)
        s.set_primitives(w)
        total = s.residual_vector().sum(axis=0)
        inflow = np.zeros(4)
        for side in ("iMin", "iMax", "jMin", "jMax"):
            f = s.boundary_fluxes(side)
            inflow -= (f[:, 5:9] * f[:, 4:5]).sum(axis=0)
        if geometry == "axisymmetric":
            prim = s.primitives()
            inflow[2] += (prim[:, 3] * s.metrics["planar"]).sum()
        assert np.allclose(total, inflow, rtol=1e-12, atol=1e-12)


def test_wall_flux_carries_no_mass_and_drag_of_a_synthetic_wall():
    """forces.ts:19-29 on a synthetic wall (tests/analysis/forces.test.ts:73
    expects 0.6 / q): faces with p n_x A summing to 0.3 and p_inf = 0."""
    faces = np.array([[0, 0, 1.0, 0, 0.1, 0, 2.0, 0, 0], [0, 0, 0.5, 0.5, 0.2, 0, 0.5, 0, 0]])
    cd = O.lib().orc_forebody_pressure_drag(O._p(np.ascontiguousarray(faces.ravel())), 2, 0.0, 1.0, 1.0)
    assert abs(cd - 2 * (2.0 * 0.1 + 0.5 * 0.2)) < 1e-15
    body = grids.cone_cylinder(15.0, 1.0)
    s, bg, free, p_inf, q_inf = _verified_case(body, 2.0, nj=12)
    s.iterate(200)
    f = s.boundary_fluxes("jMin")
    assert np.abs(f[:, 5]).max() < 1e-14


def test_taylor_maccoll_against_scipy_and_cone_surface_pressure():
    """The host Taylor-Maccoll solution against an independent scipy
    integration of the same ODE, and the oracle's cone surface pressure on
    the config-1 body (15 deg + 4 calibers, M2) within 1e-3 of it
    (tests/engine-cpu/body-validation.test.ts:72-116)."""
    mach, cone = 2.0, 15 * DEG
    tm = grids.taylor_maccoll(mach, cone, GAMMA)
    g = GAMMA

    def rhs(theta, y):
        vr, vt = y
        a = 0.5 * (g - 1) * (1 - vr * vr - vt * vt)
        return [vt, (vr * vt * vt - a * (2 * vr + vt / math.tan(theta))) / (a - vt * vt)]

    beta = tm.shock_angle
    mn1 = mach * math.sin(beta)
    defl = grids.deflection_angle(mach, beta, g)
    m2 = grids.normal_shock(mn1, g)[2] / math.sin(beta - defl)
    v = 1 / math.sqrt(1 + 2 / ((g - 1) * m2 * m2))
    ev = lambda th, y: y[1]
    ev.terminal = True
    sol = solve_ivp(rhs, [beta, 1e-3], [v * math.cos(beta - defl), -v * math.sin(beta - defl)], events=ev,
                    rtol=1e-12, atol=1e-14)
    assert abs(sol.t_events[0][0] - cone) < 1e-8

    body = grids.cone_cylinder(15.0, 4.0)
    s, bg, free, p_inf, q_inf = _verified_case(body, mach)
    s.iterate(1200)
    f = s.boundary_fluxes("jMin")
    L = body.nose.length
    sel = (f[:, 0] > 0.5 * L) & (f[:, 0] < 0.95 * L)
    p = f[sel, 6] * f[sel, 2] + f[sel, 7] * f[sel, 3]
    assert abs(p.mean() / tm.surface_pressure_ratio - 1) < 1e-3


def test_implicit_relaxations_reach_the_explicit_drag():
    """tests/engine-cpu/implicit.test.ts:201-272: LU-SGS sweeps, Jacobi (4)
    and red-black (2) reach the explicit scheme's drag (within 1e-10 here)."""
    body = grids.cone_cylinder(15.0, 1.0)
    mach = 2.0

    def make():
        bg = grids.body_grid(body, 12, mach, GAMMA)
        free, p_inf, q_inf = grids.free_stream_at(mach, GAMMA)
        bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
               "jMin": {"kind": "wall"}, "jMax": {"kind": "farField", "state": free}}
        s = O.Euler2D(bg.grid, flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled",
                      time_integrator="ssprk2", cfl=0.8, time_stepping="local", boundaries=bcs)
        s.set_primitives(np.array(free))
        return s, p_inf, q_inf

    explicit, p_inf, q_inf = make()
    explicit.iterate(6000)
    assert explicit.density_residual_norm() < 1e-12
    ref = explicit.forebody_drag(p_inf, q_inf, 0.5)
    for relaxation, sweeps, cap in (("sweeps", 1, 600), ("jacobi", 4, 3000), ("redBlack", 2, 3000),
                                    ("redBlackLinear", 2, 3000)):
        s, _, _ = make()
        cfl = 1.0
        for _ in range(cap):
            if s.density_residual_norm() <= 1e-13:
                break
            s.iterate_implicit(1, cfl, 1.5, relaxation, sweeps)
            cfl = min(1000.0, cfl * 1.2)
        assert s.density_residual_norm() < 1e-12, relaxation
        assert abs(s.forebody_drag(p_inf, q_inf, 0.5) / ref - 1) < 1e-10, relaxation


def _normal_flux(q, nx, ny, g):
    """The Euler flux along (nx, ny) of conserved q (textbook form)."""
    rho, mx, my, e = q
    p = (g - 1) * (e - 0.5 * (mx * mx + my * my) / rho)
    un = (mx * nx + my * ny) / rho
    return np.array([rho * un, mx * un + p * nx, my * un + p * ny, (e + p) * un])


def test_linear_neighbour_term_is_the_flux_jacobian():
    """euler2d-two-float.wgsl:574-598: A (1/2 J_n d - 1/2 omega lambda d)
    with J_n the flux Jacobian along n. Pinned against central differences
    of the textbook flux (error O(eps^2)), linearity, and the damping term
    alone when the normal flux does not change."""
    rng = np.random.default_rng(11)
    g = 1.4
    for _ in range(50):
        rho = 0.5 + rng.random()
        u, v = rng.normal(size=2)
        p = 0.5 + rng.random()
        q = np.array([rho, rho * u, rho * v, p / (g - 1) + 0.5 * rho * (u * u + v * v)])
        d = rng.normal(size=4)
        t = rng.random() * 2 * math.pi
        nx, ny = math.cos(t), math.sin(t)
        area, lam, omega = 0.7, 1.3, 1.5
        lin = O.neighbour_term_linear(q, d, g, nx, ny, area, lam, omega)
        eps = 1e-5
        jd = (_normal_flux(q + eps * d, nx, ny, g) - _normal_flux(q - eps * d, nx, ny, g)) / (2 * eps)
        expect = 0.5 * area * jd - 0.5 * omega * lam * d
        assert np.abs(lin - expect).max() < 1e-8 * max(1.0, np.abs(expect).max())
        lin2 = O.neighbour_term_linear(q, 2 * d, g, nx, ny, area, lam, omega)
        assert np.allclose(lin2, 2 * lin, rtol=1e-14, atol=1e-14)
        assert not np.any(O.neighbour_term_linear(q, np.zeros(4), g, nx, ny, area, lam, omega))
    # A change along the state itself scales the flux: J_n q = F_n(q) for
    # the homogeneous Euler flux (degree 1).
    q = np.array([1.2, 0.6, -0.3, 2.9])
    lin = O.neighbour_term_linear(q, q, 1.4, 0.6, 0.8, 1.0, 0.0, 1.0)
    assert np.allclose(lin, 0.5 * _normal_flux(q, 0.6, 0.8, 1.4), rtol=1e-14, atol=1e-14)


def test_taylor_maccoll_transonic_cone_value():
    """tests/analysis/transonic-drag.test.ts:173-176: Taylor-Maccoll's
    surface pressure coefficient of the 6 deg 59' cone at Mach 1.1 (its
    nose drag coefficient, the cone's pressure being uniform) is 0.0944976
    to six decimals."""
    flow = grids.taylor_maccoll(1.1, (6 + 59 / 60) * DEG, GAMMA)
    print("Cp", flow.surface_pressure_coefficient)
    assert abs(flow.surface_pressure_coefficient - 0.0944976) < 5e-7
