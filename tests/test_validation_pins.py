"""Pins of the FP64 oracle against exact solutions, the reference's own
validation cases (tests/engine-cpu/euler2d.test.ts:73-205,
tests/engine-cpu/euler2d-validation.test.ts:84-167) with their bars:

- the isentropic vortex (Shu, Lecture Notes in Mathematics 1697, 1998) on
  a sinusoidally distorted periodic grid: second order with unlimited
  slopes (1.9 to 2.2 between 32 and 64 cells), above 1.9 with van Albada
  and the shared reconstructions, limiting costing under 1.6 times the
  unlimited error;
- periodic boundaries: the vortex centred on the domain's corner equals
  the centred one shifted by half the domain, to 1e-12;
- an oblique shock on a 10 deg wedge at Mach 2.5 and a Prandtl-Meyer fan
  round a 10 deg corner at Mach 2 (60 x 30 cells): residual down 1e-10,
  wall pressure within 5e-4 of the exact ratio with a spread under 1e-3,
  the shock angle within 0.3 deg of the theta-beta-Mach relation, the
  downstream Mach number within 0.5 % away from the wall.

- a manufactured solution (manufactured.test.ts, with this suite's own
  smooth field): cell sources second-order close to the continuum source,
  and the solver's L1 error falling at order above 1.9 from 16 to 64
  cells, planar and axisymmetric.

The exact solutions are textbook gas dynamics (NACA Report 1135), solved
here with scipy's brentq.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import pyoracle as O
from bowshock_amd import grids

GAMMA = 1.4
DEG = math.pi / 180


def vortex_state(x, y, t, x0=5.0, y0=5.0, u0=1.0, v0=1.0, beta=5.0, period=10.0, g=GAMMA):
    """The isentropic vortex: T = 1 - (g - 1) b^2 / (8 g pi^2) e^(1 - r^2),
    rho = T^(1 / (g - 1)), p = rho^g, swirl b / (2 pi) e^((1 - r^2) / 2)."""
    dx = x - x0 - u0 * t
    dy = y - y0 - v0 * t
    dx = dx - period * np.floor(dx / period + 0.5)
    dy = dy - period * np.floor(dy / period + 0.5)
    r2 = dx * dx + dy * dy
    temp = 1 - ((g - 1) * beta * beta) / (8 * g * math.pi ** 2) * np.exp(1 - r2)
    rho = temp ** (1 / (g - 1))
    swirl = beta / (2 * math.pi) * np.exp(0.5 * (1 - r2))
    return np.stack([rho, u0 - swirl * dy, v0 + swirl * dx, rho ** g], axis=-1)


def wavy_grid(n, L=10.0):
    def node(i, j):
        xi, eta = L * i / n, L * j / n
        bump = 0.4 * math.sin(2 * math.pi * xi / L) * math.sin(2 * math.pi * eta / L)
        return xi + bump, eta + bump

    return grids.grid_from_function(n, n, node)


PERIODIC = {s: {"kind": "periodic"} for s in ("iMin", "iMax", "jMin", "jMax")}


def vortex_error(n, limiter, reconstruction, t_end=1.0):
    g = wavy_grid(n)
    s = O.Euler2D(g, geometry="planar", flux="hllc", limiter=limiter, reconstruction=reconstruction,
                  time_integrator="ssprk3", cfl=0.4, time_stepping="global", boundaries=PERIODIC)
    m = s.metrics
    s.set_primitives(vortex_state(m["cx"], m["cy"], 0.0))
    s.advance_to(t_end)
    exact = vortex_state(m["cx"], m["cy"], t_end)
    return float(np.sum(np.abs(s.primitives()[:, 0] - exact[:, 0]) * m["volume"]) / np.sum(m["volume"]))


def test_vortex_second_order_unlimited():
    order = math.log2(vortex_error(32, "unlimitedCentral", "perVariable")
                      / vortex_error(64, "unlimitedCentral", "perVariable"))
    assert 1.9 < order < 2.2, order


@pytest.mark.parametrize("reconstruction", ["shared", "sharedScaled"])
def test_vortex_second_order_shared(reconstruction):
    coarse = vortex_error(32, "vanAlbada", reconstruction)
    fine = vortex_error(64, "vanAlbada", reconstruction)
    assert math.log2(coarse / fine) > 1.9
    assert fine < 1.6 * vortex_error(64, "unlimitedCentral", "perVariable")


def test_periodic_translation_invariance():
    """euler2d.test.ts:176-205: the vortex centred mid-domain and on the
    corner, advanced to t = 1 on a 40 x 40 uniform grid (spacing exact in
    binary), solver defaults vanLeer per-variable SSP-RK2, CFL 0.5, global."""
    n, L = 40, 10.0

    def run(x0, y0):
        g = grids.rectangular_grid(n, n, 0.0, L, 0.0, L)
        s = O.Euler2D(g, geometry="planar", flux="hllc", limiter="vanLeer", reconstruction="perVariable",
                      time_integrator="ssprk2", cfl=0.5, time_stepping="global", boundaries=PERIODIC)
        m = s.metrics
        s.set_primitives(vortex_state(m["cx"], m["cy"], 0.0, x0=x0, y0=y0, v0=0.5))
        s.advance_to(1.0)
        return s.primitives().reshape(n, n, 4)

    a = run(5.0, 5.0)
    b = run(0.0, 0.0)
    shifted = np.roll(np.roll(b, -n // 2, axis=0), -n // 2, axis=1)
    assert np.abs(a[..., :3] - shifted[..., :3]).max() < 1e-12


def ramp_grid(ni, nj, x_min, x_corner, x_max, height, angle):
    """A channel whose floor turns by `angle` at x_corner, the corner on a
    grid line (src/grid/shapes.ts:30-48 rampGrid)."""
    i_corner = int(math.floor(ni * (x_corner - x_min) / (x_max - x_min) + 0.5))

    def x_at(i):
        if i <= i_corner:
            return x_min + (x_corner - x_min) * i / i_corner
        return x_corner + (x_max - x_corner) * (i - i_corner) / (ni - i_corner)

    def node(i, j):
        x = x_at(i)
        y0 = 0.0 if x <= x_corner else (x - x_corner) * math.tan(angle)
        return x, y0 + (height - y0) * j / nj

    return grids.grid_from_function(ni, nj, node)


def ramp_flow(mach, angle, iterations):
    free = (1.0, mach * math.sqrt(GAMMA), 0.0, 1.0)
    g = ramp_grid(60, 30, -0.5, 0.0, 1.5, 1.2, angle)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
           "jMax": {"kind": "outflow"}}
    s = O.Euler2D(g, geometry="planar", flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled",
                  time_integrator="ssprk2", cfl=0.8, time_stepping="local", boundaries=bcs)
    s.set_primitives(np.array(free))
    r0 = s.density_residual_norm()
    s.iterate(iterations)
    return s, r0


def wall_pressure(s, x0, x1):
    f = s.boundary_fluxes("jMin")
    cx = s.metrics["cx"][:s.ni]
    keep = (cx > x0) & (cx < x1)
    p = f[keep, 6] * f[keep, 2] + f[keep, 7] * f[keep, 3]
    mean = p.mean()
    return mean, np.abs(p / mean - 1).max()


def row_mach(s, j, x0, x1):
    prim = s.primitives().reshape(s.nj, s.ni, 4)[j]
    cx = s.metrics["cx"].reshape(s.nj, s.ni)[j]
    keep = (cx > x0) & (cx < x1)
    w = prim[keep]
    return float(np.mean(np.hypot(w[:, 1], w[:, 2]) / np.sqrt(GAMMA * w[:, 3] / w[:, 0])))


def oblique_shock(m1, theta, g=GAMMA):
    """Weak-shock angle from tan(theta) = 2 cot(b) (M^2 sin^2 b - 1) /
    (M^2 (g + cos 2b) + 2); pressure ratio and downstream Mach number."""
    def f(b):
        return 2 / math.tan(b) * (m1 * m1 * math.sin(b) ** 2 - 1) / (m1 * m1 * (g + math.cos(2 * b)) + 2) \
            - math.tan(theta)

    mu = math.asin(1 / m1)
    b_max = mu + 1e-9
    best = -1.0
    for k in range(1, 2000):   # the deflection's maximum over beta
        b = mu + (math.pi / 2 - mu) * k / 2000
        if f(b) > best:
            best, b_max = f(b), b
    beta = brentq(f, mu + 1e-12, b_max)
    mn1 = m1 * math.sin(beta)
    p_ratio = 1 + 2 * g / (g + 1) * (mn1 * mn1 - 1)
    mn2 = math.sqrt((1 + 0.5 * (g - 1) * mn1 * mn1) / (g * mn1 * mn1 - 0.5 * (g - 1)))
    return beta, p_ratio, mn2 / math.sin(beta - theta)


def prandtl_meyer(m, g=GAMMA):
    a = math.sqrt((g + 1) / (g - 1))
    return a * math.atan(math.sqrt(m * m - 1) / a) - math.atan(math.sqrt(m * m - 1))


def stagnation_ratio(m, g=GAMMA):
    """p0 / p."""
    return (1 + 0.5 * (g - 1) * m * m) ** (g / (g - 1))


def test_oblique_shock_on_a_wedge():
    beta, p_ratio, m2 = oblique_shock(2.5, 10 * DEG)
    assert abs(beta / DEG - 31.85) < 0.05          # NACA 1135 chart value
    s, r0 = ramp_flow(2.5, 10 * DEG, 1600)
    assert s.density_residual_norm() < 1e-10 * r0
    mean, spread = wall_pressure(s, 0.5, 1.2)
    assert abs(mean / p_ratio - 1) < 5e-4 and spread < 1e-3
    p_mid = 0.5 * (1 + p_ratio)
    prim = s.primitives().reshape(s.nj, s.ni, 4)
    cx = s.metrics["cx"].reshape(s.nj, s.ni)
    cy = s.metrics["cy"].reshape(s.nj, s.ni)
    angles = []
    for i in range(s.ni):
        if not 0.5 <= cx[0, i] <= 1.3:
            continue
        for j in range(s.nj - 1):
            p0, p1 = prim[j, i, 3], prim[j + 1, i, 3]
            if p0 > p_mid >= p1:
                y = cy[j, i] + (cy[j + 1, i] - cy[j, i]) * (p0 - p_mid) / (p0 - p1)
                angles.append(math.atan2(y, cx[j, i]))
                break
    assert len(angles) > 10
    assert max(abs(a - beta) for a in angles) / DEG < 0.3
    for j in (3, 4, 5):
        assert abs(row_mach(s, j, 0.6, 1.1) / m2 - 1) < 5e-3


def test_prandtl_meyer_expansion():
    nu2 = prandtl_meyer(2.0) + 10 * DEG
    m2 = brentq(lambda m: prandtl_meyer(m) - nu2, 1.0 + 1e-12, 10.0)
    p_ratio = stagnation_ratio(2.0) / stagnation_ratio(m2)
    s, r0 = ramp_flow(2.0, -10 * DEG, 500)
    assert s.density_residual_norm() < 1e-10 * r0
    mean, spread = wall_pressure(s, 0.6, 1.3)
    assert abs(mean / p_ratio - 1) < 5e-4 and spread < 1e-3
    for j in (3, 4):
        assert abs(row_mach(s, j, 1.0, 1.4) / m2 - 1) < 5e-3


# --------------------------------------------------------------------------
# Manufactured solution (the method of tests/engine-cpu/manufactured.test.ts:
# a smooth supersonic flow between walls, in a channel or an annulus, on a
# smoothly distorted grid, made exact by cell sources; second order pins the
# fluxes, reconstruction, metrics, axisymmetric source and walls together).
# The field below is this suite's own: supersonic in x, v odd and the other
# variables even about both walls (y = 1 and y = 2), as mirror walls need.
# --------------------------------------------------------------------------

def ms_state(x, y, g=GAMMA):
    s = y - 1.0
    ramp = np.sin(0.5 * np.pi * np.maximum(x, 0.0)) ** 4
    rho = 1 + 0.12 * ramp * np.sin(np.pi * x + 0.4) * np.cos(2 * np.pi * s)
    u = 2.4 + 0.15 * ramp * np.cos(2 * np.pi * x) * np.cos(np.pi * s)
    v = 0.08 * ramp * np.sin(2 * np.pi * x + 0.7) * np.sin(2 * np.pi * s)
    p = (1 + 0.1 * ramp * np.cos(np.pi * x + 1.1) * np.cos(np.pi * s)) / g
    return rho, u, v, p


def ms_fluxes(x, y, g=GAMMA):
    rho, u, v, p = ms_state(x, y, g)
    e = p / (g - 1) + 0.5 * rho * (u * u + v * v)
    fx = np.stack([rho * u, rho * u * u + p, rho * u * v, u * (e + p)])
    fy = np.stack([rho * v, rho * u * v, rho * v * v + p, v * (e + p)])
    return fx, fy


GAUSS2 = ((0.5 - 0.5 / math.sqrt(3), 0.5), (0.5 + 0.5 / math.sqrt(3), 0.5))


def ms_cell_sources(grid, axisymmetric, g=GAMMA):
    """The source that makes the field exact for the finite-volume balance:
    per cell, the outward face flux integrated by 2-point Gauss on each
    edge (times the radius in the annulus), less the axisymmetric pressure
    term integrated by 2 x 2 Gauss over the cell's bilinear map."""
    ni, nj = grid.ni, grid.nj
    X = grid.x.reshape(nj + 1, ni + 1)
    Y = grid.y.reshape(nj + 1, ni + 1)
    c = [(X[:-1, :-1], Y[:-1, :-1]), (X[:-1, 1:], Y[:-1, 1:]), (X[1:, 1:], Y[1:, 1:]), (X[1:, :-1], Y[1:, :-1])]
    total = np.zeros((4, nj, ni))
    for e in range(4):
        (ax, ay), (bx, by) = c[e], c[(e + 1) % 4]
        for t, w in GAUSS2:
            px, py = ax + t * (bx - ax), ay + t * (by - ay)
            fx, fy = ms_fluxes(px, py, g)
            r = py if axisymmetric else 1.0
            total += w * r * (fx * (by - ay) - fy * (bx - ax))
    if axisymmetric:
        (x0, y0), (x1, y1), (x2, y2), (x3, y3) = c
        for s, ws in GAUSS2:
            for t, wt in GAUSS2:
                px = (1 - s) * (1 - t) * x0 + s * (1 - t) * x1 + s * t * x2 + (1 - s) * t * x3
                py = (1 - s) * (1 - t) * y0 + s * (1 - t) * y1 + s * t * y2 + (1 - s) * t * y3
                jac = (((1 - t) * (x1 - x0) + t * (x2 - x3)) * ((1 - s) * (y3 - y0) + s * (y2 - y1))
                       - ((1 - s) * (x3 - x0) + s * (x2 - x1)) * ((1 - t) * (y1 - y0) + t * (y2 - y3)))
                total[2] -= ws * wt * jac * ms_state(px, py, g)[3]
    return np.moveaxis(total, 0, -1).reshape(-1, 4)


def ms_channel(n):
    def node(i, j):
        x, y = i / n, 1 + j / n
        bump = 0.04 * math.sin(math.pi * x) * math.sin(math.pi * (y - 1))
        return x + bump, y + bump

    return grids.grid_from_function(n, n, node)


def ms_solve(n, geometry, reconstruction):
    g = ms_channel(n)
    axi = geometry == "axisymmetric"
    inflow = np.array(ms_state(0.0, 1.0))
    bcs = {"iMin": {"kind": "inflow", "state": tuple(inflow)}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "wall"}}
    s = O.Euler2D(g, geometry=geometry, flux="hllc", limiter="vanAlbada" if reconstruction == "shared" else "vanLeer",
                  reconstruction=reconstruction, time_integrator="ssprk2", cfl=0.8, time_stepping="local",
                  boundaries=bcs, source=ms_cell_sources(g, axi))
    m = s.metrics
    s.set_primitives(np.stack(ms_state(m["cx"], m["cy"]), axis=1))
    r0 = s.density_residual_norm()
    s.iterate(40 * n)
    exact = np.stack(ms_state(m["cx"], m["cy"]), axis=1)
    err = float(np.sum(np.abs(s.primitives() - exact).sum(axis=1) * m["volume"]) / np.sum(m["volume"]))
    return err, s.density_residual_norm() / r0


@pytest.mark.parametrize("geometry,reconstruction", [("axisymmetric", "shared"), ("planar", "perVariable")])
def test_manufactured_solution_second_order(geometry, reconstruction):
    """manufactured.test.ts:105-119: residuals converged below 1e-10 of
    their start, L1 error orders above 1.9 from 16 to 32 and 32 to 64
    cells."""
    runs = [ms_solve(n, geometry, reconstruction) for n in (16, 32, 64)]
    for err, res in runs:
        assert res < 1e-10
    o1 = math.log2(runs[0][0] / runs[1][0])
    o2 = math.log2(runs[1][0] / runs[2][0])
    print(geometry, reconstruction, "orders", o1, o2)
    assert o1 > 1.9 and o2 > 1.9


def ms_point_source(x, y, axisymmetric, g=GAMMA, h=1e-3):
    """The continuum source at a point: div F (the radius-weighted
    divergence in the annulus, less p / r in the radial momentum), by
    fourth-order central differences of the fluxes."""
    def d(f):
        return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

    dfx = d(lambda t: ms_fluxes(x + t, y, g)[0])
    if axisymmetric:
        dfy = d(lambda t: (y + t) * ms_fluxes(x, y + t, g)[1]) / y
    else:
        dfy = d(lambda t: ms_fluxes(x, y + t, g)[1])
    src = dfx + dfy
    if axisymmetric:
        src[2] -= ms_state(x, y, g)[3] / y
    return src.T


@pytest.mark.parametrize("geometry", ["planar", "axisymmetric"])
def test_manufactured_cell_sources_integrate_the_point_source(geometry):
    """manufactured.test.ts:76-101: the quadrature-integrated cell sources
    per unit volume approach the continuum source at the centroids at
    second order (above 1.8 from 16 to 32 cells)."""
    axi = geometry == "axisymmetric"

    def deviation(n):
        g = ms_channel(n)
        m = O.compute_metrics(g, axi)
        cells = ms_cell_sources(g, axi)
        point = ms_point_source(m["cx"], m["cy"], axi)
        return np.abs(cells / m["volume"][:, None] - point).max()

    assert math.log2(deviation(16) / deviation(32)) > 1.8
