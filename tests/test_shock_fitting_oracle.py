"""The fitted-shock oracle (oracle/shock_fitting.py, restating
src/solver/shock-fitting.ts) pinned by the mathematics: the post-shock
state satisfies the Rankine-Hugoniot conditions in the shock's frame
(mass, normal momentum and energy fluxes continuous, tangential velocity
kept); M = 1 is the undisturbed stream (a Mach wave); a stationary normal
shock (W = 0) is the textbook normal shock; and the compatible shock's
state satisfies the compatibility relation it was solved from."""
import math

import numpy as np
import pytest

from shock_fitting import compatible_shock, shock_state

G = 1.4


def fluxes_in_shock_frame(w, speed, nx, ny, g):
    rho, u, v, p = w
    un = u * nx + v * ny - speed
    ut = -u * ny + v * nx
    h = g / (g - 1) * p / rho + 0.5 * (un * un + ut * ut)
    return np.array([rho * un, rho * un * un + p, rho * un * h]), ut


@pytest.mark.parametrize("mach,angle", [(1.0, 0.3), (1.7, 1.1), (3.2, -0.4), (7.5, 2.5)])
def test_rankine_hugoniot_in_the_shock_frame(mach, angle):
    free = (1.3, 2.1, -0.7, 0.9)
    nx, ny = math.cos(angle), math.sin(angle)
    w, speed, vn = shock_state(free, G, mach, nx, ny)
    a, ta = fluxes_in_shock_frame(free, speed, nx, ny, G)
    b, tb = fluxes_in_shock_frame(w, speed, nx, ny, G)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12)
    assert abs(ta - tb) < 1e-12
    assert abs(vn - (w[1] * nx + w[2] * ny)) < 1e-12
    if mach == 1.0:
        assert np.allclose(w, free, rtol=1e-13, atol=1e-13)


def test_stationary_normal_shock_is_the_textbook_one():
    """Free stream at Mach 2 along +x meeting a shock of normal -x (the
    shock faces the flow): W = 0 at M = 2, and the downstream state is
    the normal-shock table's (p2/p1 = 4.5, rho2/rho1 = 2.6667,
    M2 = 0.57735)."""
    free = (1.0, 2.0 * math.sqrt(G), 0.0, 1.0)
    w, speed, _ = shock_state(free, G, 2.0, -1.0, 0.0)
    assert abs(speed) < 1e-14
    assert w[3] == pytest.approx(4.5, rel=1e-14)
    assert w[0] == pytest.approx(8.0 / 3.0, rel=1e-14)
    m2 = abs(w[1]) / math.sqrt(G * w[3] / w[0])
    assert m2 == pytest.approx(math.sqrt(1.0 / 3.0), rel=1e-13)


def test_compatible_shock_solves_its_relation():
    free = (1.0, 2.5, 0.0, 1.0)
    nx, ny = -0.8, 0.6
    rho, a, p, un = 1.9, 1.4, 2.6, -0.2
    w, speed, vn, mach = compatible_shock(free, G, nx, ny, (rho, a, p, un))
    assert mach > 1
    assert abs(w[3] + rho * a * vn - (p + rho * a * un)) < 1e-12 * (p + abs(rho * a * un))
    # A flow behind that the undisturbed stream already satisfies: a Mach wave.
    w1, _, _, m1 = compatible_shock(free, G, nx, ny, (1.0, math.sqrt(G), 1.0, 2.5 * nx))
    assert m1 == 1.0 and np.allclose(w1, free)


# ---- the moving grid: remapAlongLines and the moves of fitShock -----------

import pyoracle as O  # noqa: E402
from shock_fitting import fitted_moves, grid_on_lines, remap_along_lines  # noqa: E402


def _fan_lines(ni):
    """ni + 1 straight lines from an arc of the unit circle, pointing
    outwards (unit directions), and fixed fractions stretched towards the
    body: a C-grid-like family."""
    th = np.linspace(0.1, 1.4, ni + 1)
    lines = np.stack([np.cos(th), np.sin(th), np.cos(th), np.sin(th)], axis=1)
    return lines


def _fractions(ni, nj):
    s = np.linspace(0.0, 1.0, nj + 1) ** 1.3
    return np.tile(s[:, None], (1, ni + 1))


def _centres_along(lines, grid, ni, nj):
    """Cell centres' distance along their column, as the remap measures it
    (the mean of the four nodes' projections on the column's lines)."""
    X, Y = grid.x.reshape(nj + 1, ni + 1), grid.y.reshape(nj + 1, ni + 1)
    proj = (X - lines[None, :, 0]) * lines[None, :, 2] + (Y - lines[None, :, 1]) * lines[None, :, 3]
    col = 0.5 * (proj[:, :-1] + proj[:, 1:])
    return 0.5 * (col[:-1] + col[1:])


def test_remap_identity_when_the_heights_do_not_change():
    ni, nj = 9, 7
    lines, eta = _fan_lines(ni), _fractions(ni, nj)
    h = 1.0 + 0.1 * np.sin(np.arange(ni + 1))
    g = grid_on_lines(lines, eta, h)
    rng = np.random.default_rng(5)
    prim = rng.uniform(0.5, 2.0, (ni * nj, 4))
    out = remap_along_lines(lines, nj, prim, g, g)
    assert np.allclose(out, prim, rtol=4e-16, atol=0)


@pytest.mark.parametrize("scale", [1.03, 0.97])
def test_remap_exact_on_fields_linear_along_the_lines_and_held_at_the_ends(scale):
    ni, nj = 9, 7
    lines, eta = _fan_lines(ni), _fractions(ni, nj)
    h = 1.0 + 0.1 * np.sin(np.arange(ni + 1))
    old, new = grid_on_lines(lines, eta, h), grid_on_lines(lines, eta, scale * h)
    c_old, c_new = _centres_along(lines, old, ni, nj), _centres_along(lines, new, ni, nj)
    # Per column a different linear function of the distance, per variable.
    a = np.linspace(1.0, 2.0, ni)[None, :, None] + np.arange(4)[None, None, :]
    b = np.linspace(-0.5, 0.7, ni)[None, :, None] * (1 + np.arange(4))[None, None, :]
    prim = (a + b * c_old[:, :, None]).reshape(ni * nj, 4)
    out = remap_along_lines(lines, nj, prim, old, new).reshape(nj, ni, 4)
    want = a + b * c_new[:, :, None]
    lo, hi = c_old[0][None, :], c_old[-1][None, :]
    inside = (c_new >= lo) & (c_new <= hi)
    assert inside.sum() > ni * (nj - 2)
    assert np.allclose(out[inside], want[inside], rtol=1e-13, atol=1e-13)
    # Beyond the old end cells' centres the end cells' values are held.
    prim3 = prim.reshape(nj, ni, 4)
    below, above = c_new < lo, c_new > hi
    assert below.any() or above.any()
    for j, i in zip(*np.nonzero(below)):
        assert np.array_equal(out[j, i], prim3[0, i])
    for j, i in zip(*np.nonzero(above)):
        assert np.allclose(out[j, i], prim3[-1, i], rtol=1e-15, atol=0)


def test_a_fitted_move_by_its_shock_speed_in_closed_form():
    """Vertical lines over a flat wall, the free stream falling on it at
    Mach 3 along -y, and the flow below the shock the Rankine-Hugoniot
    state of a shock of relative Mach number 3.5 (moving upstream): the
    compatible shock recovers that shock, and one move lifts every node
    but the held first one by W dt with dt = half its top cell's height
    over |u| + a below (fitShock, shock-fitting.ts:252-296)."""
    g = 1.4
    ni, nj = 6, 5
    lines = np.array([(0.2 * i, 0.0, 0.0, 1.0) for i in range(ni + 1)])
    eta = _fractions(ni, nj)
    h = np.full(ni + 1, 1.5)
    a_inf = math.sqrt(g)
    free = (1.0, 0.0, -3.0 * a_inf, 1.0)
    w, speed, _ = shock_state(free, g, 3.5, 0.0, 1.0)
    assert speed > 0

    def make(grid, states):
        return O.Euler2D(grid, geometry="planar", gamma=g, flux="hllc", limiter="minmod",
                         reconstruction="perVariable", time_integrator="forwardEuler", cfl=0.4,
                         time_stepping="local",
                         boundaries={"iMin": {"kind": "outflow"}, "iMax": {"kind": "outflow"},
                                     "jMin": {"kind": "outflow"}, "jMax": {"kind": "faceStates",
                                                                          "states": states}})

    h1, _, _ = fitted_moves(make, lines, eta, h, free, g, np.array(w), 1, symmetry=False)
    signal = math.hypot(w[1], w[2]) + math.sqrt(g * w[3] / w[0])
    top = h * (eta[nj] - eta[nj - 1])
    want = h + speed * 0.5 * top / signal
    want[0] = h[0]
    assert np.allclose(h1, want, rtol=1e-12, atol=0)
    # Zero shock speed: the stationary shock's state (Mach 3) leaves the
    # heights where they are.
    w0, s0, _ = shock_state(free, g, 3.0, 0.0, 1.0)
    assert abs(s0) < 1e-14
    h0, _, _ = fitted_moves(make, lines, eta, h, free, g, np.array(w0), 1, symmetry=False)
    assert np.allclose(h0, h, rtol=1e-13, atol=0)
