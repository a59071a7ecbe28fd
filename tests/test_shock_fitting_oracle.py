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
