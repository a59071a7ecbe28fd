"""The fitted-shock boundary, restated for checking (test infrastructure:
only the tests import it; the product is the HIP library's fitted shock).

Follows src/solver/shock-fitting.ts: shockState (54-82), the
Rankine-Hugoniot state behind a shock of outward normal n met by the free
stream at relative normal Mach number M, moving at W = M a + q_n;
compatibleShock (91-113), the M at which the u_n + a compatibility
relation p_s + rho a u_n,s = p_c + rho a u_n,c holds (doubling, then
bisection to 1e-15), M = 1 where it cannot; and shockBoundary (172-197),
the shock of every jMax face from the two cells below it, its state as the
face's ghost state and its physical flux as the face flux. `fitted_steps`
is fitShock's inner loop without the moves (applyBoundary, then one step).
"""
from __future__ import annotations

import math

import numpy as np

import pyoracle as O


def shock_state(free, gamma, mach, nx, ny):
    """shock-fitting.ts:54-82: (state (rho, u, v, p), speed, normal velocity)."""
    rho, u, v, p = free
    a = math.sqrt(gamma * p / rho)
    vn = u * nx + v * ny
    speed = mach * a + vn
    m2 = mach * mach
    density_ratio = ((gamma + 1) * m2) / ((gamma - 1) * m2 + 2)
    pressure_ratio = 1 + ((2 * gamma) / (gamma + 1)) * (m2 - 1)
    normal_velocity = speed - (mach * a) / density_ratio
    state = (rho * density_ratio, u + (normal_velocity - vn) * nx, v + (normal_velocity - vn) * ny,
             p * pressure_ratio)
    return state, speed, normal_velocity


def compatible_shock(free, gamma, nx, ny, behind):
    """shock-fitting.ts:91-113; behind = (rho, a, p, u_n) at the face.
    Returns (state, speed, normal velocity, mach)."""
    rho, a, p, un = behind
    impedance = rho * a
    target = p + impedance * un

    def mismatch(m):
        st, _, vn = shock_state(free, gamma, m, nx, ny)
        return st[3] + impedance * vn - target

    mach = 1.0
    if mismatch(1.0) < 0:
        low, high = 1.0, 2.0
        while mismatch(high) < 0:
            high *= 2
        k = 0
        while k < 100 and high - low > 1e-15 * high:
            mid = 0.5 * (low + high)
            if mismatch(mid) < 0:
                low = mid
            else:
                high = mid
            k += 1
        mach = 0.5 * (low + high)
    st, speed, vn = shock_state(free, gamma, mach, nx, ny)
    return st, speed, vn, mach


def shock_boundary(orc: O.Euler2D, grid, free, gamma):
    """shock-fitting.ts:172-197: (states (ni, 4), fluxes (ni, 4), speeds,
    machs, signals) of the jMax faces for the oracle's current flow; the
    signal is |u| + a in the face's cell (euler2d.wgsl:449-451)."""
    ni, nj = grid.ni, grid.nj
    m = orc.metrics
    prim = orc.primitives()
    states = np.zeros((ni, 4))
    fluxes = np.zeros((ni, 4))
    speeds, machs, signals = np.zeros(ni), np.zeros(ni), np.zeros(ni)
    for i in range(ni):
        f = i + nj * ni
        nx, ny = m["jnx"][f], m["jny"][f]
        c1 = prim[i + (nj - 1) * ni]
        c2 = prim[i + (nj - 2) * ni]
        a1 = math.sqrt(gamma * c1[3] / c1[0])
        behind = (c1[0], a1, 1.5 * c1[3] - 0.5 * c2[3],
                  1.5 * (c1[1] * nx + c1[2] * ny) - 0.5 * (c2[1] * nx + c2[2] * ny))
        w, speed, _, mach = compatible_shock(free, gamma, nx, ny, behind)
        rho, u, v, p = w
        energy = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
        un = u * nx + v * ny
        states[i] = w
        fluxes[i] = (rho * un, rho * u * un + p * nx, rho * v * un + p * ny, un * (energy + p))
        speeds[i], machs[i], signals[i] = speed, mach, math.hypot(c1[1], c1[2]) + a1
    return states, fluxes, speeds, machs, signals


def fitted_steps(orc: O.Euler2D, grid, free, gamma, count: int):
    """fitShock's applyBoundary + iterate(1), `count` times, with the
    grid held (shock-fitting.ts:242-249). Returns the last boundary."""
    last = None
    for _ in range(count):
        last = shock_boundary(orc, grid, free, gamma)
        orc.set_face_states("jMax", last[0])
        orc.set_face_fluxes("jMax", last[1])
        orc.iterate(1)
    return last
