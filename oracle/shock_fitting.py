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


def grid_on_lines(lines, eta, heights):
    """Nodes at each line's origin plus h_i eta[j, i] times its direction
    (gridOnLines, blunt-grid.ts:602-620, with fixed fractions)."""
    from bowshock_amd.grids import StructuredGrid

    lines = np.asarray(lines, dtype=np.float64)
    eta = np.asarray(eta, dtype=np.float64)
    nj, ni = eta.shape[0] - 1, eta.shape[1] - 1
    s = eta * np.asarray(heights, dtype=np.float64)[None, :]
    x = lines[None, :, 0] + s * lines[None, :, 2]
    y = lines[None, :, 1] + s * lines[None, :, 3]
    return StructuredGrid(ni, nj, x.reshape(-1), y.reshape(-1))


def remap_along_lines(lines, nj, old_prim, old_grid, new_grid):
    """shock-fitting.ts:305-338 remapAlongLines: the old primitives (n, 4)
    interpolated linearly between old cell centres along each column to
    the new cells' centres, held at the end cells."""
    ni = old_grid.ni

    def along(g, i, j):
        k = i + j * (ni + 1)
        l = lines[i]
        return (g.x[k] - l[0]) * l[2] + (g.y[k] - l[1]) * l[3]

    def column(g, i, j):
        return 0.5 * (along(g, i, j) + along(g, i + 1, j))

    out = np.zeros_like(old_prim)
    for j in range(nj):
        for i in range(ni):
            d = 0.5 * (column(new_grid, i, j) + column(new_grid, i, j + 1))

            def centre(jj):
                return 0.5 * (column(old_grid, i, jj) + column(old_grid, i, jj + 1))

            jj = 0
            while jj < nj - 2 and centre(jj + 1) < d:
                jj += 1
            t = min(max((d - centre(jj)) / (centre(jj + 1) - centre(jj)), 0.0), 1.0)
            a = old_prim[i + jj * ni]
            b = old_prim[i + (jj + 1) * ni]
            out[i + j * ni] = a + t * (b - a)
    return out


def fitted_moves(make, lines, eta, heights, free, gamma, initial, moves: int, symmetry: bool):
    """fitShock (shock-fitting.ts:215-296) with one step per move: the
    boundary, one step, each node moved by the speed of the shock face on
    its axis side times half its top cell's height per signal speed, the
    first node normal to the axis (symmetry) or held, the flow remapped
    onto the new grid. The signal speed is the one the boundary recorded
    at the start of the step, as the GPU engine's move takes it
    (line-grid.wgsl node_move); fitShock takes the cell's after the step.
    `make(grid, states)` builds the oracle solver for a grid with these
    jMax face states. Returns (heights, oracle, grid)."""
    lines = np.asarray(lines, dtype=np.float64)
    eta = np.asarray(eta, dtype=np.float64)
    nj, ni = eta.shape[0] - 1, eta.shape[1] - 1
    h = np.array(heights, dtype=np.float64)
    grid = grid_on_lines(lines, eta, h)
    orc = make(grid, np.tile(np.asarray(free, dtype=np.float64), (ni, 1)))
    orc.set_primitives(initial)
    for _ in range(moves):
        states, fluxes, speeds, machs, signals = shock_boundary(orc, grid, free, gamma)
        orc.set_face_states("jMax", states)
        orc.set_face_fluxes("jMax", fluxes)
        orc.iterate(1)
        nxt = h.copy()
        m = orc.metrics
        for k in range(1, ni + 1):
            f = k - 1
            nx, ny = m["jnx"][f + nj * ni], m["jny"][f + nj * ni]
            l = lines[k]
            size = h[k] * (eta[nj, k] - eta[nj - 1, k]) * math.hypot(l[2], l[3])
            dt = 0.5 * size / signals[min(k, ni - 1)]
            nxt[k] = h[k] + speeds[f] * dt / (l[2] * nx + l[3] * ny)
        if symmetry:
            def node(k):
                s = nxt[k] * eta[nj, k]
                return lines[k, 0] + s * lines[k, 2], lines[k, 1] + s * lines[k, 3]

            x1, y1 = node(1)
            x2, y2 = node(2)
            c = (x2 - x1) / (y2 * y2 - y1 * y1)
            nxt[0] = (x1 - c * y1 * y1 - lines[0, 0]) / (lines[0, 2] * eta[nj, 0])
        new_grid = grid_on_lines(lines, eta, nxt)
        prim = remap_along_lines(lines, nj, orc.primitives(), grid, new_grid)
        orc = make(new_grid, states)
        orc.set_primitives(prim)
        grid, h = new_grid, nxt
    return h, orc, grid
