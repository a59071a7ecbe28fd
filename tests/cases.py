"""Shared case builders for the parity tests: grids, free streams, boundary
sets and the matching oracle / GPU engines."""
from __future__ import annotations

import math

import numpy as np

from bowshock_amd import grids
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D

GAMMA = 1.4


def lcg(seed):
    """Deterministic numbers in [-0.5, 0.5): the generator of
    tests/engine-cpu/euler2d.test.ts:35-39 (a = 1664525, c = 1013904223,
    m = 2^32)."""
    state = seed

    def nxt():
        nonlocal state
        state = (state * 1664525 + 1013904223) % 2 ** 32
        return state / 2 ** 32 - 0.5

    return nxt


def distorted_grid(ni, nj, x0=0.0, x1=2.0, y0=0.5, y1=1.5, seed=2024):
    """Interior nodes moved randomly by up to 30% of the spacing, in the
    order of tests/engine-cpu/euler2d.test.ts:33-46 (i fastest, di then dj
    per interior node)."""
    rnd = lcg(seed)

    def node(i, j):
        interior = 0 < i < ni and 0 < j < nj
        di = 0.6 * rnd() if interior else 0.0
        dj = 0.6 * rnd() if interior else 0.0
        return x0 + ((x1 - x0) * (i + di)) / ni, y0 + ((y1 - y0) * (j + dj)) / nj

    return grids.grid_from_function(ni, nj, node)


def cone_case(ni, nj, mach=2.0, half_angle=15.0, cylinder=4.0):
    body = grids.cone_cylinder(half_angle, cylinder)
    bg = grids.body_grid_fixed(body, ni, nj, mach, GAMMA)
    free, p_inf, q_inf = grids.free_stream_at(mach, GAMMA)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
    return bg.grid, free, p_inf, q_inf, bcs


def ogive_case(ni, nj, mach=3.0, fineness=3.0, cylinder=4.0):
    body = grids.ogive_cylinder(fineness, cylinder)
    bg = grids.body_grid_fixed(body, ni, nj, mach, GAMMA)
    free, p_inf, q_inf = grids.free_stream_at(mach, GAMMA)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
    return bg.grid, free, p_inf, q_inf, bcs


def make_pair(ctx, grid, bcs, *, geometry="axisymmetric", flux="hllc", limiter="minmod",
              reconstruction="perVariable", integrator="ssprk3", cfl=0.5, stepping="local"):
    """The same case on the oracle and on the GPU."""
    import pyoracle as O

    orc = O.Euler2D(grid, geometry=geometry, gamma=GAMMA, flux=flux, limiter=limiter,
                    reconstruction=reconstruction, time_integrator=integrator, cfl=cfl,
                    time_stepping=stepping, boundaries=bcs)
    cfg = Euler2DConfig(grid=grid, geometry=geometry, gamma=GAMMA, flux=flux, limiter=limiter,
                        reconstruction=reconstruction, timeIntegrator=integrator, cfl=cfl,
                        timeStepping=stepping, boundaries=bcs)
    gpu = GpuEuler2D.create(ctx, cfg) if ctx is not None else None
    return orc, gpu


def gpu_conserved(gpu):
    s = gpu.readState()
    return np.stack([s.rho, s.mu, s.mv, s.energy], axis=1).astype(np.float64)


def field_errors(a, b):
    """Relative L1 and Linf differences per conserved component, each
    scaled by the component's largest magnitude in b."""
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    d = np.abs(a - b) / scale
    return d.mean(axis=0), d.max(axis=0)


def state_from_gpu(orc, gpu, gamma=GAMMA):
    """Load the GPU's FP32 state into the oracle in FP64 with the FP32
    gamma's pressure (gpuPrimitive2D, euler2d-gpu.ts:136-146)."""
    from bowshock_amd.engine import gpu_primitives

    orc.set_primitives(gpu_primitives(gpu.readState(), gamma))
