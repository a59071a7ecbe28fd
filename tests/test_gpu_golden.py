"""The HIP solver against the golden fixtures of tests/golden (written by
the FP64 oracle; no oracle runs here): the residual vector of a seeded
flow on a distorted 4 x 4 grid with all four boundary kinds, planar and
axisymmetric, and 10 SSP-RK3 steps on a distorted 8 x 8 grid. Bars: the
FP64 solver within 1e-12 of each component's largest magnitude; the FP32
solver within its transient bar (L1 5e-6, Linf 2e-4) and the FP32
residual within 1e-5."""
import os

import numpy as np
import pytest

import cases
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D
from bowshock_amd.grids import StructuredGrid

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GAMMA = 1.4
FREE = (1.0, 2.0 * np.sqrt(GAMMA), 0.0, 1.0)
SIDES = {"iMin": {"kind": "inflow", "state": FREE}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
         "jMax": {"kind": "farField", "state": FREE}}


def _load(name):
    with np.load(os.path.join(HERE, name + ".npz")) as f:
        return {k: f[k] for k in f.files}


def _solver(ctx, d, n, precision, geometry, **scheme):
    grid = StructuredGrid(n, n, d["x"], d["y"])
    cfg = Euler2DConfig(grid=grid, geometry=geometry, gamma=GAMMA, flux="hllc", timeIntegrator="ssprk3", cfl=0.5,
                        timeStepping="local", boundaries=SIDES, **scheme)
    gpu = GpuEuler2D.create(ctx, cfg, precision)
    gpu.setCellPrimitives(d["primitives"])
    return gpu


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("geometry", ["planar", "axisymmetric"])
def test_residual_4x4(ctx, precision, geometry):
    d = _load("residual_4x4")
    gpu = _solver(ctx, d, 4, precision, geometry, limiter="vanAlbada", reconstruction="sharedScaled")
    want = d[f"residual_{geometry}"]
    err = (np.abs(gpu.residualVector() - want).max(axis=0) / np.abs(want).max(axis=0)).max()
    print(precision, geometry, err)
    assert err < (1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_trajectory_8x8(ctx, precision):
    d = _load("trajectory_8x8")
    gpu = _solver(ctx, d, 8, precision, "axisymmetric", limiter="minmod", reconstruction="perVariable")
    gpu.iterate(10)
    s = gpu.readState()
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1).astype(np.float64)
    l1, linf = cases.field_errors(q, d["after_10"])
    print(precision, l1.max(), linf.max())
    if precision == "fp64":
        assert linf.max() < 1e-12
    else:
        assert l1.max() < 5e-6 and linf.max() < 2e-4
