"""The precision check of the transonic contract (bs_check_precision;
TransonicGridSolver.checkPrecision, transonic-drag.ts:56-66;
gpu-transonic.ts:108-131 residualPrecision with RESIDUAL_PRECISION = 1e-8).

An FP64 solver's residual is the double one, and a healthy state passes.
An FP32 solver's residual differs from the double one of the same state by
far more than 1e-8 of the largest residual (32-bit face fluxes), so the
check raises GpuPrecisionError. The difference it reports is the one the
oracle gives: the FP32 solver's residual vector against the oracle's FP64
residual vector of the same conserved state (widened), on a plain case and
on one with face states and an imposed flux."""
import ctypes as C

import numpy as np
import pytest

import cases
from bowshock_amd import _native as N
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D, GpuPrecisionError

pytestmark = pytest.mark.gpu


def _scheme():
    return dict(geometry="axisymmetric", gamma=1.4, flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled",
                timeIntegrator="ssprk3", cfl=0.8, timeStepping="local")


def test_an_fp64_solver_passes_the_check_on_a_healthy_state(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=3.0)
    gpu = GpuEuler2D.create(ctx, Euler2DConfig(grid=grid, boundaries=bcs, **_scheme()), "fp64")
    gpu.setCellPrimitives(np.array(free))
    gpu.setImplicitCfl(1.0)
    gpu.iterateImplicit(40, cfl=10.0, sweeps=2)
    assert gpu.checkPrecision() == 0.0
    gpu.destroy()


def _difference(gpu):
    """bs_check_precision's status and reported difference."""
    d = C.c_double(-1.0)
    rc = N.lib().bs_check_precision(gpu._h, C.byref(d))
    return rc, d.value


def _oracle_difference(orc, gpu):
    q = cases.gpu_conserved(gpu)
    orc.set_conserved(q)
    ref = orc.residual_vector()
    mine = gpu.residualVector()
    return np.abs(mine - ref).max() / np.abs(ref).max()


def test_an_fp32_solver_fails_the_check_by_the_oracle_s_difference(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=3.0)
    orc, gpu = cases.make_pair(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled", cfl=0.8)
    gpu.setCellPrimitives(np.array(free))
    gpu.iterate(60)
    with pytest.raises(GpuPrecisionError, match="differs from the double-precision one"):
        gpu.checkPrecision()
    rc, diff = _difference(gpu)
    assert rc == N.BS_ERR_PRECISION
    want = _oracle_difference(orc, gpu)
    assert 1e-8 < want < 1e-2
    assert diff == pytest.approx(want, rel=1e-3)
    gpu.destroy()


def test_the_check_carries_face_states_and_imposed_fluxes(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(40, 20, mach=2.5)
    states = np.tile(np.array(free), (20, 1))
    states[:, 3] *= 1.0 + 0.01 * np.arange(20) / 20
    bcs = {"iMin": {"kind": "faceStates", "states": states}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "farField", "state": free}}
    orc, gpu = cases.make_pair(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    rho, u, v, p = free
    e = p / 0.4 + 0.5 * rho * u * u
    flux = np.tile(np.array([rho * u, rho * u * u + p, 0.0, u * (e + p)]), (20, 1)) * 1.001
    orc.set_face_fluxes("iMin", flux)
    gpu.setFaceFluxes("iMin", flux)
    gpu.setCellPrimitives(np.array(free))
    gpu.iterate(30)
    rc, diff = _difference(gpu)
    assert rc == N.BS_ERR_PRECISION
    # The GPU holds the face states as FP32 conserved values formed with the
    # FP32 gamma (euler2d-gpu.ts:412-432); the oracle takes the primitives
    # of those values, widened, with the FP64 gamma, as the check does.
    g32 = float(np.float32(1.4))
    r, uu, vv, pp = states.T
    q32 = np.stack([r, r * uu, r * vv, pp / (g32 - 1) + 0.5 * r * (uu * uu + vv * vv)], axis=1)
    q32 = q32.astype(np.float32).astype(np.float64)
    faces = np.stack([q32[:, 0], q32[:, 1] / q32[:, 0], q32[:, 2] / q32[:, 0],
                      (1.4 - 1) * (q32[:, 3] - 0.5 * (q32[:, 1] ** 2 + q32[:, 2] ** 2) / q32[:, 0])], axis=1)
    orc.set_face_states("iMin", faces)
    orc.set_face_fluxes("iMin", flux.astype(np.float32).astype(np.float64))
    want = _oracle_difference(orc, gpu)
    assert diff == pytest.approx(want, rel=1e-3)
    gpu.destroy()
