"""Anderson acceleration on the GPU (bs_anderson_*, bs_iterate_anderson)
against the oracle's restatement of src/solver/anderson.ts and of the
accelerated loop of solveTransonic (transonic-flow.ts:184-208).

Bars: bs_anderson_next on well-conditioned histories equals the oracle's
modified Gram-Schmidt within 1e-10 (the GPU solves the same least-squares
problem through the Cholesky factor of dF^T dF, in a different summation
order); accelerated solver steps follow the oracle's within the FP32
relaxation's bar (L1 5e-6, Linf 2e-4) and converge to the oracle's drag.
"""
import math

import numpy as np
import pytest

import cases
from anderson import AndersonMixing as OracleMixing
from anderson import accelerated_steps
from bowshock_amd.engine import AndersonMixing, Euler2DConfig, GpuEuler2D, RangeError
from cases import GAMMA

pytestmark = pytest.mark.gpu


def test_next_matches_the_oracle(ctx):
    rng = np.random.default_rng(21)
    n, depth = 4096, 4
    gpu = AndersonMixing(ctx, n, depth)
    orc = OracleMixing(n, depth)
    for k in range(10):
        x = rng.normal(size=n)
        f = rng.normal(size=n) * 0.5 ** k
        a = gpu.next(x, f)
        b = orc.next(x, f)
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max(), k
        cols, used, restarts, gamma = gpu.state()
        assert cols == min(k, depth) and used == cols and restarts == 0
        assert np.allclose(gamma[:cols], orc.gamma, rtol=1e-10, atol=1e-13)
    gpu.reset()
    orc.reset()
    x, f = rng.normal(size=n), rng.normal(size=n)
    assert np.array_equal(gpu.next(x, f), x + f)
    with pytest.raises(RangeError):
        gpu.next(np.zeros(3), np.zeros(3))


def test_finite_termination_on_a_linear_iteration(ctx):
    n = 6
    rng = np.random.default_rng(3)
    a = np.eye(n) + 0.3 * rng.normal(size=(n, n))
    b = rng.normal(size=n)
    exact = np.linalg.solve(a, b)
    mix = AndersonMixing(ctx, n, n + 2)
    x = np.zeros(n)
    for _ in range(n + 2):
        x = mix.next(x, 0.2 * (b - a @ x))
    assert np.abs(x - exact).max() < 1e-9


def test_a_repeated_column_is_dropped(ctx):
    rng = np.random.default_rng(9)
    n = 30
    gpu = AndersonMixing(ctx, n, 4)
    orc = OracleMixing(n, 4)
    base, step = rng.normal(size=n), rng.normal(size=n)
    fs = [base, base + step, base + 2 * step, base + 2 * step + rng.normal(size=n)]
    xs = [rng.normal(size=n) for _ in fs]
    for x, f in zip(xs, fs):
        a = gpu.next(x, f)
        b = orc.next(x, f)
    cols, used, _, gamma = gpu.state()
    assert cols == 3 and used == 2 and gamma[0] == 0.0
    assert np.allclose(gamma[:3], orc.gamma, rtol=1e-10, atol=1e-13)
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


def _pair(ctx, grid, bcs, precision="fp64"):
    orc, _ = cases.make_pair(None, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    cfg = Euler2DConfig(grid=grid, geometry="axisymmetric", gamma=GAMMA, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", timeIntegrator="ssprk3", cfl=0.5, timeStepping="local",
                        boundaries=bcs)
    return orc, GpuEuler2D.create(ctx, cfg, precision)


def test_accelerated_steps_follow_the_oracle(ctx):
    """10 plain steps, then 12 accelerated ones (depth 8), both from the
    free stream with the CFL number ramped from 1 to 10."""
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=2.0)
    orc, gpu = _pair(ctx, grid, bcs)
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    cfl = 1.0
    for _ in range(10):
        orc.iterate_implicit(1, cfl, 1.5, "redBlackLinear", 2)
        cfl = min(10.0, cfl * 1.2)
    gpu.setImplicitCfl(1.0)
    gpu.iterateImplicit(10, cfl=10.0, sweeps=2)
    mix = AndersonMixing(ctx, 4 * gpu.n, 8)
    gpu.iterateAnderson(mix, 12, cfl=10.0, sweeps=2)
    accelerated_steps(orc, OracleMixing(4 * gpu.n, 8), 12, cfl, 10.0, 1.5, "redBlackLinear", 2)
    s = gpu.readState()
    q = np.stack([s.rho, s.mu, s.mv, s.energy], axis=1)
    l1, linf = cases.field_errors(q, orc.conserved())
    print("L1", l1, "Linf", linf, "state", mix.state()[:3])
    assert gpu.last_report.steps == 22
    assert l1.max() < 5e-6
    assert linf.max() < 2e-4


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_accelerated_solve_reaches_the_converged_drag(ctx, precision):
    """From the free stream: 30 plain steps, then accelerated ones until the
    density residual is at the precision's floor; the drag is the oracle's
    converged one (1e-10 for the FP64 solver, 1e-5 for FP32)."""
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=2.0)
    orc, gpu = _pair(ctx, grid, bcs, precision)
    orc.set_primitives(np.array(free))
    orc.iterate(2000)
    cd_o = orc.forebody_drag(p_inf, q_inf, 0.5)
    gpu.setCellPrimitives(np.array(free))
    gpu.setImplicitCfl(1.0)
    gpu.iterateImplicit(30, cfl=10.0, sweeps=2)
    mix = AndersonMixing(ctx, 4 * gpu.n, 8)
    target = 1e-11 if precision == "fp64" else 2e-4
    steps, res = 30, math.inf
    while res > target and steps < 3000:
        gpu.iterateAnderson(mix, 50, cfl=10.0, sweeps=2)
        steps += 50
        res = gpu.densityResidualNorm()
    cd = gpu.forebodyDrag(p_inf, q_inf, 0.5)
    print(precision, "steps", steps, "residual", res, "drag", cd / cd_o - 1, "state", mix.state()[:3])
    assert res <= target
    assert abs(cd / cd_o - 1) < (1e-10 if precision == "fp64" else 1e-5)


def test_iterate_anderson_checks_the_history_size(ctx):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(40, 20)
    _, gpu = _pair(ctx, grid, bcs)
    gpu.setCellPrimitives(np.array(free))
    with pytest.raises(RangeError):
        gpu.iterateAnderson(AndersonMixing(ctx, 4 * gpu.n + 1, 4), 1, cfl=10.0, sweeps=2)
    with pytest.raises(RangeError):
        AndersonMixing(ctx, 10, 17)
