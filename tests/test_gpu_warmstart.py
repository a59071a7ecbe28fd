"""What the learned warm start gains, end to end on the GPU on a small
scale (the population's warm-start run, bench.py --workload warmstart):
blunt noses (ellipsoids, 1.5 to 3 calibers, 6 to 10 calibers of cylinder,
Mach 2 and 3) on their C grids; 32 solved from the free stream to train
the network on their converged fields, then 8 others solved from the free
stream and from the network's prediction (interpolated on the GPU,
bs_batch_set_coarse), all as FP64 batches solved to a density residual of
1e-6 with the drag settled to 1e-5. Bars: the warm starts need fewer
cell-steps (SAVING), and their drags agree with the free-stream starts'
within 1e-5."""
import math

import numpy as np
import pytest

from bowshock_amd import grids
from bowshock_amd import population as P
from bowshock_amd import warmstart as W
from bowshock_amd.batch import GpuBatch2D, SolvePlan, solve
from bowshock_amd.engine import Euler2DConfig

pytestmark = pytest.mark.gpu


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        f, cyl = 1.5 + 1.5 * rng.random(), 6 + 4 * rng.random()
        out.append(P.Case(P.Rocket(k, "ellipsoid", f, 50.0, cyl), (2.0, 3.0)[k % 2], 160, 80, cyl))
    P.build_grids(out, workers=1)
    assert all(c.grid is not None for c in out)
    return out


def _batch(ctx, cs):
    cfgs, frees, ps, qs = [], [], [], []
    for c in cs:
        free, p, q = grids.free_stream_at(c.mach, 1.4)
        cfgs.append(Euler2DConfig(grid=c.grid, boundaries=P.boundaries(c, free), geometry="axisymmetric",
                                  gamma=1.4, flux="rusanov", limiter="vanAlbada", reconstruction="sharedScaled",
                                  timeIntegrator="ssprk3", cfl=0.8, timeStepping="local"))
        frees.append(free)
        ps.append(p)
        qs.append(q)
    b = GpuBatch2D(ctx, cfgs, precision="fp64")
    b.setUniform(frees)
    return b, np.array(frees), np.array(ps), np.array(qs)


PLAN = SolvePlan(implicit=True, chunk=60, max_steps=12000, cfl=10.0, cfl_start=1.0, sweeps=2, residual=1e-6,
                 drop=0.0, stall=10 ** 9)
SAVING = 0.25  # measured: 38 % fewer cell-steps, drags 1.2e-7 apart


def _features(cs):
    return np.array([np.append(c.features(), c.cylinder) for c in cs])


def test_warm_start_saves_steps_on_blunt_noses(ctx):
    train_cases, held = _cases(1, 32), _cases(2, 8)
    b, frees, ps, qs = _batch(ctx, train_cases)
    res = solve(b, ps, qs, 0.5, PLAN)
    rows, targets = [], []
    for k, (c, r) in enumerate(zip(train_cases, res)):
        if r.converged:
            prim = W.primitives_from_state(b.readState(k), c.ni, c.nj)
            targets.append(W.to_target(prim, float(np.float32(c.mach * math.sqrt(1.4)))))
            rows.append(k)
    b.destroy()
    assert len(rows) >= 28
    ws = W.train(ctx, _features([train_cases[k] for k in rows]), np.array(targets), steps=1500, batch=64,
                 lr=1e-3, seed=0)
    # Held out: from the free stream, then from the prediction.
    b, frees, ps, qs = _batch(ctx, held)
    free_norm = b.densityResidualNorms()
    plain = solve(b, ps, qs, 0.5, PLAN, reference=free_norm)
    b.destroy()
    b, frees, ps, qs = _batch(ctx, held)
    pred = ws.predict_targets(_features(held))
    u = np.array([float(np.float32(c.mach * math.sqrt(1.4))) for c in held])
    b.setCoarse(pred, np.stack([np.ones_like(u), u, u, u * u], axis=1), shape=(W.COARSE_NJ, W.COARSE_NI))
    used = b.densityResidualNorms() <= free_norm
    for k in np.nonzero(~used)[0]:
        b.setCellPrimitives(k, np.broadcast_to(frees[k], (held[k].cells, 4)))
    warm = solve(b, ps, qs, 0.5, PLAN, reference=free_norm, cfl_start=np.where(used, PLAN.cfl, PLAN.cfl_start))
    b.destroy()
    cells = np.array([c.cells for c in held])
    w_plain = float((cells * np.array([r.steps for r in plain])).sum())
    w_warm = float((cells * np.array([r.steps for r in warm])).sum())
    dcd = max(abs(a.cd / p.cd - 1) for a, p in zip(warm, plain))
    print("used", int(used.sum()), "cell-steps plain", w_plain, "warm", w_warm, "ratio", w_warm / w_plain,
          "max CD change", dcd)
    assert all(r.converged and r.residual <= 1e-6 for r in plain + warm)
    assert w_warm < (1 - SAVING) * w_plain
    assert dcd < 1e-5


ROCKETS = 160


def test_warm_start_on_a_population_sample(ctx):
    """The same measure on the rocket population's own blunt noses (the
    seeded population of 160 rockets at Mach 1.5, 2, 3, 4: its accepted C
    grids, up to 936 x 160): trained on the converged FP64 fields of the
    rockets with index % 4 != 3, the held-out quarter solved from the free
    stream and from the warm start, all to a density residual of 1e-6."""
    pop = P.population_cases(ROCKETS)
    blunt = [c for c in pop if c.rocket.kind in P.BLUNT_KINDS]
    P.build_grids(blunt, workers=8)
    acc = [c for c in blunt if c.grid is not None]
    train_cases = [c for c in acc if c.rocket.index % 4 != 3]
    held = [c for c in acc if c.rocket.index % 4 == 3]
    assert len(train_cases) >= 20 and len(held) >= 6
    b, frees, ps, qs = _batch(ctx, train_cases)
    res = solve(b, ps, qs, 0.5, PLAN)
    rows, targets = [], []
    for k, (c, r) in enumerate(zip(train_cases, res)):
        if r.converged:
            prim = W.primitives_from_state(b.readState(k), c.ni, c.nj)
            targets.append(W.to_target(prim, float(np.float32(c.mach * math.sqrt(1.4)))))
            rows.append(k)
    b.destroy()
    assert len(rows) >= len(train_cases) - 1
    ws = W.train(ctx, _features([train_cases[k] for k in rows]), np.array(targets), steps=1500, batch=64,
                 lr=1e-3, seed=0)
    b, frees, ps, qs = _batch(ctx, held)
    free_norm = b.densityResidualNorms()
    plain = solve(b, ps, qs, 0.5, PLAN, reference=free_norm)
    b.destroy()
    b, frees, ps, qs = _batch(ctx, held)
    pred = ws.predict_targets(_features(held))
    u = np.array([float(np.float32(c.mach * math.sqrt(1.4))) for c in held])
    b.setCoarse(pred, np.stack([np.ones_like(u), u, u, u * u], axis=1), shape=(W.COARSE_NJ, W.COARSE_NI))
    used = b.densityResidualNorms() <= free_norm
    for k in np.nonzero(~used)[0]:
        b.setCellPrimitives(k, np.broadcast_to(frees[k], (held[k].cells, 4)))
    warm = solve(b, ps, qs, 0.5, PLAN, reference=free_norm, cfl_start=np.where(used, PLAN.cfl, PLAN.cfl_start))
    b.destroy()
    cells = np.array([c.cells for c in held])
    w_plain = float((cells * np.array([r.steps for r in plain])).sum())
    w_warm = float((cells * np.array([r.steps for r in warm])).sum())
    dcd = max(abs(a.cd / p.cd - 1) for a, p in zip(warm, plain))
    print(len(rows), "training fields,", len(held), "held out; used", int(used.sum()), "cell-steps plain", w_plain,
          "warm", w_warm, "ratio", w_warm / w_plain, "max CD change", dcd)
    assert all(r.converged and r.residual <= 1e-6 for r in plain + warm)
    assert w_warm < 0.9 * w_plain  # measured: 19 % fewer cell-steps, drags 4.6e-7 apart
    assert dcd < 1e-5
