"""Batched solves (bs_batch_*) against the same cases solved one at a time
by the single-grid engine, whose parity with the oracle the other GPU
tests establish: the per-case kernels are the same, so every case's state
is bitwise equal to its single solve, on either tile class, with cases
retired in between; norms and drags agree to the reduction order."""
import math

import numpy as np
import pytest

import cases
from bowshock_amd.batch import GpuBatch2D, SolvePlan, solve
from bowshock_amd.engine import Euler2DConfig, GpuEuler2D, RangeError

pytestmark = pytest.mark.gpu

SHAPES = [((72, 36), 2.5), ((40, 20), 3.0), ((130, 64), 1.8), ((24, 12), 4.0)]


def _cases(**scheme):
    out = []
    for (ni, nj), mach in SHAPES:
        grid, free, p_inf, q_inf, bcs = cases.cone_case(ni, nj, mach=mach)
        kw = dict(geometry="axisymmetric", gamma=cases.GAMMA, flux="hllc", limiter="minmod",
                  reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5, timeStepping="local")
        kw.update(scheme)
        out.append((Euler2DConfig(grid=grid, boundaries=bcs, **kw), free, p_inf, q_inf))
    return out


def _singles(ctx, cs):
    solvers = []
    for cfg, free, _, _ in cs:
        s = GpuEuler2D.create(ctx, cfg)
        s.setCellPrimitives(np.array(free))
        solvers.append(s)
    return solvers


def _state(st):
    return np.stack([st.rho, st.mu, st.mv, st.energy])


@pytest.mark.parametrize("small_below", [0, 5000])
def test_explicit_batch_matches_single_solves(ctx, small_below):
    cs = _cases()
    batch = GpuBatch2D(ctx, [c[0] for c in cs])
    batch.setTiling(small_below)
    batch.setUniform([c[1] for c in cs])
    singles = _singles(ctx, cs)
    for count in (37, 24):
        batch.iterate(count)
        for s in singles:
            s.iterate(count)
        for k, s in enumerate(singles):
            assert np.array_equal(_state(batch.readState(k)), _state(s.readState())), k
    norms = batch.densityResidualNorms()
    p = np.array([c[2] for c in cs])
    q = np.array([c[3] for c in cs])
    drags = batch.forebodyDrags(p, q, 0.5)
    for k, s in enumerate(singles):
        assert norms[k] == pytest.approx(s.densityResidualNorm(), rel=1e-12, abs=0)
        assert drags[k] == pytest.approx(s.forebodyDrag(p[k], q[k], 0.5), rel=1e-13, abs=1e-15)
    status, steps = batch.status()
    assert status.tolist() == [0] * len(cs) and steps.tolist() == [61] * len(cs)


def test_retired_cases_stay_put(ctx):
    cs = _cases()
    batch = GpuBatch2D(ctx, [c[0] for c in cs])
    batch.setUniform([c[1] for c in cs])
    singles = _singles(ctx, cs)
    batch.iterate(30)
    for s in singles:
        s.iterate(30)
    frozen = _state(batch.readState(1))
    batch.setActive([1, 0, 1, 1])
    batch.iterate(25)
    for k in (0, 2, 3):
        singles[k].iterate(25)
    assert np.array_equal(_state(batch.readState(1)), frozen)
    for k in (0, 2, 3):
        assert np.array_equal(_state(batch.readState(k)), _state(singles[k].readState())), k
    norms = batch.densityResidualNorms()
    assert math.isnan(norms[1]) and all(math.isfinite(norms[k]) for k in (0, 2, 3))
    _, steps = batch.status()
    assert steps.tolist() == [55, 30, 55, 55]
    # Back in: it resumes where it stopped.
    batch.setActive([1, 1, 1, 1])
    batch.iterate(5)
    singles[1].iterate(5)
    assert np.array_equal(_state(batch.readState(1)), _state(singles[1].readState()))


@pytest.mark.parametrize("small_below,form", [(0, "fluxDifference"), (5000, "fluxDifference"), (0, "linearized")])
def test_implicit_batch_matches_single_solves(ctx, small_below, form):
    cs = _cases(limiter="vanAlbada", reconstruction="sharedScaled")
    batch = GpuBatch2D(ctx, [c[0] for c in cs])
    batch.setTiling(small_below)
    batch.setRelaxation(form)
    batch.setUniform([c[1] for c in cs])
    batch.setImplicitCfl(1.0)
    singles = _singles(ctx, cs)
    for s in singles:
        s.setRelaxation(form)
        s.setImplicitCfl(1.0)
    for count, sweeps in ((30, 2), (13, 2), (12, 1)):
        batch.iterateImplicit(count, 10.0, sweeps)
        for s in singles:
            s.iterateImplicit(count, cfl=10.0, sweeps=sweeps)
        for k, s in enumerate(singles):
            assert np.array_equal(_state(batch.readState(k)), _state(s.readState())), (count, k)


def test_timed_passes_run_the_same_steps(ctx):
    """bs_batch_enable_timing: the steps run eagerly with an event pair per
    pass, bitwise equal to the graph-replayed steps of an untimed batch, and
    every pass is counted."""
    cs = _cases(limiter="vanAlbada", reconstruction="sharedScaled")
    timed, plain = (GpuBatch2D(ctx, [c[0] for c in cs]) for _ in range(2))
    for b in (timed, plain):
        b.setUniform([c[1] for c in cs])
        b.setImplicitCfl(1.0)
    timed.enableTiming(True)
    for b in (timed, plain):
        b.iterateImplicit(24, 10.0, 2)
        b.iterate(12)
    for k in range(len(cs)):
        assert np.array_equal(_state(timed.readState(k)), _state(plain.readState(k))), k
    prep, n_prep = timed.kernelTiming("implicitPrepare")
    relax, n_relax = timed.kernelTiming("implicitRelax")
    stage, n_stage = timed.kernelTiming("stage")
    # Two sweeps: the prepare pass and one relaxation pass of three
    # half-sweeps per step; SSP-RK3: three stage passes per step.
    assert (n_prep, n_relax, n_stage) == (24, 24, 36)
    assert prep > 0 and relax > 0 and stage > 0
    assert timed.kernelTiming("implicitRelax") == (0.0, 0)
    timed.enableTiming(False)
    for b in (timed, plain):
        b.destroy()


def test_batch_solve_matches_single_converged_drag(ctx):
    """The driver loop retires each case once its residual and drag have
    settled; the drags equal a single solve run for the same step count."""
    cs = _cases()[:2]
    batch = GpuBatch2D(ctx, [c[0] for c in cs])
    batch.setUniform([c[1] for c in cs])
    p = np.array([c[2] for c in cs])
    q = np.array([c[3] for c in cs])
    res = solve(batch, p, q, 0.5, SolvePlan(chunk=240, max_steps=24000))
    for k, r in enumerate(res):
        assert r.converged, (k, r.reason, r.residual, r.steps)
        s = _singles(ctx, [cs[k]])[0]
        s.iterate(r.steps)
        assert s.forebodyDrag(p[k], q[k], 0.5) == pytest.approx(r.cd, rel=1e-13)


def test_batch_rejects_mixed_schemes(ctx):
    a = _cases()[0][0]
    b = _cases(limiter="vanLeer")[1][0]
    with pytest.raises(RangeError):
        GpuBatch2D(ctx, [a, b])


def _mixed_cases(**scheme):
    """Two H grids (cones, inflow iMin) and two blunt-body C grids (the
    axis ahead of the nose a wall) in one batch."""
    from bowshock_amd import grids
    from bowshock_amd.blunt_grid import blunt_body_grid

    out = _cases(**scheme)[:2]
    kw = dict(geometry="axisymmetric", gamma=cases.GAMMA, flux="hllc", limiter="minmod",
              reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5, timeStepping="local")
    kw.update(scheme)
    for body, mach in ((grids.Body(grids.Nose("ellipsoid", 0.5, 0.5), 4.0), 3.0),
                       (grids.Body(grids.Nose("ellipsoid", 1.5, 0.5), 6.0), 2.0)):
        grid = blunt_body_grid(body, 16, mach).grid
        free, p_inf, q_inf = grids.free_stream_at(mach, cases.GAMMA)
        bcs = {"iMin": {"kind": "wall"}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
               "jMax": {"kind": "inflow", "state": free}}
        out.append((Euler2DConfig(grid=grid, boundaries=bcs, **kw), free, p_inf, q_inf))
    return out


@pytest.mark.parametrize("implicit", [False, True])
def test_mixed_topologies_match_single_solves(ctx, implicit):
    """H and C grids, whose boundary kinds differ, in one batch: each case
    bitwise equal to its single solve, drags to the reduction order."""
    scheme = dict(limiter="vanAlbada", reconstruction="sharedScaled") if implicit else {}
    cs = _mixed_cases(**scheme)
    batch = GpuBatch2D(ctx, [c[0] for c in cs])
    batch.setUniform([c[1] for c in cs])
    singles = _singles(ctx, cs)
    if implicit:
        batch.setImplicitCfl(1.0)
        for s in singles:
            s.setImplicitCfl(1.0)
    for count in (25, 16):
        if implicit:
            batch.iterateImplicit(count, 10.0, 2)
        else:
            batch.iterate(count)
        for s in singles:
            if implicit:
                s.iterateImplicit(count, cfl=10.0, sweeps=2)
            else:
                s.iterate(count)
        for k, s in enumerate(singles):
            assert np.array_equal(_state(batch.readState(k)), _state(s.readState())), (count, k)
    p = np.array([c[2] for c in cs])
    q = np.array([c[3] for c in cs])
    drags = batch.forebodyDrags(p, q, 0.5)
    for k, s in enumerate(singles):
        assert drags[k] == pytest.approx(s.forebodyDrag(p[k], q[k], 0.5), rel=1e-13, abs=1e-15)


def test_coarse_fields_set_on_the_gpu_equal_the_host_resampling(ctx):
    """bs_batch_set_coarse against bs_batch_set_primitives of the host's
    resampling (warmstart.from_target) of the same seeded fields: every
    case's state bit for bit, H and C grids alike; a NaN field is refused
    as non-physical."""
    from bowshock_amd import warmstart as W
    from bowshock_amd.engine import NonPhysicalStateError

    cs = _mixed_cases()
    cfgs = [c[0] for c in cs]
    rng = np.random.default_rng(5)
    n_coarse = W.COARSE_NI * W.COARSE_NJ
    base = np.array([1.0, 0.9, 0.05, 0.6])
    targets = np.stack([np.concatenate([b + 0.3 * rng.standard_normal(n_coarse) for b in base]) for _ in cs])
    targets[0, :n_coarse] = -0.5          # density below the floor everywhere in case 0
    u_inf = np.array([c[1][1] for c in cs])
    scale = np.stack([np.ones_like(u_inf), u_inf, u_inf, u_inf * u_inf], axis=1)
    gpu = GpuBatch2D(ctx, cfgs)
    gpu.setCoarse(targets, scale, 0.05, shape=(W.COARSE_NJ, W.COARSE_NI))
    host = GpuBatch2D(ctx, cfgs)
    for k, cfg in enumerate(cfgs):
        host.setCellPrimitives(k, W.from_target(targets[k], u_inf[k], cfg.grid.ni, cfg.grid.nj, floor=0.05))
    for k in range(len(cs)):
        assert np.array_equal(_state(gpu.readState(k)), _state(host.readState(k))), k
    assert np.all(gpu.readState(0).rho == np.float32(0.05))
    bad = targets.copy()
    bad[1, 7] = np.nan
    with pytest.raises(NonPhysicalStateError):
        gpu.setCoarse(bad, scale, 0.05, shape=(W.COARSE_NJ, W.COARSE_NI))
    with pytest.raises(RangeError):
        gpu.setCoarse(targets[:, :-1], scale)


# ---- FP64 batches (bs_batch_create_fp64) -----------------------------------


def _singles64(ctx, cs):
    solvers = []
    for cfg, free, _, _ in cs:
        s = GpuEuler2D.create(ctx, cfg, "fp64")
        s.setCellPrimitives(np.array(free))
        solvers.append(s)
    return solvers


@pytest.mark.parametrize("small_below", [0, 5000])
def test_fp64_batch_matches_fp64_single_solves(ctx, small_below):
    """Explicit and implicit steps of an FP64 batch: every case bitwise
    equal to the FP64 solver's single solve (FP64 tiles, linearized FP32
    relaxation, FP64 update), H and C grids, either tile class; FP64 state
    read back as FP64."""
    cs = _mixed_cases(limiter="vanAlbada", reconstruction="sharedScaled")
    batch = GpuBatch2D(ctx, [c[0] for c in cs], precision="fp64")
    batch.setTiling(small_below)
    batch.setUniform([c[1] for c in cs])
    singles = _singles64(ctx, cs)
    batch.iterate(14)
    for s in singles:
        s.iterate(14)
    batch.setImplicitCfl(1.0)
    for s in singles:
        s.setImplicitCfl(1.0)
    for count in (25, 13):
        batch.iterateImplicit(count, 10.0, 2)
        for s in singles:
            s.iterateImplicit(count, cfl=10.0, sweeps=2)
        for k, s in enumerate(singles):
            st = batch.readState(k)
            assert st.rho.dtype == np.float64
            assert np.array_equal(_state(st), _state(s.readState())), (count, k)
    p = np.array([c[2] for c in cs])
    q = np.array([c[3] for c in cs])
    drags = batch.forebodyDrags(p, q, 0.5)
    norms = batch.densityResidualNorms()
    for k, s in enumerate(singles):
        assert drags[k] == pytest.approx(s.forebodyDrag(p[k], q[k], 0.5), rel=1e-13, abs=1e-15)
        assert norms[k] == pytest.approx(s.densityResidualNorm(), rel=1e-12, abs=0)
    with pytest.raises(Exception):
        batch.setRelaxation("fluxDifference")


def test_fp64_batch_converges_below_the_fp32_floor(ctx):
    """Population-style implicit solves (CFL 1 -> 10, 2 red-black sweeps)
    of an FP64 batch reach a density residual of 1e-6 (absolute), which the
    FP32 batch's state cannot (its floor), with drags settled to 1e-5."""
    cs = _cases(limiter="vanAlbada", reconstruction="sharedScaled")
    p = np.array([c[2] for c in cs])
    q = np.array([c[3] for c in cs])
    plan = SolvePlan(implicit=True, chunk=60, max_steps=6000, cfl=10.0, cfl_start=1.0, sweeps=2, residual=1e-6,
                     drop=0.0, stall=10 ** 9)
    batch = GpuBatch2D(ctx, [c[0] for c in cs], precision="fp64")
    batch.setUniform([c[1] for c in cs])
    res = solve(batch, p, q, 0.5, plan)
    for k, r in enumerate(res):
        assert r.converged and r.residual <= 1e-6, (k, r.reason, r.residual, r.steps)


@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("integrator", ["ssprk3", "ssprk2", "forwardEuler"])
def test_persistent_launches_equal_per_stage_launches(ctx, tile, integrator):
    """Persistent launches (every stage of every active case in one launch,
    a case's blocks meeting at device-side barriers between stages) give
    the states of one launch per stage bit for bit: H and C grids, runs of
    steps that cross the 60-step launch boundary, a case retired between
    runs (its state left as it was), and the same status words and step
    counts."""
    cs = _mixed_cases(timeIntegrator=integrator)
    ref = GpuBatch2D(ctx, [c[0] for c in cs])
    per = GpuBatch2D(ctx, [c[0] for c in cs])
    for b in (ref, per):
        b.setUniform([c[1] for c in cs])
    per.setPersistent(tile)
    for count in (37, 71):
        ref.iterate(count)
        per.iterate(count)
        used, blocks, capacity = per.persistentInfo()
        assert used and 0 < blocks <= capacity, (used, blocks, capacity)
        for k in range(len(cs)):
            assert np.array_equal(_state(per.readState(k)), _state(ref.readState(k))), (count, k)
    mask = [1, 0, 1, 1]
    for b in (ref, per):
        b.setActive(mask)
        b.iterate(25)
    assert per.persistentInfo()[0]
    for k in range(len(cs)):
        assert np.array_equal(_state(per.readState(k)), _state(ref.readState(k))), ("retired", k)
    s_ref, n_ref = ref.status()
    s_per, n_per = per.status()
    assert s_per.tolist() == s_ref.tolist() == [0] * len(cs)
    assert n_per.tolist() == n_ref.tolist() == [133, 108, 133, 133]
    ref.destroy()
    per.destroy()


def test_persistent_launches_fall_back_when_the_cases_do_not_fit(ctx):
    """A case of 1280 x 640 cells on 32 x 4 tiles needs more blocks than the
    chip holds at once: the batch runs its per-stage launches instead, with
    the same states."""
    grid, free, _, _, bcs = cases.cone_case(1280, 640, mach=3.0)
    cfg = Euler2DConfig(grid=grid, boundaries=bcs, geometry="axisymmetric", gamma=cases.GAMMA, flux="hllc",
                        limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5,
                        timeStepping="local")
    ref = GpuBatch2D(ctx, [cfg])
    per = GpuBatch2D(ctx, [cfg])
    for b in (ref, per):
        b.setUniform([free])
    per.setPersistent(1)
    used, blocks, capacity = per.persistentInfo()
    assert blocks == 0 and capacity > 0
    ref.iterate(6)
    per.iterate(6)
    assert not per.persistentInfo()[0]
    assert np.array_equal(_state(per.readState(0)), _state(ref.readState(0)))
    ref.destroy()
    per.destroy()


def test_fp64_batches_refuse_persistent_launches(ctx):
    cs = _cases()
    b = GpuBatch2D(ctx, [c[0] for c in cs], "fp64")
    with pytest.raises(RangeError):
        b.setPersistent(1)
    b.destroy()


@pytest.mark.parametrize("fusion,sweeps", [(1, 2), (2, 2), (1, 3), (2, 3)])
def test_relaxation_fusion_keeps_the_states(ctx, fusion, sweeps):
    """Fewer red-black half-sweeps per relaxation launch (more launches)
    give the states of the default three per launch bit for bit."""
    cs = _mixed_cases(limiter="vanAlbada", reconstruction="sharedScaled")
    ref = GpuBatch2D(ctx, [c[0] for c in cs])
    fused = GpuBatch2D(ctx, [c[0] for c in cs])
    fused.setRelaxFusion(fusion)
    for b in (ref, fused):
        b.setUniform([c[1] for c in cs])
        b.setImplicitCfl(1.0)
        b.iterateImplicit(13, 10.0, sweeps)
    for k in range(len(cs)):
        assert np.array_equal(_state(fused.readState(k)), _state(ref.readState(k))), k
    with pytest.raises(RangeError):
        fused.setRelaxFusion(4)
    ref.destroy()
    fused.destroy()


@pytest.mark.parametrize("implicit", [False, True])
def test_the_autotuner_picks_a_launch_plan_with_the_same_states(ctx, tmp_path, implicit):
    """autotune times the launch plans on a throwaway batch, caches the
    fastest by the batch's size histogram, and a batch run under the plan
    (read back from the cache) gives the states of the default plan."""
    from bowshock_amd.batch import Tuning, autotune

    scheme = dict(limiter="vanAlbada", reconstruction="sharedScaled") if implicit else {}
    cs = _mixed_cases(**scheme)
    cache = str(tmp_path / "tiling.json")
    probe = GpuBatch2D(ctx, [c[0] for c in cs])
    probe.setUniform([c[1] for c in cs])
    if implicit:
        probe.setImplicitCfl(1.0)
    t = autotune(probe, steps=12, implicit=implicit, cache=cache)
    assert isinstance(t, Tuning) and t.relax_fusion in (1, 2, 3) and t.persistent in (0, 1, 2)
    probe.destroy()
    tuned = GpuBatch2D(ctx, [c[0] for c in cs])
    assert autotune(tuned, implicit=implicit, cache=cache) == t  # from the cache
    ref = GpuBatch2D(ctx, [c[0] for c in cs])
    for b in (ref, tuned):
        b.setUniform([c[1] for c in cs])
        if implicit:
            b.setImplicitCfl(1.0)
            b.iterateImplicit(20, 10.0, 2)
        else:
            b.iterate(40)
    for k in range(len(cs)):
        assert np.array_equal(_state(tuned.readState(k)), _state(ref.readState(k))), k
    ref.destroy()
    tuned.destroy()


def test_two_batches_driven_from_two_threads_equal_serial_runs(ctx):
    """Each batch owns a stream for its kernels and copies. Two batches
    advanced from two host threads at once (graph captures, work-list
    uploads after a case retires, status and state reads, all overlapping)
    end in the same states, bit for bit, as the same batches advanced one
    after the other."""
    from concurrent.futures import ThreadPoolExecutor

    cs = _cases()
    halves = (cs[:2], cs[2:])

    def run(batch, group, implicit):
        batch.setUniform([c[1] for c in group])
        for chunk, active in ((40, (1, 1)), (80, (1, 1)), (60, (0, 1)), (50, (1, 1))):
            batch.setActive(np.array(active, dtype=np.uint8))
            if implicit:
                batch.iterateImplicit(chunk, 10.0, 2)
            else:
                batch.iterate(chunk)
            batch.status()
            batch.densityResidualNorms()
        return [_state(batch.readState(k)) for k in range(batch.count)]

    serial = []
    for implicit, group in zip((False, True), halves):
        b = GpuBatch2D(ctx, [c[0] for c in group])
        serial.append(run(b, group, implicit))
        b.destroy()
    batches = [GpuBatch2D(ctx, [c[0] for c in group]) for group in halves]
    with ThreadPoolExecutor(2) as ex:
        futures = [ex.submit(run, b, group, implicit) for b, group, implicit in zip(batches, halves, (False, True))]
        concurrent = [f.result() for f in futures]
    for b in batches:
        b.destroy()
    for s, c in zip(serial, concurrent):
        for a, d in zip(s, c):
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, d)


def test_mixed_precision_solve_reaches_the_fp64_target(ctx):
    """FP32 steps to the FP32 stopping rule, then FP64 steps from the
    widened states: every case reaches the FP64 residual target, with the
    drag of the FP64 solve from the free stream within the drag tolerance,
    and the steps of both phases counted."""
    from bowshock_amd.batch import solve_mixed

    cs = _cases()
    cfgs = [c[0] for c in cs]
    frees = [c[1] for c in cs]
    p = np.array([c[2] for c in cs])
    q = np.array([c[3] for c in cs])
    plan32 = SolvePlan(implicit=True, chunk=60, max_steps=6000, cfl=10.0, cfl_start=1.0, sweeps=2)
    plan64 = SolvePlan(implicit=True, chunk=60, max_steps=12000, cfl=10.0, cfl_start=1.0, sweeps=2, residual=1e-6,
                       drop=0.0, stall=10 ** 9)
    b64 = GpuBatch2D(ctx, cfgs, "fp64")
    b64.setUniform(frees)
    plain = solve(b64, p, q, 0.5, plan64)
    b32 = GpuBatch2D(ctx, cfgs)
    b32.setUniform(frees)
    b64.setUniform(frees)
    with pytest.raises(RangeError):
        solve_mixed(b64, b32, frees, p, q, 0.5, plan32, plan64)
    mixed = solve_mixed(b32, b64, frees, p, q, 0.5, plan32, plan64)
    for k, (a, m) in enumerate(zip(plain, mixed)):
        assert a.converged and m.converged, (k, a.reason, m.reason)
        assert m.residual <= 1e-6
        assert m.cd == pytest.approx(a.cd, rel=1e-5)
        assert len(m.history) * plan64.chunk == m.steps
    b32.destroy()
    b64.destroy()
