"""Steps replayed from captured HIP graphs against the same steps launched
kernel by kernel: bitwise equal after every batch, with eager work (state
reads, residual norms, drags) between the replays."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _pair(ctx, **scheme):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(72, 36, mach=2.5)
    _, a = cases.make_pair(ctx, grid, bcs, **scheme)
    _, b = cases.make_pair(ctx, grid, bcs, **scheme)
    b.enableGraphs(False)
    for s in (a, b):
        s.setCellPrimitives(np.array(free))
    return a, b, p_inf, q_inf


@pytest.mark.parametrize("integrator,stepping", [("ssprk3", "local"), ("ssprk2", "global"), ("forwardEuler", "local")])
def test_explicit_graph_replays_match_eager(ctx, integrator, stepping):
    a, b, p_inf, q_inf = _pair(ctx, integrator=integrator, stepping=stepping, cfl=0.4)
    for count in (37, 24, 12, 50):
        a.iterate(count)
        b.iterate(count)
        assert np.array_equal(cases.gpu_conserved(a), cases.gpu_conserved(b))
        assert a.densityResidualNorm() == b.densityResidualNorm()
        assert a.forebodyDrag(p_inf, q_inf, 0.5) == b.forebodyDrag(p_inf, q_inf, 0.5)
    assert a.last_report.steps == b.last_report.steps == 123
    assert a.last_report.time == b.last_report.time


def test_implicit_graph_replays_match_eager(ctx):
    a, b, p_inf, q_inf = _pair(ctx, limiter="vanAlbada", reconstruction="sharedScaled")
    for s in (a, b):
        s.setImplicitCfl(1.0)
    for count in (30, 13, 40):
        a.iterateImplicit(count, cfl=10.0, sweeps=2)
        b.iterateImplicit(count, cfl=10.0, sweeps=2)
        assert np.array_equal(cases.gpu_conserved(a), cases.gpu_conserved(b))
        assert a.forebodyDrag(p_inf, q_inf, 0.5) == b.forebodyDrag(p_inf, q_inf, 0.5)
    # A change of sweeps recaptures.
    a.iterateImplicit(24, cfl=10.0, sweeps=1)
    b.iterateImplicit(24, cfl=10.0, sweeps=1)
    assert np.array_equal(cases.gpu_conserved(a), cases.gpu_conserved(b))


def test_region_timing_counts_stage_launches(ctx):
    """Region timing (events around whole batches of explicit local steps,
    graphs on) counts every stage launch and agrees with per-launch events
    to within the launch gaps."""
    a, _, _, _ = _pair(ctx)
    a.enableTiming("region")
    a.iterate(300)
    ms_region, n_region = a.kernelTiming("stage")
    assert n_region == 900 and ms_region > 0
    a.enableTiming(True)
    a.iterate(60)
    ms_launch, n_launch = a.kernelTiming("stage")
    assert n_launch == 180 and ms_launch > 0
    a.enableTiming(False)
    assert 0.2 * ms_launch < ms_region < 5 * ms_launch
