"""Cell-by-cell parity at the roofline ladder's full sizes: a few explicit
and implicit steps from the free stream on 640 x 320 and 1280 x 640
cone- and ogive-cylinder grids against the FP64 oracle on the same grid
(the converged-drag checks run on smaller grids, where the oracle reaches
convergence in seconds)."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("body,ni,nj,mach", [("cone", 640, 320, 1.5), ("ogive", 640, 320, 3.0),
                                             ("cone", 1280, 640, 3.0)])
def test_explicit_steps_at_ladder_size(ctx, body, ni, nj, mach):
    make = cases.cone_case if body == "cone" else cases.ogive_case
    grid, free, p_inf, q_inf, bcs = make(ni, nj, mach=mach)
    orc, gpu = cases.make_pair(ctx, grid, bcs)
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    orc.iterate(3)
    gpu.iterate(3)
    l1, linf = cases.field_errors(cases.gpu_conserved(gpu), orc.conserved())
    assert l1.max() < 5e-6 and linf.max() < 2e-4, (l1, linf)
    assert gpu.forebodyDrag(p_inf, q_inf, 0.5) == pytest.approx(orc.forebody_drag(p_inf, q_inf, 0.5), rel=1e-4)


@pytest.mark.parametrize("body,ni,nj,mach", [("cone", 640, 320, 3.0), ("ogive", 1280, 640, 1.5)])
def test_implicit_steps_at_ladder_size(ctx, body, ni, nj, mach):
    make = cases.cone_case if body == "cone" else cases.ogive_case
    grid, free, p_inf, q_inf, bcs = make(ni, nj, mach=mach)
    orc, gpu = cases.make_pair(ctx, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc.set_primitives(np.array(free))
    gpu.setCellPrimitives(np.array(free))
    gpu.setImplicitCfl(1.0)
    cfl = 1.0
    for _ in range(2):
        orc.iterate_implicit(1, cfl, 1.5, relaxation="redBlack", sweeps=2)
        cfl = min(10.0, cfl * 1.2)
    gpu.iterateImplicit(2, cfl=10.0, sweeps=2)
    l1, linf = cases.field_errors(cases.gpu_conserved(gpu), orc.conserved())
    assert l1.max() < 5e-6 and linf.max() < 2e-4, (l1, linf)
