"""Batched solves against the FP64 oracle case by case, on the batched
configurations' own seeded cases (mulberry32, seed 20261020):

- the Mach sweep (one nose on a 4-caliber cylinder at Mach 1.2, 1.3, ...,
  5.0; SSP-RK3, minmod per variable, HLLC, CFL 0.5) as one batch;
- a slice of the rocket population as one batch of uneven grids: sharp
  noses on H grids up to 622 x 311 and blunt noses on C grids (the axis a
  wall), implicit red-black LU-SGS with the CFL ramp, with HLLC and with
  Rusanov's flux (the population's flux for the C grids).

Both as FP32 batches and as FP64 ones (bs_batch_create_fp64). After a
number of steps every case's conserved fields are compared cell by cell
with the oracle run for the same steps (the parity bars of the single
solves, per component relative to its largest value: L1 < 5e-6, Linf <
2e-4 for FP32 states and for FP64 implicit steps, whose relaxation runs in
FP32 against the oracle's linearized red-black steps; L1 < 1e-12, Linf <
1e-10 for FP64 explicit steps), and the batch's FP64 drag and density
residual norm with the oracle's on the batch's own state (1e-12, 1e-4)."""
import math

import numpy as np
import pytest

import cases
import pyoracle as O
from bowshock_amd import grids
from bowshock_amd import population as P
from bowshock_amd.batch import GpuBatch2D
from bowshock_amd.engine import Euler2DConfig, gpu_primitives

pytestmark = pytest.mark.gpu


def _run_oracle(c, bcs, scheme, free, steps, implicit, relaxation="redBlack"):
    orc = O.Euler2D(c.grid, geometry="axisymmetric", gamma=1.4, flux=scheme["flux"], limiter=scheme["limiter"],
                    reconstruction=scheme["reconstruction"], time_integrator=scheme["timeIntegrator"],
                    cfl=scheme["cfl"], time_stepping="local", boundaries=bcs)
    orc.set_primitives(np.array(free))
    if implicit:
        cfl = 1.0
        for _ in range(steps):
            orc.iterate_implicit(1, cfl, 1.5, relaxation, 2)
            cfl = min(10.0, cfl * 1.2)
    else:
        orc.iterate(steps)
    return orc


def _check(batch, accepted, scheme, steps, implicit, bars=(5e-6, 2e-4)):
    frees = [grids.free_stream_at(c.mach, 1.4) for c in accepted]
    fp64 = batch.precision == "fp64"
    worst = (0.0, 0.0)
    drags = batch.forebodyDrags(np.array([f[1] for f in frees]), np.array([f[2] for f in frees]), 0.5)
    norms = batch.densityResidualNorms()
    for k, c in enumerate(accepted):
        free, p_inf, q_inf = frees[k]
        bcs = P.boundaries(c, free)
        orc = _run_oracle(c, bcs, scheme, free, steps, implicit, "redBlackLinear" if fp64 else "redBlack")
        st = batch.readState(k)
        mine = np.stack([st.rho, st.mu, st.mv, st.energy], axis=1).astype(np.float64)
        l1, linf = cases.field_errors(mine, orc.conserved())
        assert l1.max() < bars[0] and linf.max() < bars[1], (k, c.ni, c.nj, c.topology, l1, linf)
        worst = (max(worst[0], l1.max()), max(worst[1], linf.max()))
        # The batch's FP64 reductions on its own state.
        if fp64:
            orc.set_conserved(mine)
        else:
            orc.set_primitives(gpu_primitives(st, 1.4))
        cd = orc.forebody_drag(p_inf, q_inf, 0.5)
        assert abs(drags[k] - cd) <= 1e-12 * max(1.0, abs(cd)), (k, drags[k], cd)
        assert abs(norms[k] / orc.density_residual_norm() - 1) < 1e-4, k
    print(f"{len(accepted)} cases, worst L1 {worst[0]:.2e}, Linf {worst[1]:.2e}")


def _batch(ctx, accepted, scheme, precision="fp32"):
    cfgs = []
    for c in accepted:
        free, _, _ = grids.free_stream_at(c.mach, 1.4)
        cfgs.append(Euler2DConfig(grid=c.grid, boundaries=P.boundaries(c, free), geometry="axisymmetric", gamma=1.4,
                                  timeStepping="local", **scheme))
    b = GpuBatch2D(ctx, cfgs, precision)
    b.setUniform([grids.free_stream_at(c.mach, 1.4)[0] for c in accepted])
    return b


@pytest.mark.parametrize("precision,persistent", [("fp32", 0), ("fp64", 0), ("fp32", 1), ("fp32", 2)])
def test_the_mach_sweep_as_one_batch_against_the_oracle(ctx, precision, persistent):
    """FP64 batches to the FP64 parity bar (L1 1e-12, Linf 1e-10); FP32
    batches also in persistent launches (both tile shapes)."""
    sweep = P.build_grids(P.sweep_cases(), workers=1)
    accepted = [c for c in sweep if c.grid is not None]
    assert len(sweep) == 39 and len(accepted) >= 30
    scheme = dict(flux="hllc", limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5)
    batch = _batch(ctx, accepted, scheme, precision)
    batch.setTiling(8192)  # the sweep's grids take the small tile
    if persistent:
        batch.setPersistent(persistent)
    batch.iterate(30)
    if persistent:
        assert batch.persistentInfo()[0]
    _check(batch, accepted, scheme, 30, implicit=False, bars=(1e-12, 1e-10) if precision == "fp64" else (5e-6, 2e-4))
    batch.destroy()


def _population_slice():
    """Sharp noses on H grids (the largest of the first 120 rockets' cases,
    622 x 311, the smallest, and two between) and blunt ellipsoids on C
    grids at two refinements."""
    pop = P.population_cases(120)
    sharp = [c for c in pop if c.rocket.kind not in P.BLUNT_KINDS]
    sharp.sort(key=lambda c: c.ni)
    pick = [sharp[-1], sharp[0], sharp[len(sharp) // 3], sharp[2 * len(sharp) // 3]]
    blunt = [c for c in pop if c.rocket.kind == "ellipsoid"]
    by_ref = {}
    for c in blunt:
        by_ref.setdefault(P.c_grid_refinement(c.ni), c)
    pick += [by_ref[r] for r in sorted(by_ref)[:2]]
    P.build_grids(pick, workers=1)
    return [c for c in pick if c.grid is not None]


@pytest.mark.parametrize("flux,precision", [("hllc", "fp32"), ("rusanov", "fp32"), ("hllc", "fp64"),
                                            ("rusanov", "fp64")])
def test_a_population_slice_as_one_batch_against_the_oracle(ctx, flux, precision):
    """FP64 batches relax in FP32 with the linearized term, against the
    oracle's linearized red-black steps (the FP32 relaxation's bar)."""
    accepted = _population_slice()
    assert max(c.ni for c in accepted if c.topology == "H") >= 600
    assert sum(c.topology == "C" for c in accepted) == 2
    scheme = dict(flux=flux, limiter="vanAlbada", reconstruction="sharedScaled", timeIntegrator="ssprk3", cfl=0.8)
    batch = _batch(ctx, accepted, scheme, precision)
    batch.setImplicitCfl(1.0)
    batch.iterateImplicit(6, cfl=10.0, sweeps=2)
    _check(batch, accepted, scheme, 6, implicit=True)
    batch.destroy()
