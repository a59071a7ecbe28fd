"""Single-solve throughput of the HIP Euler engine on one MI355X.

Workload (the largest single-GPU grid of the single-solve ladder): the
15 deg cone with a 4-caliber cylinder, Mach 3, on a 1280 x 640 body-fitted
H grid; MUSCL with the minmod limiter per variable, HLLC, SSP-RK3 at CFL
0.5, local time steps, FP32 state. A step is one SSP-RK3 pseudo-time step
of every cell; the metric is cell-updates per second (cells x steps / s).

The timed region starts with the state resident in HBM (after `--warmup`
steps from the free stream, so the bow shock has formed) and brackets
exactly `--steps` steps with device synchronisations. The dominant kernel's
roofline comes from HIP events around its launches on the solver's stream
in a second, separate pass; the CPU baseline times the FP64 oracle on one
host core on a bounded sample of the same workload.

`--workload sweep` and `--workload population` measure the batched
solves instead: a 39-point Mach sweep of one seeded nose (SSP-RK3), and a
seeded population of nose-plus-cylinder bodies at Mach 1.5, 2, 3 and 4
(implicit red-black LU-SGS), every case run to a converged drag in one
batch; the metric is converged solves per second and the aggregate
cell-update rate. Their CPU baseline runs the oracle on all host cores,
one case per worker.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
# Algorithmic HBM bytes per cell-update of SSP-RK3 with local steps: stage 1
# 60 B, stages 2 and 3 76 B each (state 16 B per copy, metrics 24 B, dt 4 B).
BYTES_PER_CU_RK3 = 212.0
BYTES_PER_STAGE = BYTES_PER_CU_RK3 / 3.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--ni", type=int, default=1280)
    ap.add_argument("--nj", type=int, default=640)
    ap.add_argument("--mach", type=float, default=3.0)
    ap.add_argument("--cpu-steps", type=int, default=40, help="oracle steps timed for the CPU baseline (0: skip)")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernel by kernel instead of HIP graphs")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "stage_kernel_1280x640_pmc.txt"))
    ap.add_argument("--workload", choices=("single", "sweep", "population"), default="single")
    ap.add_argument("--rockets", type=int, default=2000, help="population: rockets drawn (4 Mach numbers each)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="batched CPU baseline: seconds of oracle work")
    return ap.parse_args()


def read_traffic(path):
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or
    None when no summary exists."""
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}); "
              "the batched solves would shard by case, one process per GPU, with no exchange.", file=sys.stderr)
        sys.exit(2)

    if a.workload != "single":
        return run_batch(a)

    import numpy as np

    from bowshock_amd import grids
    from bowshock_amd.engine import Context, Euler2DConfig, GpuEuler2D

    gamma = 1.4
    body = grids.cone_cylinder(15.0, 4.0)
    t0 = time.time()
    bg = grids.body_grid_fixed(body, a.ni, a.nj, a.mach, gamma)
    free, p_inf, q_inf = grids.free_stream_at(a.mach, gamma)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
    scheme = dict(geometry="axisymmetric", gamma=gamma, flux="hllc", limiter="minmod", reconstruction="perVariable",
                  timeIntegrator="ssprk3", cfl=0.5, timeStepping="local")
    print(f"grid {a.ni}x{a.nj} built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    ctx = Context(0)
    gpu = GpuEuler2D.create(ctx, Euler2DConfig(grid=bg.grid, boundaries=bcs, **scheme))
    gpu.setCellPrimitives(np.array(free))
    if a.no_graphs:
        gpu.enableGraphs(False)
    gpu.iterate(a.warmup)
    cells = a.ni * a.nj

    # Timed region: exactly `steps` steps, synchronised on both sides
    # (bs_iterate returns after the device has finished its last batch).
    gpu.iterate(1)
    t = time.perf_counter()
    gpu.iterate(a.steps)
    elapsed = time.perf_counter() - t
    ms_per_step = 1e3 * elapsed / a.steps
    value = cells * a.steps / elapsed
    print(f"{a.steps} steps in {elapsed:.3f}s: {ms_per_step * 1e3:.1f} us/step, {value:.3e} CU/s",
          file=sys.stderr, flush=True)

    # Kernel timing pass: HIP events on the solver's stream around each
    # batch of steps (explicit local steps launch the stage kernel only),
    # graphs on as in the timed region.
    gpu.enableTiming("region")
    gpu.iterate(min(a.steps, 1200))
    ms_launch, launches = gpu.kernelTiming("stage")
    gpu.enableTiming(False)
    bytes_per_launch = BYTES_PER_STAGE * cells
    achieved = bytes_per_launch / (ms_launch * 1e-3) if ms_launch > 0 else 0.0
    traffic = read_traffic(a.traffic)
    cd = gpu.forebodyDrag(p_inf, q_inf, 0.5)
    res = gpu.densityResidualNorm()
    print(f"stage kernel {ms_launch * 1e3:.2f} us/launch over {launches} launches; "
          f"{achieved / 1e9:.0f} GB/s algorithmic; CD {cd:.6f}; residual {res:.3e}", file=sys.stderr, flush=True)

    cpu = None
    if a.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        orc = O.Euler2D(bg.grid, geometry="axisymmetric", gamma=gamma, flux="hllc", limiter="minmod",
                        reconstruction="perVariable", time_integrator="ssprk3", cfl=0.5, time_stepping="local",
                        boundaries=bcs)
        orc.set_primitives(np.array(free))
        orc.iterate(1)
        t = time.perf_counter()
        orc.iterate(a.cpu_steps)
        ce = time.perf_counter() - t
        cpu = {"value": cells * a.cpu_steps / ce, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"{a.cpu_steps} SSP-RK3 steps of the same {a.ni}x{a.nj} case on one core, "
                         f"FP64 C oracle (gcc -O3), {ce:.1f}s"}
        print(f"cpu baseline {cpu['value']:.3e} CU/s ({ce:.1f}s)", file=sys.stderr, flush=True)

    line = {
        "metric": "Euler cell-updates/s, %HBM roofline; converged CD solves/s warm-start vs plain",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: generated cone-cylinder grid, free-stream start",
        "config": {"workload": f"single solve, 15 deg cone + 4 cal cylinder, M{a.mach:g}, {a.ni}x{a.nj}, "
                               "MUSCL minmod per-variable, HLLC, SSP-RK3 CFL 0.5, local dt",
                   "cells": cells, "grid": [a.ni, a.nj], "scheme": "ssprk3", "reconstruction": "perVariable/minmod",
                   "flux": "hllc", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "tile_kernel (explicit stage)", "bytes_per_cell_update": BYTES_PER_CU_RK3,
                     "us_per_launch": ms_launch * 1e3},
        "cpu_baseline": cpu,
        "drag": cd,
    }
    print(json.dumps(line), flush=True)
    gpu.destroy()


# Algorithmic bytes per cell-update of an implicit step with s red-black
# sweeps: flux and prepare 60, each sweep 116 (two half-sweeps of 58),
# update 48.
def bytes_per_cu_implicit(sweeps):
    return 108.0 + 116.0 * sweeps


def _cpu_batch_job(args):
    """One oracle case for a bounded number of steps (a worker of the
    batched CPU baseline); returns cell-updates and seconds."""
    grid, bcs, free, scheme, implicit, steps = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O

    orc = O.Euler2D(grid, geometry="axisymmetric", gamma=1.4, flux=scheme["flux"], limiter=scheme["limiter"],
                    reconstruction=scheme["reconstruction"], time_integrator=scheme["timeIntegrator"],
                    cfl=scheme["cfl"], time_stepping="local", boundaries=bcs)
    orc.set_primitives(np.array(free))
    t = time.perf_counter()
    cfl = 1.0
    for _ in range(steps):
        if implicit:
            orc.iterate_implicit(1, cfl, 1.5, relaxation="redBlack", sweeps=2)
            cfl = min(10.0, cfl * 1.2)
        else:
            orc.iterate(1)
    return grid.ni * grid.nj * steps, time.perf_counter() - t


def run_batch(a):
    import numpy as np
    from concurrent.futures import ProcessPoolExecutor

    from bowshock_amd import grids
    from bowshock_amd import population as P
    from bowshock_amd.batch import GpuBatch2D, SolvePlan, autotune_tiling, solve
    from bowshock_amd.engine import Context, Euler2DConfig

    sweep = a.workload == "sweep"
    cases = P.sweep_cases() if sweep else P.population_cases(a.rockets)
    t0 = time.time()
    P.build_grids(cases)
    t_grids = time.time() - t0
    refusals = P.refusal_counts(cases)
    ok = [c for c in cases if c.grid is not None]
    print(f"{len(cases)} cases, grids in {t_grids:.1f}s: {refusals}", file=sys.stderr, flush=True)
    if sweep:
        scheme = dict(flux="hllc", limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5)
        plan = SolvePlan(implicit=False, chunk=240, max_steps=20000)
    else:
        scheme = dict(flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled", timeIntegrator="ssprk3",
                      cfl=0.8)
        plan = SolvePlan(implicit=True, chunk=60, max_steps=6000, cfl=10.0, cfl_start=1.0, sweeps=2)
    frees, p_inf, q_inf, bcs_all, configs = [], [], [], [], []
    for c in ok:
        free, p, q = grids.free_stream_at(c.mach, 1.4)
        bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
               "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
        frees.append(free)
        p_inf.append(p)
        q_inf.append(q)
        bcs_all.append(bcs)
        configs.append(Euler2DConfig(grid=c.grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4,
                                     timeStepping="local", **scheme))
    cells = np.array([c.cells for c in ok])
    ctx = Context(0)
    t0 = time.time()
    batch = GpuBatch2D(ctx, configs)
    t_create = time.time() - t0
    batch.setUniform(frees)
    tile = autotune_tiling(batch, steps=24 if sweep else 12, implicit=not sweep)
    batch.setUniform(frees)
    print(f"batch of {len(ok)} cases ({cells.sum():.3e} cells) created in {t_create:.1f}s; small tile below "
          f"{tile} cells", file=sys.stderr, flush=True)
    t = time.perf_counter()
    results = solve(batch, p_inf, q_inf, 0.5, plan)
    wall = time.perf_counter() - t
    steps = np.array([r.steps for r in results])
    converged = sum(r.converged for r in results)
    cu = float((cells * steps).sum())
    solves = converged / wall
    bpc = BYTES_PER_CU_RK3 if sweep else bytes_per_cu_implicit(plan.sweeps)
    achieved = bpc * cu / wall
    reasons = {}
    for r in results:
        reasons[r.reason] = reasons.get(r.reason, 0) + 1
    print(f"{converged}/{len(ok)} converged in {wall:.2f}s: {solves:.2f} solves/s, {cu / wall:.3e} CU/s; "
          f"{reasons}; steps median {int(np.median(steps))}", file=sys.stderr, flush=True)

    # CPU baseline: the oracle on every host core, one case per worker,
    # each for a bounded number of steps; solves/s from the GPU run's
    # cell-updates per converged solve.
    cpu = None
    if a.cpu_seconds > 0:
        workers = min(16, os.cpu_count() or 1)
        per_cu = 1.0 / 6e5 if sweep else 1.0 / 1e6   # rough oracle seconds per cell-update, one core
        pick = np.linspace(0, len(ok) - 1, workers).astype(int)
        jobs = []
        for k in pick:
            st = max(2, int(a.cpu_seconds / (per_cu * cells[k])))
            jobs.append((ok[k].grid, bcs_all[k], frees[k], scheme, not sweep, st))
        t = time.perf_counter()
        with ProcessPoolExecutor(workers) as ex:
            out = list(ex.map(_cpu_batch_job, jobs))
        cpu_wall = time.perf_counter() - t
        cpu_cu = sum(o[0] for o in out)
        rate = cpu_cu / cpu_wall
        cu_per_solve = cu / max(1, converged)
        cpu = {"value": rate / cu_per_solve, "unit": "solves/s", "cores": workers, "kind": "port",
               "sample": f"FP64 C oracle, {workers} processes, one case each ({len(jobs)} cases spread over the "
                         f"batch), {cpu_cu:.3e} cell-updates in {cpu_wall:.1f}s = {rate:.3e} CU/s; solves/s = "
                         f"that rate / the GPU run's {cu_per_solve:.3e} cell-updates per converged solve"}
        print(f"cpu baseline {rate:.3e} CU/s on {workers} cores", file=sys.stderr, flush=True)
    cds = [r.cd for r in results if r.converged]
    line = {
        "metric": "Euler cell-updates/s, %HBM roofline; converged CD solves/s warm-start vs plain",
        "value": solves,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": int(steps.max()) if len(steps) else 0,
        "warmup": 0,
        "ms_per_step": 1e3 * wall / max(1, int(steps.max())),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded bodies (mulberry32, seed 20261020), generated H grids, free-stream start",
        "config": {"workload": ("Mach sweep 1.2:0.1:5.0 of one seeded nose on a 4-caliber cylinder, SSP-RK3"
                                if sweep else f"population of {a.rockets} seeded rockets x M 1.5/2/3/4, "
                                              "LU-SGS red-black 2 sweeps, CFL 1 -> 10 x1.2"),
                   "cases": len(cases), "accepted": len(ok), "refusals": refusals, "converged": converged,
                   "outcomes": reasons, "cells": int(cells.sum()), "cell_updates": cu,
                   "cell_updates_per_s": cu / wall, "grid_seconds": t_grids, "create_seconds": t_create,
                   "small_tile_below": tile, "parallelism": "single GPU, one batch"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "whole batched step (all kernels)",
                     "bytes_per_cell_update": bpc},
        "cpu_baseline": cpu,
        "drag_range": [min(cds), max(cds)] if cds else None,
    }
    print(json.dumps(line), flush=True)
    batch.destroy()


if __name__ == "__main__":
    main()
