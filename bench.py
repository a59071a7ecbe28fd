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
one case per worker. `--workload warmstart` trains the warm-start network
on the converged fields of three quarters of the population's rockets and
reports converged solves/s of the held-out quarter's cases started from
the network's prediction (the blunt noses; the sharp noses' H grids
start from the free stream, where the prediction costs steps) against the
same cases all started from the free stream, with the network's dominant
GEMM against the BF16 MFMA peak; the prediction for every case is
reported beside it.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
# Steps of the kernel-timing pass, fixed so the per-launch figure does not
# depend on --steps.
TIMING_STEPS = 1200
# Untimed steps before the warm-up (about 30 ms at 1280 x 640): the GPU
# clocks down when idle and takes that long to reach its steady clock, and
# the step graphs are captured. Measured: 20 steps right after 10 ms of
# idle take 7 % longer per step than back to back.
PREWARM_STEPS = 600
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
    ap.add_argument("--precision", choices=("fp32", "fp64", "mixed"), default="fp32",
                    help="the FP32 solver (batches), or the FP64 one (state, metrics and residual in FP64; the "
                         "batched workloads then stop on a density residual of 1e-6); population only: mixed, "
                         "FP32 batches to their residual floor, then FP64 batches from their states to 1e-6")
    ap.add_argument("--traffic", default=None, help="rocprofv3 PMC summary of the stage kernel")
    ap.add_argument("--workload", choices=("single", "implicit", "sweep", "population", "warmstart", "ladder", "verified",
                                           "fitted"), default="single")
    ap.add_argument("--train-steps", type=int, default=3000, help="warmstart: Adam steps of the network")
    ap.add_argument("--warm-snap", type=float, default=0.0,
                    help="warmstart: predicted cells within this of the free stream start as the free stream")
    ap.add_argument("--rockets", type=int, default=2000, help="population: rockets drawn (4 Mach numbers each)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="batched CPU baseline: seconds of oracle work")
    ap.add_argument("--no-solves", action="store_true",
                    help="single: skip the converged-solves leg (the seeded Mach sweep solved to convergence)")
    ap.add_argument("--persistent", choices=("auto", "0", "1", "2"), default="auto",
                    help="explicit batches: the autotuner's launch plan, or persistent launches on 32 x 4 (1) or "
                         "32 x 8 (2) tiles, or per-stage launches (0)")
    ap.add_argument("--no-fp64-sample", action="store_true",
                    help="batched workloads: skip the FP32-vs-FP64 drag sample (FP64 single solves)")
    return ap.parse_args()


def host_info():
    """The host's logical CPU count and CPU model (os.cpu_count and
    /proc/cpuinfo, the figures lscpu reports)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


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
    if a.precision == "mixed" and a.workload != "population":
        sys.exit("--precision mixed applies to the population workload")

    if a.workload == "warmstart":
        print(json.dumps(run_warmstart(a)), flush=True)
        return None
    if a.workload == "ladder":
        return run_ladder(a)
    if a.workload == "implicit":
        return run_implicit(a)
    if a.workload == "verified":
        return run_verified(a)
    if a.workload == "fitted":
        return run_fitted(a)
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
    gpu = GpuEuler2D.create(ctx, Euler2DConfig(grid=bg.grid, boundaries=bcs, **scheme), a.precision)
    gpu.setCellPrimitives(np.array(free))
    if a.no_graphs:
        gpu.enableGraphs(False)
    # Untimed: the pre-warm, one call of the timed region's length (which
    # captures the graphs it replays), then the `--warmup` steps.
    gpu.iterate(PREWARM_STEPS)
    gpu.iterate(a.steps)
    gpu.iterate(a.warmup)
    cells = a.ni * a.nj

    # Timed region: exactly `steps` steps, synchronised on both sides
    # (bs_iterate returns after the device has finished its last batch).
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
    gpu.iterate(TIMING_STEPS)
    ms_launch, launches = gpu.kernelTiming("stage")
    gpu.enableTiming(False)
    width = 2 if a.precision == "fp64" else 1   # every algorithmic byte doubles in FP64
    bytes_per_launch = width * BYTES_PER_STAGE * cells
    achieved = bytes_per_launch / (ms_launch * 1e-3) if ms_launch > 0 else 0.0
    tag = "_fp64" if a.precision == "fp64" else ""
    traffic = read_traffic(a.traffic or os.path.join(ROOT, "profiles", f"stage_kernel{tag}_{a.ni}x{a.nj}_pmc.txt"))
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
                         f"FP64 C oracle (gcc -O3), {ce:.1f}s", "host": host_info()}
        print(f"cpu baseline {cpu['value']:.3e} CU/s ({ce:.1f}s)", file=sys.stderr, flush=True)

    solves = None
    if not a.no_solves:
        solves = {"plain_sweep": converged_sweep(ctx), "warm_start_sample": warm_start_sample(a)}

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
        "dtype": a.precision,
        "data": "synthetic: generated cone-cylinder grid, free-stream start",
        "config": {"workload": f"single solve, 15 deg cone + 4 cal cylinder, M{a.mach:g}, {a.ni}x{a.nj}, "
                               "MUSCL minmod per-variable, HLLC, SSP-RK3 CFL 0.5, local dt"
                               + (", FP64 solver" if a.precision == "fp64" else ""),
                   "cells": cells, "grid": [a.ni, a.nj], "scheme": "ssprk3", "reconstruction": "perVariable/minmod",
                   "flux": "hllc", "parallelism": "single GPU",
                   "untimed_steps_before_warmup": PREWARM_STEPS + a.steps},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "tile_kernel (explicit stage)", "bytes_per_cell_update": width * BYTES_PER_CU_RK3,
                     "us_per_launch": ms_launch * 1e3},
        "cpu_baseline": cpu,
        "drag": cd,
        "converged_solves": solves,
    }
    print(json.dumps(line), flush=True)
    gpu.destroy()


def converged_sweep(ctx):
    """The metric's solves/s half on the batched Mach sweep (the seeded
    nose at Mach 1.2 ... 5.0, SSP-RK3, FP32): one batch, launch plan from
    the autotuner, solved from the free stream until every case has
    converged (the SolvePlan stopping rule), wall time from the first
    step to the last check."""
    import numpy as np

    from bowshock_amd import population as P

    t0 = time.time()
    cases = P.build_grids(P.sweep_cases(), workers=1)
    S = _setup_cases(cases, implicit=False)
    cells = np.array([c.cells for c in S["ok"]])
    batch = _CaseGroups(ctx, S)
    batch.setUniform(S["frees"])
    tuned = [dataclasses.asdict(t) for t in batch.autotune(steps=24, implicit=False)]
    batch.setUniform(S["frees"])
    t = time.perf_counter()
    results = batch.solve(S["p_inf"], S["q_inf"], S["plan"])
    wall = time.perf_counter() - t
    batch.destroy()
    steps = np.array([r.steps for r in results])
    converged = sum(r.converged for r in results)
    cu = float((cells * steps).sum())
    out = {"workload": "batched Mach sweep 1.2:0.1:5.0 of one seeded nose (mulberry32, seed 20261020), SSP-RK3 "
                       "FP32, free-stream start, to the SolvePlan stopping rule",
           "cases": len(cases), "accepted": len(S["ok"]), "converged": converged, "seconds": wall,
           "solves_per_s": converged / wall, "cell_updates_per_s": cu / wall,
           "hbm_roofline_frac": BYTES_PER_CU_RK3 * cu / wall / HBM_PEAK,
           "median_steps": int(np.median(steps)), "launch_tuning": tuned, "setup_seconds": time.time() - t0 - wall}
    print(f"converged sweep: {converged}/{len(S['ok'])} in {wall:.3f}s = {converged / wall:.1f} solves/s, "
          f"{cu / wall:.3e} CU/s, tuning {tuned}", file=sys.stderr, flush=True)
    return out


WARM_SAMPLE_ROCKETS = 120
WARM_SAMPLE_TRAIN_STEPS = 1500


def warm_start_sample(a):
    """The metric's warm-start half on a reduced population: the
    warmstart workload at 120 rockets (480 cases; the network trained for
    1,500 Adam steps on the converged fields of the training rockets), its
    held-out quarter solved from the free stream and from the warm start,
    as FP64 batches to a density residual of 1e-6, where the drags of the
    two starts can be compared at the 1e-5 bar (FP32 batches stop at their
    residual floor, where they differ by more). The full population
    (2,000 rockets, 6,000 training fields) takes minutes: bench.py
    --workload warmstart."""
    b = argparse.Namespace(**vars(a))
    b.rockets, b.train_steps, b.cpu_seconds, b.precision = WARM_SAMPLE_ROCKETS, WARM_SAMPLE_TRAIN_STEPS, 0.0, "fp64"
    t0 = time.time()
    line = run_warmstart(b)
    c = line["config"]
    out = {"workload": c["workload"], "precision": "fp64 batches to a density residual of 1e-6",
           "held_out_cases": c["held_out_cases"], "training_fields": c["training_fields"],
           "train_steps": c["train_steps"], "solves_per_s_plain": c["solves_per_s_plain"],
           "solves_per_s_warm": c["solves_per_s_warm"], "speedup": c["speedup"],
           "median_steps_plain": c["median_steps_plain"], "median_steps_warm": c["median_steps_warm"],
           "max_rel_cd_change_warm_vs_plain": c["max_rel_cd_change_warm_vs_plain"],
           "speedup_coarser_grid_first_from_network": c["sequenced_warm"]["speedup"],
           "held_out_field_rel_l2": c["held_out_field_rel_l2"], "seconds": time.time() - t0}
    print(f"warm-start sample: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


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


def _setup_cases(cases, implicit, precision="fp32"):
    """Scheme, stopping plan, free streams, boundaries and configs of the
    accepted cases. FP64 batches stop on the residual target itself (1e-6,
    absolute) with the drag settled; FP32 ones also accept their residual
    floor."""
    from bowshock_amd import grids
    from bowshock_amd import population as P
    from bowshock_amd.batch import SolvePlan
    from bowshock_amd.engine import Euler2DConfig

    ok = [c for c in cases if c.grid is not None]
    if not implicit:
        scheme = dict(flux="hllc", limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5)
        plan = SolvePlan(implicit=False, chunk=240, max_steps=20000)
    else:
        scheme = dict(flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled", timeIntegrator="ssprk3",
                      cfl=0.8)
        plan = SolvePlan(implicit=True, chunk=60, max_steps=6000, cfl=10.0, cfl_start=1.0, sweeps=2)
    if precision == "fp64":
        plan.residual, plan.drop, plan.stall = 1e-6, 0.0, 10 ** 9
        plan.max_steps = 12000 if implicit else 40000
    out = dict(ok=ok, scheme=scheme, plan=plan, frees=[], p_inf=[], q_inf=[], bcs=[], configs=[], schemes=[],
               precision=precision)
    for c in ok:
        free, p, q = grids.free_stream_at(c.mach, 1.4)
        bcs = P.boundaries(c, free)
        # The blunt tips' captured bow shocks take Rusanov's flux: with HLLC
        # the strong normal shock ahead of the nose goes unstable on the
        # refined C grids (the carbuncle) and the solve never settles.
        sch = scheme if c.topology == "H" else dict(scheme, flux="rusanov")
        out["frees"].append(free)
        out["p_inf"].append(p)
        out["q_inf"].append(q)
        out["bcs"].append(bcs)
        out["schemes"].append(sch)
        out["configs"].append(Euler2DConfig(grid=c.grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4,
                                            timeStepping="local", **sch))
    return out


class _CaseGroups:
    """The accepted cases in one batch per scheme (the H grids with HLLC,
    the blunt tips' C grids with Rusanov's flux), solved one batch after
    the other; indices are the cases' own."""

    def __init__(self, ctx, S):
        from bowshock_amd.batch import GpuBatch2D

        by_flux = {}
        for k, cfg in enumerate(S["configs"]):
            by_flux.setdefault(cfg.flux, []).append(k)
        self.groups = [(ks, GpuBatch2D(ctx, [S["configs"][k] for k in ks], S["precision"])) for ks in by_flux.values()]
        self.where = {k: (g, local) for g, (ks, _) in enumerate(self.groups) for local, k in enumerate(ks)}
        self.n = len(S["configs"])

    def setUniform(self, frees):
        for ks, b in self.groups:
            b.setUniform([frees[k] for k in ks])

    def densityResidualNorms(self):
        import numpy as np

        out = np.empty(self.n)
        for ks, b in self.groups:
            out[ks] = b.densityResidualNorms()
        return out

    def setCellPrimitives(self, k, prim):
        g, local = self.where[k]
        self.groups[g][1].setCellPrimitives(local, prim)

    def setCoarse(self, coarse, scale, shape, which=None):
        """Coarse fields for the cases `which` selects (all by default); a
        batch none of whose cases is selected is left as it is."""
        import numpy as np

        for ks, b in self.groups:
            pick = np.ones(len(ks), dtype=bool) if which is None else np.asarray(which)[ks]
            if pick.all():
                b.setCoarse(coarse[ks], scale[ks], shape=shape)
            elif pick.any():
                raise ValueError("a warm start for part of a batch")

    def readState(self, k):
        g, local = self.where[k]
        return self.groups[g][1].readState(local)

    def autotune(self, steps, implicit):
        from bowshock_amd.batch import autotune

        return [autotune(b, steps=steps, implicit=implicit) for _, b in self.groups]

    def kernel_sample(self, steps, implicit, cells):
        """Per batch: `steps` steps with HIP-event timing of every pass (no
        graphs), the mean microseconds per pass of each kernel and the
        relaxation's (or stage's) fraction of the HBM roofline by the byte
        model of the single-grid workloads. Explicit batches (the small
        ones, whose passes take microseconds and whose means a few slow
        launches can multiply) keep the lowest of three samples. Advances
        the state."""
        out = []
        for ks, b in self.groups:
            us = {}
            for _ in range(1 if implicit else 3):
                b.enableTiming(True)
                if implicit:
                    b.iterateImplicit(steps, 10.0, 2)
                else:
                    b.iterate(steps)
                for k in ("stage", "implicitPrepare", "implicitRelax"):
                    v = 1e3 * b.kernelTiming(k)[0]
                    us[k] = min(us.get(k, v), v)
                b.enableTiming(False)
            n = float(cells[ks].sum())
            if implicit:
                frac = 3 * 58.0 * n / (us["implicitRelax"] * 1e-6) / HBM_PEAK
            else:
                frac = BYTES_PER_STAGE * n / (us["stage"] * 1e-6) / HBM_PEAK
            out.append({"cases": len(ks), "cells": int(n), "us_per_pass": {k: v for k, v in us.items() if v > 0},
                        ("relaxation" if implicit else "stage") + "_roofline_frac": frac})
        return out

    def solve(self, p_inf, q_inf, plan, reference=None, cfl_start=None):
        import numpy as np

        from bowshock_amd.batch import solve

        results = [None] * self.n
        t0 = time.perf_counter()
        last = [t0]

        def progress(steps, active):
            # A line every 20 s of a long solve.
            now = time.perf_counter()
            if now - last[0] > 20:
                last[0] = now
                print(f"  {now - t0:.0f}s: step {steps}, {active} cases running", file=sys.stderr, flush=True)

        for ks, b in self.groups:
            offset = time.perf_counter() - t0
            pick = (lambda v: None if v is None else np.broadcast_to(np.asarray(v, dtype=np.float64), (self.n,))[ks])
            res = solve(b, np.asarray(p_inf)[ks], np.asarray(q_inf)[ks], 0.5, plan, reference=pick(reference),
                        cfl_start=pick(cfl_start), progress=progress)
            for k, r in zip(ks, res):
                r.seconds += offset
                results[k] = r
        return results

    def solve_mixed(self, first, frees, p_inf, q_inf, plan32, plan64):
        """`first` (FP32 batches of the same cases) to its stopping rule,
        then these FP64 batches from its states to plan64's target
        (batch.solve_mixed), one flux's batches after the other; steps and
        seconds count both phases."""
        import numpy as np

        from bowshock_amd.batch import solve_mixed

        results = [None] * self.n
        t0 = time.perf_counter()
        last = [t0]

        def progress(steps, active):
            now = time.perf_counter()
            if now - last[0] > 20:
                last[0] = now
                print(f"  {now - t0:.0f}s: step {steps}, {active} cases running", file=sys.stderr, flush=True)

        for (ks, b64), (ks32, b32) in zip(self.groups, first.groups):
            assert ks == ks32
            offset = time.perf_counter() - t0
            res = solve_mixed(b32, b64, [frees[k] for k in ks], np.asarray(p_inf)[ks], np.asarray(q_inf)[ks], 0.5,
                              plan32, plan64, progress=progress)
            for k, r in zip(ks, res):
                r.seconds += offset
                results[k] = r
        return results

    def destroy(self):
        for _, b in self.groups:
            b.destroy()


def _cpu_baseline(S, implicit, seconds, cu_per_solve):
    """The oracle on every host core, one case per worker, each for a
    bounded number of steps; solves/s from the GPU run's cell-updates per
    converged solve."""
    import numpy as np
    from concurrent.futures import ProcessPoolExecutor

    ok = S["ok"]
    workers = min(16, os.cpu_count() or 1)
    per_cu = 1.0 / 1e6 if implicit else 1.0 / 6e5   # rough oracle seconds per cell-update, one core
    pick = np.linspace(0, len(ok) - 1, workers).astype(int)
    jobs = []
    for k in pick:
        st = max(2, int(seconds / (per_cu * ok[k].cells)))
        jobs.append((ok[k].grid, S["bcs"][k], S["frees"][k], S["schemes"][k], implicit, st))
    t = time.perf_counter()
    with ProcessPoolExecutor(workers) as ex:
        out = list(ex.map(_cpu_batch_job, jobs))
    cpu_wall = time.perf_counter() - t
    cpu_cu = sum(o[0] for o in out)
    rate = cpu_cu / cpu_wall
    print(f"cpu baseline {rate:.3e} CU/s on {workers} cores", file=sys.stderr, flush=True)
    return {"value": rate / cu_per_solve, "unit": "solves/s", "cores": workers, "kind": "port", "host": host_info(),
            "sample": f"FP64 C oracle, {workers} processes, one case each ({len(jobs)} cases spread over the "
                      f"batch), {cpu_cu:.3e} cell-updates in {cpu_wall:.1f}s = {rate:.3e} CU/s; solves/s = "
                      f"that rate / the GPU run's {cu_per_solve:.3e} cell-updates per converged solve"}


def _outcomes(results):
    reasons = {}
    for r in results:
        reasons[r.reason] = reasons.get(r.reason, 0) + 1
    return reasons


def _build_grids_logged(P, cases, chunk=1000):
    """The population's grids in chunks, a progress line after each (the C
    grids of the blunt tips take about a second each)."""
    t0 = time.time()
    for k in range(0, len(cases), chunk):
        P.build_grids(cases[k:k + chunk])
        print(f"  grids {min(len(cases), k + chunk)}/{len(cases)} in {time.time() - t0:.0f}s", file=sys.stderr,
              flush=True)


def _fp64_sample(ctx, S, results, plan, count=16):
    """The FP32 results' drag error against FP64 on a sample of the
    converged cases spread over the batch: each case solved again by the
    FP64 solver (pinned to the oracle at 1e-12) for the same steps."""
    import numpy as np

    from bowshock_amd.engine import GpuEuler2D

    conv = [k for k, r in enumerate(results) if r.converged]
    if not conv:
        return None
    pick = sorted({conv[int(t)] for t in np.linspace(0, len(conv) - 1, min(count, len(conv)))})
    diffs = []
    for k in pick:
        g = GpuEuler2D.create(ctx, S["configs"][k], "fp64")
        g.setCellPrimitives(np.array(S["frees"][k]))
        steps = results[k].steps
        if plan.implicit:
            g.setImplicitCfl(plan.cfl_start)
            g.iterateImplicit(steps, cfl=plan.cfl, sweeps=plan.sweeps, omega=plan.omega, growth=plan.growth)
        else:
            g.iterate(steps)
        cd64 = g.forebodyDrag(S["p_inf"][k], S["q_inf"][k], 0.5)
        diffs.append(abs(results[k].cd / cd64 - 1))
        g.destroy()
    d = np.array(diffs)
    print(f"FP32 vs FP64 drag on {len(d)} sampled cases: median {np.median(d):.2e}, max {d.max():.2e}",
          file=sys.stderr, flush=True)
    return {"cases": len(d), "median_rel_cd_diff": float(np.median(d)), "max_rel_cd_diff": float(d.max()),
            "method": "FP64 solver, same configuration and step count; its implicit steps relax with the "
                      "linearized term (same steady state)"}


def run_batch(a):
    import numpy as np

    from bowshock_amd import population as P
    from bowshock_amd.engine import Context

    sweep = a.workload == "sweep"
    cases = P.sweep_cases() if sweep else P.population_cases(a.rockets)
    t0 = time.time()
    _build_grids_logged(P, cases)
    t_grids = time.time() - t0
    refusals = P.refusal_counts(cases)
    print(f"{len(cases)} cases, grids in {t_grids:.1f}s: {refusals}", file=sys.stderr, flush=True)
    mixed = a.precision == "mixed"
    if mixed and sweep:
        sys.exit("--precision mixed applies to the population workload")
    S = _setup_cases(cases, implicit=not sweep, precision="fp64" if mixed else a.precision)
    S32 = _setup_cases(cases, implicit=True, precision="fp32") if mixed else None
    ok, plan = S["ok"], S["plan"]
    cells = np.array([c.cells for c in ok])
    ctx = Context(0)
    t0 = time.time()
    batch = _CaseGroups(ctx, S)
    first = _CaseGroups(ctx, S32) if mixed else None
    t_create = time.time() - t0
    batch.setUniform(S["frees"])
    tuned = [dataclasses.asdict(t) for t in batch.autotune(steps=24 if sweep else 12, implicit=not sweep)]
    if mixed:
        first.setUniform(S32["frees"])
        tuned += [dataclasses.asdict(t) for t in first.autotune(steps=12, implicit=True)]
    if a.persistent != "auto" and a.precision == "fp32":
        for t, (_, b) in zip(tuned, batch.groups):
            b.setPersistent(int(a.persistent))
            t["persistent"] = int(a.persistent)
    kernels = batch.kernel_sample(24 if sweep else 12, not sweep, cells)
    print(f"kernel passes (HIP events, no graphs): {kernels}", file=sys.stderr, flush=True)
    batch.setUniform(S["frees"])
    print(f"{len(batch.groups)} batch(es) of {[len(g[0]) for g in batch.groups]} cases ({cells.sum():.3e} cells) "
          f"created in {t_create:.1f}s; launch tuning {tuned}", file=sys.stderr, flush=True)
    if mixed:
        first.setUniform(S32["frees"])
    t = time.perf_counter()
    if mixed:
        results = batch.solve_mixed(first, S["frees"], S["p_inf"], S["q_inf"], S32["plan"], plan)
    else:
        results = batch.solve(S["p_inf"], S["q_inf"], plan)
    wall = time.perf_counter() - t
    if mixed:
        first.destroy()
    steps = np.array([r.steps for r in results])
    converged = sum(r.converged for r in results)
    cu = float((cells * steps).sum())
    solves = converged / wall
    bpc = BYTES_PER_CU_RK3 if sweep else bytes_per_cu_implicit(plan.sweeps)
    achieved = bpc * cu / wall
    reasons = _outcomes(results)
    done_t = np.sort([r.seconds for r in results if r.converged])
    pct = {f"p{q}": float(np.percentile(done_t, q)) for q in (50, 90, 99)} if len(done_t) else {}
    print(f"{converged}/{len(ok)} converged in {wall:.2f}s: {solves:.2f} solves/s, {cu / wall:.3e} CU/s; "
          f"{reasons}; steps median {int(np.median(steps))}, p99 {int(np.percentile(steps, 99))}; "
          f"converged by (s) {pct}", file=sys.stderr, flush=True)
    fp64 = _fp64_sample(ctx, S, results, plan) if a.precision == "fp32" and not a.no_fp64_sample else None
    below = sum(r.converged and r.residual <= 1e-6 for r in results)
    print(f"{below}/{len(ok)} cases at a density residual <= 1e-6", file=sys.stderr, flush=True)
    cpu = _cpu_baseline(S, not sweep, a.cpu_seconds, cu / max(1, converged)) if a.cpu_seconds > 0 else None
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
        "dtype": "fp32 then fp64" if mixed else a.precision,
        "data": "synthetic: seeded bodies (mulberry32, seed 20261020), generated H grids (sharp tips) and C grids (blunt tips), free-stream start",
        "config": {"workload": ("Mach sweep 1.2:0.1:5.0 of one seeded nose on a 4-caliber cylinder, SSP-RK3"
                                if sweep else f"population of {a.rockets} seeded rockets x M 1.5/2/3/4, "
                                              "LU-SGS red-black 2 sweeps, CFL 1 -> 10 x1.2"
                                              + (", FP32 batches to their residual floor, then FP64 batches "
                                                 "from their states to a density residual of 1e-6"
                                                 if mixed else "")),
                   "cases": len(cases), "accepted": len(ok), "refusals": refusals, "converged": converged,
                   "outcomes": reasons, "cells": int(cells.sum()), "cell_updates": cu,
                   "cell_updates_per_s": cu / wall, "grid_seconds": t_grids, "create_seconds": t_create,
                   "launch_tuning": tuned, "kernel_passes": kernels, "converged_by_seconds": pct,
                   "residual_at_most_1e-6": below,
                   "steps_percentiles": {f"p{q}": float(np.percentile(steps, q)) for q in (50, 90, 99, 100)},
                   "batches": [len(g[0]) for g in batch.groups],
                   "parallelism": "single GPU, one batch per flux (H grids HLLC, C grids Rusanov)"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "whole batched step (all kernels)",
                     "bytes_per_cell_update": bpc},
        "cpu_baseline": cpu,
        "drag_range": [min(cds), max(cds)] if cds else None,
        "fp32_vs_fp64": fp64,
    }
    print(json.dumps(line), flush=True)
    batch.destroy()


def run_warmstart(a):
    """Population solved from the free stream, the network trained on the
    converged fields of the training rockets, then the held-out rockets'
    cases solved three times: from the free stream, from the warm start
    everywhere, and from the warm start on the blunt noses' C grids only
    (inference, interpolation and the fallback check inside the timing)."""
    import numpy as np

    from bowshock_amd import population as P
    from bowshock_amd import warmstart as W
    from bowshock_amd.engine import Context

    cases = P.population_cases(a.rockets)
    t0 = time.time()
    _build_grids_logged(P, cases)
    print(f"{len(cases)} cases, grids in {time.time() - t0:.1f}s: {P.refusal_counts(cases)}", file=sys.stderr,
          flush=True)
    accepted = [c for c in cases if c.grid is not None]
    held = [c for c in accepted if c.rocket.index % 4 == 3]
    train_cases = [c for c in accepted if c.rocket.index % 4 != 3]
    ctx = Context(0)

    def features(cs):
        return np.array([np.append(c.features(), c.cylinder) for c in cs])

    def run(cs, starts=None, only=None):
        S = _setup_cases(cs, implicit=True, precision=a.precision)
        batch = _CaseGroups(ctx, S)
        t = time.perf_counter()
        batch.setUniform(S["frees"])
        free_norm = batch.densityResidualNorms()
        used = np.zeros(len(cs), dtype=bool)
        if starts is not None:
            # Warm start for the cases `only` selects (all by default),
            # interpolated to their cells on the GPU in one call, kept where
            # its density residual beats the free stream's.
            only_ = np.ones(len(cs), dtype=bool) if only is None else np.asarray(only)
            targets, scale = starts()
            batch.setCoarse(targets, scale, (W.COARSE_NJ, W.COARSE_NI), which=only_)
            warm_norm = batch.densityResidualNorms()
            used = (warm_norm <= free_norm) & only_
            for k in np.nonzero(~used & only_)[0]:
                batch.setCellPrimitives(k, np.broadcast_to(np.array(S["frees"][k]), (cs[k].cells, 4)))
        # A case started from a predicted field starts at the final CFL
        # number, as the reference's driver does for a given start.
        plan = S["plan"]
        cfl0 = np.where(used, plan.cfl, plan.cfl_start)
        results = batch.solve(S["p_inf"], S["q_inf"], plan, reference=free_norm, cfl_start=cfl0)
        wall = time.perf_counter() - t
        return S, batch, results, wall, used

    def run_sequenced(cs, coarse_cs, starts_fn):
        """The cases with a coarser grid: solved there (to a density
        residual of 1e-5) from the free stream or the network's prediction,
        then each fine case started from its coarse solution at the final
        CFL number and solved to the plan's target; the others as run()."""
        seq = [k for k, cc in enumerate(coarse_cs) if cc is not None]
        rest = [k for k, cc in enumerate(coarse_cs) if cc is None]
        out = [None] * len(cs)
        t0 = time.perf_counter()
        Sc = _setup_cases([coarse_cs[k] for k in seq], implicit=True, precision=a.precision)
        Sc["plan"].residual = 1e-5
        gc = _CaseGroups(ctx, Sc)
        gc.setUniform(Sc["frees"])
        free_c = gc.densityResidualNorms()
        used = np.zeros(len(seq), dtype=bool)
        if starts_fn is not None:
            targets, scale = starts_fn()
            gc.setCoarse(targets[seq], scale[seq], (W.COARSE_NJ, W.COARSE_NI))
            warm_c = gc.densityResidualNorms()
            used = warm_c <= free_c
            for j in np.nonzero(~used)[0]:
                gc.setCellPrimitives(j, np.broadcast_to(np.array(Sc["frees"][j]), (Sc["ok"][j].cells, 4)))
        plan_c = Sc["plan"]
        cfl0 = np.where(used, plan_c.cfl, plan_c.cfl_start)
        res_c = gc.solve(Sc["p_inf"], Sc["q_inf"], plan_c, reference=free_c, cfl_start=cfl0)
        t_coarse = time.perf_counter() - t0
        Sf = _setup_cases([cs[k] for k in seq], implicit=True, precision=a.precision)
        gf = _CaseGroups(ctx, Sf)
        free_f = np.empty(len(seq))
        gf.setUniform(Sf["frees"])
        free_f[:] = gf.densityResidualNorms()
        for (ks, fb), (kc, cb) in zip(gf.groups, gc.groups):
            assert ks == kc
            fb.setFromCoarser(cb)
        gc.destroy()
        plan_f = Sf["plan"]
        res_f = gf.solve(Sf["p_inf"], Sf["q_inf"], plan_f, reference=free_f,
                         cfl_start=np.full(len(seq), plan_f.cfl))
        gf.destroy()
        for j, k in enumerate(seq):
            out[k] = res_f[j]
        if rest:
            _, b2, res_r, _, _ = run([cs[k] for k in rest], None if starts_fn is None else
                                     (lambda: tuple(v[rest] for v in starts_fn())))
            b2.destroy()
            for j, k in enumerate(rest):
                out[k] = res_r[j]
        wall = time.perf_counter() - t0
        cells_f = np.array([cs[k].cells for k in seq])
        cells_c = np.array([coarse_cs[k].cells for k in seq])
        cu = float((cells_c * np.array([r.steps for r in res_c])).sum() +
                   (cells_f * np.array([r.steps for r in res_f])).sum())
        conv = sum(r.converged for r in out)
        info = {"solves_per_s": conv / wall, "seconds": wall, "coarse_seconds": t_coarse, "converged": conv,
                "sequenced_cases": len(seq), "coarse_converged": int(sum(r.converged for r in res_c)),
                "median_coarse_steps": int(np.median([r.steps for r in res_c])) if seq else 0,
                "median_fine_steps": int(np.median([r.steps for r in res_f])) if seq else 0,
                "cell_updates": cu, "warm_start_used_on_coarse": int(used.sum())}
        print(f"sequenced ({'network' if starts_fn else 'free stream'} on the coarser grids): {json.dumps(info)}",
              file=sys.stderr, flush=True)
        return out, info

    def fields(batch, cs, results):
        import math as _m

        rows, targets = [], []
        for k, (c, r) in enumerate(zip(cs, results)):
            if not r.converged:
                continue
            prim = W.primitives_from_state(batch.readState(k), c.ni, c.nj)
            targets.append(W.to_target(prim, float(np.float32(c.mach * _m.sqrt(1.4)))))
            rows.append(k)
        return rows, np.array(targets)

    # 1. Training data: the training rockets from the free stream.
    S, batch, res_train, wall_train, _ = run(train_cases)
    rows, T = fields(batch, train_cases, res_train)
    batch.destroy()
    F = features([train_cases[k] for k in rows])
    print(f"training set: {len(rows)}/{len(train_cases)} converged in {wall_train:.1f}s", file=sys.stderr, flush=True)

    # 2. Training.
    t = time.perf_counter()
    ws = W.train(ctx, F, T, steps=a.train_steps, batch=256, lr=1e-3, seed=0,
                 log=lambda s, l: print(f"  step {s} loss {l:.4e}", file=sys.stderr, flush=True))
    t_train = time.perf_counter() - t
    ms_gemm, flops_gemm = ws.net.time_gemm(ws.net.depth, reps=50)
    tflops = flops_gemm / (ms_gemm * 1e-3) / 1e12
    print(f"trained in {t_train:.1f}s; output-layer GEMM {ms_gemm * 1e3:.1f} us = {tflops:.1f} TFLOP/s",
          file=sys.stderr, flush=True)

    # 3. Held-out rockets from the free stream.
    S_plain, batch, res_plain, wall_plain, _ = run(held)
    rows_h, T_h = fields(batch, held, res_plain)
    batch.destroy()

    # 4. Held-out rockets from the warm start.
    Fh = features(held)

    def starts():
        pred = ws.predict_targets(Fh)
        u = np.array([float(np.float32(c.mach * np.sqrt(1.4))) for c in held])
        if a.warm_snap > 0:
            pred = W.snap_to_free_stream(pred, u, a.warm_snap)
        return pred, np.stack([np.ones_like(u), u, u, u * u], axis=1)

    S_warm, batch, res_warm, wall_warm, used = run(held, starts)
    batch.destroy()
    # The warm start for the blunt noses only (their C grids), the sharp
    # noses (H grids) from the free stream.
    blunt = np.array([c.topology == "C" for c in held])
    _, batch, res_blunt, wall_blunt, used_blunt = run(held, starts, only=blunt)
    batch.destroy()
    steps_blunt = np.array([r.steps for r in res_blunt])
    conv_blunt = sum(r.converged for r in res_blunt)

    # 5. Held-out rockets solved on the grids half as fine first, from the
    # network's prediction or the free stream, then started from those
    # solutions (solveTransonic's coarse-to-fine start, transonic-flow.ts:
    # 157-168) at the final CFL number.
    t = time.perf_counter()
    coarse_cs = P.coarser_cases(held)
    t_coarse_grids = time.perf_counter() - t
    seq_plain = run_sequenced(held, coarse_cs, None)
    seq_warm = run_sequenced(held, coarse_cs, starts)
    def cd_change(results):
        both = [k for k in range(len(held)) if res_plain[k].converged and results[k].converged]
        return max((abs(results[k].cd - res_plain[k].cd) / abs(res_plain[k].cd) for k in both), default=float("nan"))

    pred_h = ws.predict_targets(Fh[rows_h])
    field_err = float(np.linalg.norm(pred_h - T_h) / np.linalg.norm(T_h))

    conv_plain = sum(r.converged for r in res_plain)
    conv_warm = sum(r.converged for r in res_warm)
    sp, sw = conv_plain / wall_plain, conv_warm / wall_warm
    steps_plain = np.array([r.steps for r in res_plain])
    steps_warm = np.array([r.steps for r in res_warm])
    both = [k for k in range(len(held)) if res_plain[k].converged and res_warm[k].converged]
    dcd = max((abs(res_warm[k].cd - res_plain[k].cd) / abs(res_plain[k].cd) for k in both), default=float("nan"))
    cells = np.array([c.cells for c in held])
    cu_plain = float((cells * steps_plain).sum())
    print(f"held-out: plain {conv_plain}/{len(held)} in {wall_plain:.2f}s ({sp:.2f}/s), warm {conv_warm}/{len(held)} "
          f"in {wall_warm:.2f}s ({sw:.2f}/s); median steps {int(np.median(steps_plain))} -> "
          f"{int(np.median(steps_warm))}; warm start used for {int(used.sum())}; field error {field_err:.3f}; "
          f"max CD change {dcd:.2e}", file=sys.stderr, flush=True)
    for m in sorted({c.mach for c in held}):
        ks = [k for k, c in enumerate(held) if c.mach == m]
        print(f"  M{m:g}: {len(ks)} cases, median steps plain {int(np.median(steps_plain[ks]))}, warm "
              f"{int(np.median(steps_warm[ks]))}", file=sys.stderr, flush=True)
    # Where the work goes: cell-steps by grid topology, and the cases whose
    # warm start cost more steps than the free stream.
    work = {}
    for topo in ("H", "C"):
        ks = np.array([k for k, c in enumerate(held) if c.topology == topo], dtype=int)
        if len(ks):
            work[topo] = {"cases": int(len(ks)), "cell_steps_plain": float((cells[ks] * steps_plain[ks]).sum()),
                          "cell_steps_warm": float((cells[ks] * steps_warm[ks]).sum()),
                          "cell_steps_warm_blunt_only": float((cells[ks] * steps_blunt[ks]).sum()),
                          "warm_slower": int((steps_warm[ks] > steps_plain[ks]).sum())}
            print(f"  {topo} grids: {work[topo]}", file=sys.stderr, flush=True)
    sb = conv_blunt / wall_blunt
    both_b = [k for k in range(len(held)) if res_plain[k].converged and res_blunt[k].converged]
    dcd_b = max((abs(res_blunt[k].cd - res_plain[k].cd) / abs(res_plain[k].cd) for k in both_b), default=float("nan"))
    print(f"warm start on the blunt noses only: {conv_blunt}/{len(held)} in {wall_blunt:.2f}s ({sb:.2f}/s), "
          f"median steps {int(np.median(steps_blunt))}, used for {int(used_blunt.sum())}, max CD change {dcd_b:.2e}",
          file=sys.stderr, flush=True)
    cpu = (_cpu_baseline(S_plain, True, a.cpu_seconds, cu_plain / max(1, conv_plain)) if a.cpu_seconds > 0
           else None)
    line = {
        "metric": "Euler cell-updates/s, %HBM roofline; converged CD solves/s warm-start vs plain",
        "value": sb,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": int(steps_blunt.max()),
        "warmup": 0,
        "ms_per_step": 1e3 * wall_blunt / max(1, int(steps_blunt.max())),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 solver, bf16 network (fp32 accumulation)",
        "data": "synthetic: seeded rockets (mulberry32, seed 20261020); the network trained on this run's converged "
                "fields of the training rockets",
        "config": {"workload": f"held-out rockets (index % 4 == 3) of a {a.rockets}-rocket population x M 1.5/2/3/4, "
                               "LU-SGS red-black 2 sweeps; warm start from an MLP 9 -> 512 x 4 -> 4 x 80 x 40 for "
                               "the blunt noses (C grids), the sharp noses (H grids) from the free stream",
                   "held_out_cases": len(held), "training_fields": len(rows), "train_steps": a.train_steps,
                   "train_seconds": t_train, "solves_per_s_plain": sp, "solves_per_s_warm_all": sw,
                   "speedup_all": sw / sp if sp > 0 else None, "median_steps_plain": int(np.median(steps_plain)),
                   "median_steps_warm_all": int(np.median(steps_warm)), "warm_start_used_all": int(used.sum()),
                   "held_out_field_rel_l2": field_err, "max_rel_cd_change_warm_all_vs_plain": dcd,
                   "outcomes_plain": _outcomes(res_plain), "outcomes_warm_all": _outcomes(res_warm),
                   "work_by_topology": work, "solves_per_s_warm": sb,
                   "speedup": sb / sp if sp > 0 else None, "warm_start_used": int(used_blunt.sum()),
                   "median_steps_warm": int(np.median(steps_blunt)), "outcomes_warm": _outcomes(res_blunt),
                   "max_rel_cd_change_warm_vs_plain": dcd_b,
                   "coarser_grid_seconds": t_coarse_grids,
                   "sequenced_free_stream": dict(seq_plain[1], speedup=seq_plain[1]["solves_per_s"] / sp if sp else None,
                                                 max_rel_cd_change_vs_plain=cd_change(seq_plain[0])),
                   "sequenced_warm": dict(seq_warm[1], speedup=seq_warm[1]["solves_per_s"] / sp if sp else None,
                                          max_rel_cd_change_vs_plain=cd_change(seq_warm[0])),
                   "parallelism": "single GPU, one batch per flux (H grids HLLC, C grids Rusanov)"},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": tflops / 2500.0, "traffic": None,
                     "kernel": "gemm_bf16_kernel, output layer 256 x 512 x 12800 (forward)",
                     "us_per_launch": ms_gemm * 1e3},
        "cpu_baseline": cpu,
    }
    return line


def run_fitted(a):
    """The fitted bow shock on the GPU (bs_set_fitted_shock and the moving
    line grid): the hemisphere-cylinder (radius 0.5, 4 calibers of
    cylinder) at Mach 3 on the blunt-body C grid, 20 cells across refined
    by 4; captured by the FP64 solver, the fitted start on its lines, then
    moves of two steps in blocks of 500 until the shock speeds fall below
    1e-9 of the sound speed (FP64) or stop falling (FP32). The rate counts
    the fitted loop's cell-updates over its wall time."""
    import math

    import numpy as np

    from bowshock_amd import grids
    from bowshock_amd.blunt_grid import blunt_body_grid
    from bowshock_amd.engine import Context, Euler2DConfig, GpuEuler2D, gpu_primitives
    from bowshock_amd.fitted_start import fitted_start

    ctx = Context(0)
    mach, g = 3.0, 1.4
    body = grids.Body(grids.Nose("ellipsoid", 0.5, 0.5), 4.0)
    bg = blunt_body_grid(body, 20, mach, refinement=4)
    grid = bg.grid
    ni, nj = grid.ni, grid.nj
    free, p_inf, q_inf = grids.free_stream_at(mach, g)
    scheme = dict(geometry="axisymmetric", gamma=g, flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled",
                  timeIntegrator="ssprk2", timeStepping="local")
    bcs = {"iMin": {"kind": "wall"}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
           "jMax": {"kind": "inflow", "state": free}}
    t = time.perf_counter()
    cap = GpuEuler2D.create(ctx, Euler2DConfig(grid=grid, boundaries=bcs, cfl=0.8, **scheme), "fp64")
    cap.setCellPrimitives(np.array(free))
    r0 = cap.densityResidualNorm()
    steps_cap = 0
    while steps_cap < 40000:
        cap.iterate(2000)
        steps_cap += 2000
        if cap.densityResidualNorm() < 1e-8 * r0:
            break
    cd_cap = cap.forebodyDrag(p_inf, q_inf, 0.5)
    prim = gpu_primitives(cap.readState(), g)
    t_cap = time.perf_counter() - t
    print(f"captured start: {ni}x{nj}, {steps_cap} FP64 steps in {t_cap:.1f}s, CD {cd_cap:.6f}", file=sys.stderr,
          flush=True)
    st = fitted_start(bg.lines, grid, prim, 0.5)
    cap.destroy()
    out = {}
    for precision in ("fp32", "fp64"):
        fbcs = dict(bcs, jMax={"kind": "faceStates", "states": np.tile(np.array(free), (ni, 1))})
        fit = GpuEuler2D.create(ctx, Euler2DConfig(grid=st.grid, boundaries=fbcs, cfl=0.5, **scheme), precision)
        fit.setLineGrid(st.lines.origins_and_directions(), st.lines.fractions, "symmetry")
        fit.setHeights(st.heights)
        fit.setCellPrimitives(st.primitives)
        fit.setFittedShock(free)
        fit.iterateFitted(4, 2)   # graphs and first launches outside the timing
        speeds, moves = [], 0
        t = time.perf_counter()
        while moves < 20000:
            fit.iterateFitted(500, 2)
            moves += 500
            speeds.append(float(np.abs(fit.readShock()["speed"]).max()) / math.sqrt(g))
            if speeds[-1] < 1e-9 or (len(speeds) > 4 and speeds[-1] > 0.5 * min(speeds[:-4])):
                break
        wall = time.perf_counter() - t
        h = fit.readHeights()
        steps = 2 * moves
        row = {"moves": moves, "steps": steps, "seconds": wall, "cell_updates_per_s": ni * nj * steps / wall,
               "shock_speed": speeds[-1], "residual": fit.densityResidualNorm(),
               "cd": fit.forebodyDrag(p_inf, q_inf, 0.5), "standoff": float(h[0])}
        print(f"fitted {precision}: {json.dumps(row)}", file=sys.stderr, flush=True)
        out[precision] = row
        fit.destroy()
    line = {
        "metric": "Euler cell-updates/s, %HBM roofline; converged CD solves/s warm-start vs plain",
        "value": out["fp32"]["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": out["fp32"]["steps"],
        "warmup": 8,
        "ms_per_step": 1e3 * out["fp32"]["seconds"] / out["fp32"]["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 beside it)",
        "data": "synthetic: generated blunt-body C grid, captured FP64 start",
        "config": {"workload": f"fitted bow shock, hemisphere-cylinder M3, C grid {ni}x{nj}, SSP-RK2 moves of two "
                               "steps until the shock settles", "captured_steps": steps_cap,
                   "captured_seconds": t_cap, "captured_cd": cd_cap, "fitted": out, "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_implicit(a):
    """The 1280 x 640 cone at Mach 3 with implicit red-black LU-SGS steps
    (CFL ramp 1 -> 10 by 1.2, 2 sweeps, omega 1.5): cell-updates/s of
    whole implicit steps, and the relaxation kernel (three half-sweeps per
    launch) against the HBM roofline."""
    import numpy as np

    from bowshock_amd import grids
    from bowshock_amd.engine import Context, Euler2DConfig, GpuEuler2D

    body = grids.cone_cylinder(15.0, 4.0)
    bg = grids.body_grid_fixed(body, a.ni, a.nj, a.mach, 1.4)
    free, p_inf, q_inf = grids.free_stream_at(a.mach, 1.4)
    bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
           "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
    cfg = Euler2DConfig(grid=bg.grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4, flux="hllc",
                        limiter="vanAlbada", reconstruction="sharedScaled", timeIntegrator="ssprk3", cfl=0.8,
                        timeStepping="local")
    ctx = Context(0)
    gpu = GpuEuler2D.create(ctx, cfg, a.precision)
    gpu.setCellPrimitives(np.array(free))
    gpu.setImplicitCfl(1.0)
    steps = max(12, a.steps // 10)
    gpu.iterateImplicit(max(a.warmup // 2, 12), cfl=10.0, sweeps=2)
    t = time.perf_counter()
    gpu.iterateImplicit(steps, cfl=10.0, sweeps=2)
    elapsed = time.perf_counter() - t
    cells = a.ni * a.nj
    value = cells * steps / elapsed
    gpu.enableTiming(True)
    gpu.iterateImplicit(64, cfl=10.0, sweeps=2)
    timing = {k: gpu.kernelTiming(k) for k in ("implicitPrepare", "implicitRelax", "implicitUpdate")}
    gpu.enableTiming(False)
    ms_relax = timing["implicitRelax"][0]
    # Three half-sweeps of 58 B per cell each (half the residual and
    # diagonal, q, the other colour's dQ, face vectors, the dQ written).
    relax_bytes = 3 * 58.0 * cells
    achieved = relax_bytes / (ms_relax * 1e-3)
    print(f"{steps} implicit steps in {elapsed:.3f}s: {1e6 * elapsed / steps:.1f} us/step, {value:.3e} CU/s; "
          + ", ".join(f"{k} {v[0] * 1e3:.2f} us" for k, v in timing.items()), file=sys.stderr, flush=True)
    cpu = None
    if a.cpu_steps > 0:
        # The oracle's red-black LU-SGS (2 sweeps, the same CFL ramp) on one
        # core for a bounded number of steps of the same case.
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        orc = O.Euler2D(bg.grid, geometry="axisymmetric", gamma=1.4, flux="hllc", limiter="vanAlbada",
                        reconstruction="sharedScaled", time_integrator="ssprk3", cfl=0.8, time_stepping="local",
                        boundaries=bcs)
        orc.set_primitives(np.array(free))
        cpu_steps = max(2, a.cpu_steps // 4)
        cfl = 1.0
        t = time.perf_counter()
        for _ in range(cpu_steps):
            orc.iterate_implicit(1, cfl, 1.5, relaxation="redBlack", sweeps=2)
            cfl = min(10.0, cfl * 1.2)
        ce = time.perf_counter() - t
        cpu = {"value": cells * cpu_steps / ce, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"{cpu_steps} red-black LU-SGS steps (2 sweeps, CFL 1 -> 10 x1.2) of the same "
                         f"{a.ni}x{a.nj} case on one core, FP64 C oracle (gcc -O3), {ce:.1f}s", "host": host_info()}
        print(f"cpu baseline {cpu['value']:.3e} CU/s ({ce:.1f}s)", file=sys.stderr, flush=True)
    line = {
        "metric": "Euler cell-updates/s, %HBM roofline; converged CD solves/s warm-start vs plain",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": max(a.warmup // 2, 12),
        "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": a.precision, "data": "synthetic: generated cone-cylinder grid, free-stream start",
        "config": {"workload": f"single solve, 15 deg cone + 4 cal cylinder, M{a.mach:g}, {a.ni}x{a.nj}, "
                               "implicit red-black LU-SGS (2 sweeps, CFL 10), MUSCL van Albada shared-scaled, HLLC"
                               + (", FP64 solver (FP32 relaxation)" if a.precision == "fp64" else ""),
                   "cells": cells, "kernel_us": {k: v[0] * 1e3 for k, v in timing.items()},
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "relax_tile_kernel<3> (three red-black half-sweeps)", "bytes_per_cell": 174.0,
                     "bytes_model": "FP32 operands" + (" (a lower bound: the FP64 state and metrics are read "
                                                       "at twice the width)" if a.precision == "fp64" else ""),
                     "us_per_launch": ms_relax * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_verified(a):
    """The forebody drag of the 15 and 20 deg cones (4-caliber cylinders)
    with its measured discretisation error: the verified-drag driver on
    nested bodyGrid families (20 cells across, refinement 1 to 8) on the
    HIP engine."""
    import math

    from bowshock_amd import grids
    from bowshock_amd.engine import Context
    from bowshock_amd.verified_drag import verified_forebody_drag

    ctx = Context(0)
    rows = []
    t_all = time.perf_counter()
    for half_angle, mach in ((15.0, 2.0), (20.0, 3.0)):
        body = grids.Body(grids.Nose("conical", 0.5 / math.tan(math.radians(half_angle)), 0.5), 4.0)
        t = time.perf_counter()
        res = verified_forebody_drag(ctx, body, mach=mach, gamma=1.4, tolerance=1e-4, max_refinement=8)
        wall = time.perf_counter() - t
        est = res.result.estimate if res.result else None
        row = {"cone_deg": half_angle, "mach": mach, "kind": res.kind, "seconds": wall,
               "levels": [[r.refinement, r.ni, r.nj, r.iterations, r.drag] for r in res.reports],
               "estimate": est.kind if est else None, "value": est.value if est else None,
               "extrapolated": est.extrapolated if est else None, "order": est.order if est else None,
               "band": res.result.uncertainty if res.result else None,
               "verified": res.result.verified if res.result else False, "reason": res.reason or
               (est.reason if est else "")}
        print(json.dumps(row), file=sys.stderr, flush=True)
        rows.append(row)
    line = {"metric": "Euler cell-updates/s, %HBM roofline; converged CD solves/s warm-start vs plain",
            "value": len([r for r in rows if r["verified"]]) / (time.perf_counter() - t_all), "unit": "verified drags/s",
            "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32 solve, fp64 drag",
            "data": "synthetic: generated cone-cylinder grid families",
            "config": {"workload": "verified forebody drag, nested grids nj = 20..160", "cases": rows,
                       "parallelism": "single GPU"},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def run_ladder(a):
    """Single solves to convergence: the 160 x 80 cone at Mach 2 (FP64
    oracle beside it, same step count), then the 15 deg cone and the
    3-caliber tangent ogive (4-caliber cylinders) at Mach 1.5 and 3 on
    160 x 80, 640 x 320 and 1280 x 640, explicit SSP-RK3 and implicit
    red-black LU-SGS."""
    import numpy as np

    from bowshock_amd import grids
    from bowshock_amd.batch import SolvePlan, solve_single
    from bowshock_amd.engine import Context, Euler2DConfig, GpuEuler2D

    ctx = Context(0)
    rows = []

    def one(body_name, body, mach, ni, nj, implicit, plan=None, oracle_seconds=0.0):
        bg = grids.body_grid_fixed(body, ni, nj, mach, 1.4)
        free, p_inf, q_inf = grids.free_stream_at(mach, 1.4)
        bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
               "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
        if implicit:
            scheme = dict(flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled", timeIntegrator="ssprk3",
                          cfl=0.8)
            plan = plan or SolvePlan(implicit=True, chunk=60, max_steps=20000, cfl=10.0, cfl_start=1.0, sweeps=2)
        else:
            scheme = dict(flux="hllc", limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3",
                          cfl=0.5)
            plan = plan or SolvePlan(implicit=False, chunk=240, max_steps=120000)
        cfg = Euler2DConfig(grid=bg.grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4,
                            timeStepping="local", **scheme)
        gpu = GpuEuler2D.create(ctx, cfg)
        gpu.setCellPrimitives(np.array(free))
        ref = gpu.densityResidualNorm()
        t = time.perf_counter()
        r = solve_single(gpu, p_inf, q_inf, 0.5, plan, reference=ref)
        wall = time.perf_counter() - t
        label = f"lu-sgs {plan.sweeps} sweeps" if implicit else "ssprk3"
        row = {"body": body_name, "mach": mach, "grid": [ni, nj], "scheme": label,
               "steps": r.steps, "seconds": wall, "cell_updates_per_s": ni * nj * r.steps / wall, "cd": r.cd,
               "residual": r.residual, "outcome": r.reason}
        # The FP32 result's error against FP64: the FP64 solver (pinned to
        # the oracle at 1e-12) run for the same steps; its implicit steps
        # relax with the linearized neighbour term, which changes the path
        # but not the steady state.
        g64 = GpuEuler2D.create(ctx, cfg, "fp64")
        g64.setCellPrimitives(np.array(free))
        if implicit:
            g64.setImplicitCfl(plan.cfl_start)
            g64.iterateImplicit(r.steps, cfl=plan.cfl, sweeps=plan.sweeps, omega=plan.omega, growth=plan.growth)
        else:
            g64.iterate(r.steps)
        cd64 = g64.forebodyDrag(p_inf, q_inf, 0.5)
        row.update({"cd_fp64_same_steps": cd64, "cd_fp32_vs_fp64": abs(r.cd / cd64 - 1),
                    "residual_fp64": g64.densityResidualNorm()})
        g64.destroy()
        if oracle_seconds > 0:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as O

            orc = O.Euler2D(bg.grid, geometry="axisymmetric", gamma=1.4, flux=scheme["flux"],
                            limiter=scheme["limiter"], reconstruction=scheme["reconstruction"],
                            time_integrator="ssprk3", cfl=scheme["cfl"], time_stepping="local", boundaries=bcs)
            orc.set_primitives(np.array(free))
            t = time.perf_counter()
            orc.iterate(r.steps)
            ce = time.perf_counter() - t
            cd_o = orc.forebody_drag(p_inf, q_inf, 0.5)
            row.update({"oracle_cd_same_steps": cd_o, "cd_rel_diff": abs(r.cd / cd_o - 1), "oracle_seconds": ce,
                        "oracle_cell_updates_per_s": ni * nj * r.steps / ce,
                        "oracle_residual": orc.density_residual_norm()})
        print(json.dumps(row), file=sys.stderr, flush=True)
        rows.append(row)
        gpu.destroy()
        return row

    def converged_fp64(body_name, body, mach, ni, nj):
        from bowshock_amd.engine import AndersonMixing

        bg = grids.body_grid_fixed(body, ni, nj, mach, 1.4)
        free, p_inf, q_inf = grids.free_stream_at(mach, 1.4)
        bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
               "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
        cfg = Euler2DConfig(grid=bg.grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4, flux="hllc",
                            limiter="vanAlbada", reconstruction="sharedScaled", timeIntegrator="ssprk3", cfl=0.8,
                            timeStepping="local")
        out = {}
        for scheme in ("lu-sgs 2 sweeps fp64 + anderson", "ssprk3 fp64"):
            g = GpuEuler2D.create(ctx, cfg, "fp64")
            g.setCellPrimitives(np.array(free))
            t = time.perf_counter()
            steps = 0
            if scheme.startswith("lu-sgs"):
                g.setImplicitCfl(1.0)
                g.iterateImplicit(300, cfl=10.0, sweeps=2)
                steps = 300
                aa = AndersonMixing(ctx, 4 * ni * nj, 8)
                while steps < 20000 and g.densityResidualNorm() > 1e-6:
                    g.iterateAnderson(aa, 600, 10.0, 2)
                    steps += 600
                aa.destroy()
            else:
                while steps < 60000 and (steps == 0 or g.densityResidualNorm() > 1e-6):
                    g.iterate(2400)
                    steps += 2400
            wall = time.perf_counter() - t
            out[scheme] = {"steps": steps, "seconds": wall, "cell_updates_per_s": ni * nj * steps / wall,
                           "residual": g.densityResidualNorm(), "cd": g.forebodyDrag(p_inf, q_inf, 0.5)}
            g.destroy()
        imp, ex = out["lu-sgs 2 sweeps fp64 + anderson"], out["ssprk3 fp64"]
        row = {"body": body_name, "mach": mach, "grid": [ni, nj], "scheme": "lu-sgs 2 sweeps fp64 + anderson",
               "steps": imp["steps"], "seconds": imp["seconds"], "cell_updates_per_s": imp["cell_updates_per_s"],
               "cd": imp["cd"], "residual": imp["residual"],
               "outcome": "converged" if imp["residual"] <= 1e-6 else "step limit",
               "fp64_explicit": ex, "cd_vs_fp64_explicit": abs(imp["cd"] / ex["cd"] - 1)}
        print(json.dumps(row), file=sys.stderr, flush=True)
        rows.append(row)

    cone = grids.cone_cylinder(15.0, 4.0)
    ogive = grids.ogive_cylinder(3.0, 4.0)
    # The small configuration: residual 1e-6 or 20,000 steps.
    one("cone15", cone, 2.0, 160, 80, False,
        SolvePlan(implicit=False, chunk=240, max_steps=20000, residual=1e-6, drop=0.0, stall=10 ** 9, settle=0),
        oracle_seconds=1.0 if a.cpu_seconds > 0 else 0.0)
    for name, body in (("cone15", cone), ("ogive3", ogive)):
        for mach in (1.5, 3.0):
            for ni, nj in ((160, 80), (640, 320), (1280, 640)):
                for implicit in (False, True):
                    one(name, body, mach, ni, nj, implicit)
                if ni == 1280:
                    # Two red-black sweeps at CFL 10 settle into a limit
                    # cycle on this grid in FP32; three sweeps converge.
                    one(name, body, mach, ni, nj, True,
                        SolvePlan(implicit=True, chunk=60, max_steps=20000, cfl=10.0, cfl_start=1.0, sweeps=3))
                    # The FP64 solver with two sweeps as solveTransonic runs
                    # them: plain steps, then Anderson mixing (depth 8), to a
                    # density residual of 1e-6; against the FP64 explicit
                    # solve to the same residual.
                    converged_fp64(name, body, mach, ni, nj)
    head = [r for r in rows if r["grid"] == [1280, 640] and r["scheme"] == "ssprk3"]
    line = {
        "metric": "Euler cell-updates/s, %HBM roofline; converged CD solves/s warm-start vs plain",
        "value": float(np.mean([r["cell_updates_per_s"] for r in head])),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": int(sum(r["steps"] for r in rows)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: generated cone-cylinder and ogive-cylinder grids, free-stream starts",
        "config": {"workload": "single-solve ladder to convergence (value: mean over the 1280x640 explicit solves, "
                               "including the driver's residual and drag checks)", "solves": rows,
                   "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
