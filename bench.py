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
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--ni", type=int, default=1280)
    ap.add_argument("--nj", type=int, default=640)
    ap.add_argument("--mach", type=float, default=3.0)
    ap.add_argument("--cpu-steps", type=int, default=8, help="oracle steps timed for the CPU baseline (0: skip)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "stage_kernel_1280x640_pmc.txt"))
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

    # Kernel timing pass: HIP events around each stage-kernel launch.
    gpu.enableTiming(True)
    gpu.iterate(min(a.steps, 300))
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


if __name__ == "__main__":
    main()
