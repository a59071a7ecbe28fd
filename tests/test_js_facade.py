"""The Node facade (host/index.js over the N-API addon): host logic on the
CPU, and a solve on the GPU checked against the FP64 oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def test_host_logic_without_gpu():
    r = subprocess.run([NODE, os.path.join(ROOT, "host", "test", "host_logic.test.js")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "FAIL" not in r.stdout


@pytest.mark.gpu
def test_node_solve_matches_oracle(tmp_path):
    grid, free, p_inf, q_inf, bcs = cases.cone_case(48, 24, mach=2.0)
    config = dict(geometry="axisymmetric", gamma=1.4, flux="hllc", limiter="vanAlbada",
                  reconstruction="sharedScaled", timeIntegrator="ssprk3", cfl=0.5, timeStepping="local",
                  boundaries={"iMin": {"kind": "inflow"}, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
                              "jMax": {"kind": "inflow"}})
    case = dict(ni=grid.ni, nj=grid.nj, x=grid.x.tolist(), y=grid.y.tolist(), mach=2.0, config=config, steps=60,
                implicitSteps=20)
    path = tmp_path / "case.json"
    path.write_text(json.dumps(case))
    r = subprocess.run([NODE, os.path.join(ROOT, "host", "test", "gpu_solve.js"), str(path)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = json.loads(r.stdout.strip().splitlines()[-1])
    orc, _ = cases.make_pair(None, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc.set_primitives(np.array(free))
    orc.iterate(60)
    cd = orc.forebody_drag(p_inf, q_inf, 0.5)
    assert abs(out["drag"] / cd - 1) < 1e-5
    assert abs(out["gpuDrag"] - out["drag"]) < 1e-12
    # The FP64 solver through the facade: the oracle's drag within FP64
    # round-off of 60 steps (the FP64 parity bar of the Python tests).
    assert abs(out["drag64"] / cd - 1) < 1e-11
    assert out["state64"] == "Float64Array"
    assert "await the previous operation first" in out["overlap"]
    assert out["steps"] == 60
    assert out["nonPhysical"] == "NonPhysicalStateError"
    # The transonic adapter runs the FP64 solver (as the reference's runs
    # the two-float one), whose red-black relaxation is the linearized one.
    orc2, _ = cases.make_pair(None, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc2.set_primitives(np.array(free))
    cfl = 1.0
    for _ in range(20):
        orc2.iterate_implicit(1, cfl, 1.5, "redBlackLinear", 2)
        cfl = min(10.0, cfl * 1.2)
    assert abs(out["implicitDrag"] / orc2.forebody_drag(p_inf, q_inf, 0.5) - 1) < 1e-5


@pytest.mark.gpu
def test_node_fitted_shock_matches_the_python_facade(tmp_path, ctx):
    """The fitted shock and its moving grid through the Node facade
    (GpuLineGrid, iterateFitted, readShock) on the FP64 solver: the same
    heights and shock speeds as the Python facade, which the oracle's
    fitShock loop pins (test_gpu_shock_fitting.py)."""
    from test_gpu_shock_fitting import _fitted_pair, _line_cone

    lines, eta, h0, _ = _line_cone()
    case = dict(lines=lines.tolist(), fractions=eta.T.tolist(), heights=h0.tolist(), nj=eta.shape[0] - 1,
                mach=2.0, moves=30)
    path = tmp_path / "fitted.json"
    path.write_text(json.dumps(case))
    r = subprocess.run([NODE, os.path.join(ROOT, "host", "test", "gpu_fitted.js"), str(path)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = json.loads(r.stdout.strip().splitlines()[-1])
    grid, free, make, gpu = _fitted_pair(ctx, lines, eta, h0, "fp64", 2.0)
    gpu.setLineGrid(lines, eta, "fixed")
    gpu.setHeights(h0)
    gpu.setCellPrimitives(np.array(free))
    gpu.setFittedShock(free)
    gpu.iterateFitted(30, 1)
    assert np.abs(np.array(out["heights"]) - gpu.readHeights()).max() < 1e-13
    assert np.abs(np.array(out["speed"]) - gpu.readShock()["speed"]).max() < 1e-12
    assert out["steps"] == 30
    assert "fitted shock" in out["implicit"]


@pytest.mark.gpu
def test_node_batch_matches_the_python_facade(tmp_path, ctx):
    """GpuBatch2D and solveBatch through the Node facade: two H grids and
    a blunt-body C grid in one batch; the states after implicit steps
    equal the Python facade's batch bit for bit, and the stopping rule
    gives the same outcomes and drags."""
    from bowshock_amd import grids
    from bowshock_amd.batch import GpuBatch2D, SolvePlan, solve
    from bowshock_amd.blunt_grid import blunt_body_grid
    from bowshock_amd.engine import Euler2DConfig

    specs = []
    for ni, nj, mach in ((48, 24, 2.0), (40, 20, 3.0)):
        g = grids.body_grid_fixed(grids.cone_cylinder(15.0, 4.0), ni, nj, mach, 1.4).grid
        specs.append((g, mach, False))
    specs.append((blunt_body_grid(grids.Body(grids.Nose("ellipsoid", 0.5, 0.5), 4.0), 12, 2.0).grid, 2.0, True))
    from bowshock_amd import warmstart as W

    rng = np.random.default_rng(8)
    n_coarse = W.COARSE_NI * W.COARSE_NJ
    coarse = np.stack([np.concatenate([b + 0.2 * rng.standard_normal(n_coarse) for b in (1.0, 0.9, 0.0, 0.6)])
                       for _ in specs])
    u_inf = np.array([grids.free_stream_at(m, 1.4)[0][1] for _, m, _ in specs])
    scales = np.stack([np.ones_like(u_inf), u_inf, u_inf, u_inf * u_inf], axis=1)
    case = dict(steps=30, maxSteps=4800, coarse=coarse.ravel().tolist(), scales=scales.tolist(),
                cases=[dict(ni=g.ni, nj=g.nj, x=g.x.tolist(), y=g.y.tolist(), mach=m, axis=a) for g, m, a in specs])
    path = tmp_path / "batch.json"
    path.write_text(json.dumps(case))
    r = subprocess.run([NODE, os.path.join(ROOT, "host", "test", "gpu_batch.js"), str(path)], capture_output=True,
                       text=True, timeout=300)
    print(r.stderr[-2000:])
    assert r.returncode == 0
    out = json.loads(r.stdout.strip().splitlines()[-1])
    cfgs, frees, ps, qs = [], [], [], []
    for g, m, axis in specs:
        free, p, q = grids.free_stream_at(m, 1.4)
        inflow = {"kind": "inflow", "state": free}
        bcs = {"iMin": {"kind": "wall"} if axis else inflow, "iMax": {"kind": "outflow"}, "jMin": {"kind": "wall"},
               "jMax": inflow}
        cfgs.append(Euler2DConfig(grid=g, geometry="axisymmetric", gamma=1.4, flux="hllc", limiter="vanAlbada",
                                  reconstruction="sharedScaled", timeIntegrator="ssprk3", cfl=0.5,
                                  timeStepping="local", boundaries=bcs))
        frees.append(free)
        ps.append(p)
        qs.append(q)
    batch = GpuBatch2D(ctx, cfgs)
    batch.setUniform(frees)
    batch.setImplicitCfl(1.0)
    batch.iterateImplicit(30, 10.0, 2)
    for k in range(len(specs)):
        assert np.array_equal(np.array(out["rho"][k], dtype=np.float32), batch.readState(k).rho), k
    assert np.array_equal(np.array(out["norms"]), batch.densityResidualNorms())
    assert np.array_equal(np.array(out["drags"]), batch.forebodyDrags(np.array(ps), np.array(qs), 0.5))
    assert "await the previous operation" in out["overlap"]
    assert "faceStates" in out["faceStates"]
    batch.setCoarse(coarse, scales)
    for k in range(len(specs)):
        assert np.array_equal(np.array(out["coarseRho"][k], dtype=np.float32), batch.readState(k).rho), k
    batch.destroy()
    b2 = GpuBatch2D(ctx, cfgs)
    b2.setUniform(frees)
    res = solve(b2, ps, qs, 0.5, SolvePlan(chunk=240, max_steps=4800))
    for k, rr in enumerate(res):
        js = out["solve"][k]
        assert js["reason"] == rr.reason and js["steps"] == rr.steps and js["converged"] == rr.converged, k
        assert js["cd"] == rr.cd, k
