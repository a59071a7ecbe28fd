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
    orc2, _ = cases.make_pair(None, grid, bcs, limiter="vanAlbada", reconstruction="sharedScaled")
    orc2.set_primitives(np.array(free))
    cfl = 1.0
    for _ in range(20):
        orc2.iterate_implicit(1, cfl, 1.5, "redBlack", 2)
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
