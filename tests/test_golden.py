"""The golden fixtures of tests/golden (written by make_golden.py from the
FP64 oracle) against the oracle re-run now: any change to the oracle's
arithmetic shows as a difference here, bit for bit."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import make_golden  # noqa: E402


@pytest.mark.parametrize("name", sorted(make_golden.MAKERS))
def test_oracle_reproduces_the_fixture(name):
    with np.load(os.path.join(HERE, name + ".npz")) as f:
        saved = {k: f[k] for k in f.files}
    now = make_golden.MAKERS[name]()
    assert sorted(saved) == sorted(now)
    for k in saved:
        assert np.array_equal(saved[k], now[k]), (name, k)


def test_fixtures_are_nontrivial():
    """Every regime the fixtures claim is in them: subsonic and supersonic
    faces, limited and unlimited reconstructions that differ, a residual
    that is not zero, a trajectory that moved."""
    with np.load(os.path.join(HERE, "face_fluxes.npz")) as f:
        left, nx, ny, hllc, rus = f["left"], f["nx"], f["ny"], f["hllc"], f["rusanov"]
    un = left[:, 1] * nx + left[:, 2] * ny
    mach = np.abs(un) / np.sqrt(1.4 * left[:, 3] / left[:, 0])
    assert (mach < 1).any() and (mach > 1).any()
    assert not np.allclose(hllc, rus)
    with np.load(os.path.join(HERE, "reconstruct.npz")) as f:
        assert not np.allclose(f["perVariable_minmod_+1"], f["perVariable_unlimitedCentral_+1"])
        assert np.array_equal(f["perVariable_firstOrder_+1"], f["centre"])
    with np.load(os.path.join(HERE, "residual_4x4.npz")) as f:
        assert np.abs(f["residual_axisymmetric"]).max() > 1e-3
        assert not np.allclose(f["residual_planar"], f["residual_axisymmetric"])
    with np.load(os.path.join(HERE, "trajectory_8x8.npz")) as f:
        assert np.abs(f["after_10"] - f["start"]).max() > 1e-3
