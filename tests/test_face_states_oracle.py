"""Per-face boundary states and imposed face fluxes of the oracle
(orc_set_face_states / orc_set_face_fluxes, restating euler2d.ts:357-398),
pinned by the reference's own closed-form tests
(tests/engine-cpu/euler2d.test.ts:305-406) with their cases and bars:

- a faceStates side given the inflow state on every face reproduces the
  inflow side's run exactly;
- per-face states are taken, can be replaced (the solution changes), and a
  wrong count or a side of another kind is refused;
- the free stream's own flux imposed on a side keeps a uniform stream
  uniform to 1e-12, a wrong count is refused, and null goes back to the
  Riemann flux.
"""
import numpy as np
import pytest

import pyoracle as O
from bowshock_amd import grids

GAMMA = 1.4
FREE = (1.0, 2.5, 0.1, 0.8)  # rho, u, v, p


def ramp_grid():
    """euler2d.test.ts:308: a 16 x 8 box whose top rises with i."""
    return grids.grid_from_function(16, 8, lambda i, j: (i / 16, (j / 8) * (0.5 + 0.02 * i)))


def config(jmax, jmin=None):
    return dict(geometry="planar", gamma=GAMMA, flux="hllc", limiter="vanAlbada", reconstruction="sharedScaled",
                time_integrator="ssprk2", cfl=0.5, time_stepping="local",
                boundaries={"iMin": {"kind": "inflow", "state": FREE}, "iMax": {"kind": "outflow"},
                            "jMin": jmin or {"kind": "wall"}, "jMax": jmax})


def free_stream_fluxes(metrics, ni, nj):
    """The free stream's physical flux through each jMax face, per unit
    area along the face normal (euler2d.test.ts:377-392)."""
    rho, u, v, p = FREE
    energy = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    out = np.zeros((ni, 4))
    for i in range(ni):
        f = i + nj * ni
        nx, ny = metrics["jnx"][f], metrics["jny"][f]
        un = u * nx + v * ny
        out[i] = (rho * un, rho * u * un + p * nx, rho * v * un + p * ny, un * (energy + p))
    return out.reshape(-1)


def test_face_states_equal_to_the_inflow_reproduce_the_inflow_side_exactly():
    grid = ramp_grid()
    a = O.Euler2D(grid, **config({"kind": "inflow", "state": FREE}))
    b = O.Euler2D(grid, **config({"kind": "faceStates", "states": np.tile(FREE, (16, 1))}))
    for s in (a, b):
        s.set_primitives(np.array(FREE))
        s.iterate(200)
    # Not a trivial pass: the wall and the rising top make the flow move.
    assert np.abs(a.primitives() - np.array(FREE)).max() > 1e-2
    assert np.array_equal(b.primitives(), a.primitives())


def test_face_states_per_face_replaceable_and_refusals():
    grid = ramp_grid()
    states = np.array([(FREE[0], FREE[1], FREE[2], 0.8 + 0.01 * i) for i in range(16)])
    s = O.Euler2D(grid, **config({"kind": "faceStates", "states": states}))
    s.set_primitives(np.array(FREE))
    s.iterate(50)
    cell = 10 + 7 * 16
    before = s.primitives()[cell, 3]
    doubled = states.copy()
    doubled[:, 3] *= 2
    s.set_face_states("jMax", doubled)
    s.iterate(50)
    assert s.primitives()[cell, 3] != before
    with pytest.raises(O.OracleError, match="16 face states"):
        s.set_face_states("jMax", states[1:])
    with pytest.raises(O.OracleError, match="does not take face states"):
        s.set_face_states("iMin", states)


def test_imposed_free_stream_flux_keeps_a_uniform_stream():
    grid = ramp_grid()
    states = np.tile(FREE, (16, 1))
    s = O.Euler2D(grid, **config({"kind": "faceStates", "states": states}, jmin={"kind": "inflow", "state": FREE}))
    fluxes = free_stream_fluxes(s.metrics, 16, 8)
    s.set_face_fluxes("jMax", fluxes)
    s.set_primitives(np.array(FREE))
    s.iterate(50)
    w = s.primitives()
    assert np.abs(w[:, 3] - FREE[3]).max() < 1e-12
    assert np.abs(w[:, 2] - FREE[2]).max() < 1e-12
    with pytest.raises(O.OracleError, match="64 flux values"):
        s.set_face_fluxes("jMax", fluxes[4:])
    s.set_face_fluxes("jMax", None)


def test_imposed_flux_is_used_instead_of_the_riemann_flux():
    """A planted check that the imposed flux is the one the residual uses:
    a doubled mass flux through jMax changes the residual of exactly the
    cells under the top row, by the flux difference over their volume."""
    grid = ramp_grid()
    states = np.tile(FREE, (16, 1))
    s = O.Euler2D(grid, **config({"kind": "faceStates", "states": states}, jmin={"kind": "inflow", "state": FREE}))
    fluxes = free_stream_fluxes(s.metrics, 16, 8)
    s.set_face_fluxes("jMax", fluxes)
    s.set_primitives(np.array(FREE))
    r0 = s.residual_vector()
    bumped = fluxes.copy().reshape(16, 4)
    bumped[:, 0] *= 2
    s.set_face_fluxes("jMax", bumped.reshape(-1))
    r1 = s.residual_vector()
    d = (r1 - r0).reshape(8, 16, 4)
    assert np.abs(d[:7]).max() == 0.0
    assert np.abs(d[7, :, 1:]).max() == 0.0
    area = s.metrics["jarea"][8 * 16:9 * 16]
    assert np.allclose(d[7, :, 0], -fluxes.reshape(16, 4)[:, 0] * area, rtol=1e-12, atol=1e-14)
