"""Case generators for batched solves (CPU): the seeded draws, the grid
sizing rules, refusals, and the vectorized grid generator against the
scalar one."""
import json
import math
import shutil
import subprocess

import numpy as np
import pytest

from bowshock_amd import grids
from bowshock_amd import population as P


def test_mulberry32_is_a_32_bit_generator():
    """Uniform in [0, 1), reproducible, period-free over a long draw, and
    each draw a multiple of 2^-32."""
    a = P.mulberry32(P.SEED)
    b = P.mulberry32(P.SEED)
    xs = np.array([a() for _ in range(20000)])
    assert np.array_equal(xs, [b() for _ in range(20000)])
    assert xs.min() >= 0 and xs.max() < 1
    assert np.all(np.mod(xs * 2 ** 32, 1) == 0)
    assert abs(xs.mean() - 0.5) < 0.01 and len(np.unique(xs)) == len(xs)
    assert P.mulberry32(1)() != P.mulberry32(2)()


JS_MULBERRY32 = """
// Tommy Ettinger's published Mulberry32, in JavaScript's own 32-bit
// integer operations (Math.imul, >>>, ^), which population.ts:64-72 uses.
const seeds = JSON.parse(process.argv[1]);
const out = seeds.map((seed) => {
  let s = seed >>> 0;
  const draws = [];
  for (let k = 0; k < 2000; k++) {
    s = (s + 0x6d2b79f5) | 0;
    let z = Math.imul(s ^ (s >>> 15), 1 | s);
    z = (z + Math.imul(z ^ (z >>> 7), 61 | z)) ^ z;
    draws.push(((z ^ (z >>> 14)) >>> 0) / 4294967296);
  }
  return draws;
});
console.log(JSON.stringify(out));
"""


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_mulberry32_matches_javascript_integer_semantics():
    """The Python draw emulates JavaScript's signed 32-bit Math.imul, the
    unsigned shifts and the double-valued sum folded back by ^=; Node runs
    the generator with the real operations (the state kept as a signed
    int32 here, as an unsigned one in population.ts: the same bits), for
    seeds that cross the sign bit, wrap, and come in negative."""
    seeds = [P.SEED, 0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, -1, -123456789, 2 ** 31 + 12345]
    r = subprocess.run([shutil.which("node"), "-e", JS_MULBERRY32, json.dumps(seeds)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    expected = json.loads(r.stdout)
    for seed, js in zip(seeds, expected):
        g = P.mulberry32(seed)
        assert [g() for _ in range(len(js))] == js, seed


def test_population_sizes_follow_the_rule():
    cs = P.population_cases(50)
    assert len(cs) == 200
    for c in cs:
        L = c.rocket.fineness + c.cylinder
        assert 160 <= c.ni <= 640 and c.ni % 2 == 0 and c.nj == c.ni // 2
        assert abs(c.ni - (160 + 480 * (L - 7) / 18)) <= 1
        assert 1 <= c.rocket.fineness <= 5 and 6 <= c.cylinder <= 20 and 24 <= c.rocket.diameter_mm <= 100
    sw = P.sweep_cases()
    assert len(sw) == 39 and sw[0].mach == 1.2 and sw[-1].mach == 5.0
    assert len({(c.rocket, c.ni) for c in sw}) == 1


def test_refusals():
    blunt = P.Case(P.Rocket(0, "ellipsoid", 2.0, 50, 8.0), 3.0, 160, 80, 8.0)
    gated = P.Case(P.Rocket(4, "von-karman", 4.85, 50, 6.4), 2.0, 274, 137, 6.4)
    steep = P.Case(P.Rocket(1, "parabolic", 1.2, 50, 8.0), 3.0, 160, 80, 8.0)
    detached = P.Case(P.Rocket(2, "tangent-ogive", 1.95, 50, 8.0), 1.2, 160, 80, 8.0)
    fine = P.Case(P.Rocket(3, "conical", 3.0, 50, 8.0), 2.0, 40, 20, 8.0)
    all_ = [blunt, gated, steep, detached, fine]
    P.build_grids(all_, workers=1)
    counts = P.refusal_counts(all_)
    assert counts == {"accepted (C grid)": 1, "C grid fails its quality gates": 1, "tip half-angle above 30 deg": 1,
                      "detached tip shock": 1, "accepted": 1}
    assert fine.grid is not None and fine.grid.ni == 40 and fine.grid.nj == 20 and fine.topology == "H"
    # The blunt tip's C grid: 20 cells across refined by 2 (ni 160 / 80).
    assert blunt.topology == "C" and blunt.nj == 40 and blunt.grid.nj == 40 and blunt.grid.ni == blunt.ni
    free = grids.free_stream_at(3.0, 1.4)[0]
    assert P.boundaries(blunt, free)["iMin"] == {"kind": "wall"}
    assert P.boundaries(fine, free)["iMin"]["kind"] == "inflow"
    assert [P.c_grid_refinement(n) for n in (160, 226, 240, 320, 640)] == [2, 2, 4, 4, 8]


@pytest.mark.parametrize("body,mach", [(grids.ogive_cylinder(3.0, 15.0), 2.0), (grids.cone_cylinder(15, 4), 3.0),
                                       (grids.Body(grids.Nose("parabolic", 2.5, 0.5, 1.0), 8.0), 1.5)])
def test_array_generator_matches_the_scalar_one(body, mach):
    a = grids.body_grid_fixed(body, 96, 48, mach, 1.4)
    b = grids.body_grid_array(body, 96, 48, grids.tip_shock_angle(body, mach, 1.4))
    assert a.i_tip == b.i_tip and a.outer_angle == b.outer_angle
    assert np.allclose(a.grid.x, b.grid.x, rtol=0, atol=1e-13)
    assert np.allclose(a.grid.y, b.grid.y, rtol=0, atol=1e-13)
    for x in np.linspace(-0.5, grids.body_length(body) + 0.1, 57):
        assert grids.body_radius_array(body, x) == pytest.approx(grids.body_radius(body, x), abs=1e-15)
