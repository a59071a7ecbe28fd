"""The H grid of sharp noses (bowshock_amd/grids.py body_grid) against
the properties of the reference's tests/grid/body-grid.test.ts, on
numpy-seeded draws of its random rockets (noses of 2 to 8 calibers,
conical, ogive 0.2-1, parabolic 0-1; no cylinder or 1 to 25 calibers),
and the column positions of tests/grid/spacing.test.ts with the
reference's measured cell counts.
"""
import math

import numpy as np
import pytest

from bowshock_amd import grids
from bowshock_amd.blunt_grid import column_positions_array
from bowshock_amd.grids import Body, GridQualityError, Nose

GAMMA = 1.4


def _rockets(seed, count, mach_lo, mach_hi):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        kind = ("conical", "ogive", "parabolic")[rng.integers(3)]
        parameter = {"conical": 0.0, "ogive": rng.integers(200, 1001) / 1000,
                     "parabolic": rng.integers(0, 1001) / 1000}[kind]
        length = rng.integers(2000, 8001) / 1000
        cyl = 0.0 if rng.random() < 0.5 else rng.integers(1000, 25001) / 1000
        mach = rng.integers(int(1000 * mach_lo), int(1000 * mach_hi) + 1) / 1000
        body = Body(Nose(kind, length, 0.5, parameter), cyl)
        # The generator's own draws keep tips below the cone that detaches at
        # the Mach number drawn; the gate refuses the rest.
        try:
            grids.tip_shock_angle(body, mach, GAMMA, 1e-10)
        except ValueError:
            continue
        out.append((body, mach))
    return out


@pytest.mark.parametrize("seed,lo,hi", [(121, 2.0, 5.0), (123, 1.5, 2.0)])
def test_random_rockets_pass_every_gate(seed, lo, hi):
    cases = _rockets(seed, 40, lo, hi)
    assert len(cases) > 20
    for body, mach in cases:
        g = grids.body_grid(body, 20, mach, GAMMA).grid
        assert grids.quality_failures(grids.grid_quality(g)) == [], (body, mach)


def test_wall_nodes_on_the_profile_tip_and_shoulder_on_lines():
    for body, mach in _rockets(122, 20, 2.0, 5.0):
        bg = grids.body_grid(body, 20, mach, GAMMA)
        g = bg.grid
        xs = g.x[:g.ni + 1]
        for i in range(g.ni + 1):
            assert abs(g.y[i] - grids.body_radius(body, g.x[i])) < 1e-12
        assert xs[bg.i_tip] == 0.0
        assert body.nose.length in xs.tolist()
        assert abs(xs[g.ni] - (body.nose.length + body.cylinder_length)) < 1e-12
        assert np.all(np.diff(xs) > 0)


def test_refined_grids_are_nested():
    for body, mach in _rockets(124, 8, 2.0, 5.0):
        coarse = grids.body_grid(body, 20, mach, GAMMA).grid
        fine = grids.body_grid(body, 20, mach, GAMMA, refinement=2).grid
        assert fine.ni == 2 * coarse.ni
        C = np.stack([coarse.x, coarse.y]).reshape(2, coarse.nj + 1, coarse.ni + 1)
        F = np.stack([fine.x, fine.y]).reshape(2, fine.nj + 1, fine.ni + 1)[:, ::2, ::2]
        scale = 1 + np.abs(C).sum(axis=0)
        assert (np.abs(F - C).max(axis=0) / scale).max() < 1e-12


def test_tip_shock_stays_three_cells_below_the_outer_boundary():
    for body, mach in _rockets(125, 20, 2.0, 5.0):
        bg = grids.body_grid(body, 20, mach, GAMMA)
        g = bg.grid
        shock = grids.taylor_maccoll(mach, math.atan(grids.nose_tip_slope(body.nose)), GAMMA).shock_angle
        Y = g.y.reshape(g.nj + 1, g.ni + 1)
        for i in range(bg.i_tip + 1, g.ni + 1):
            line = g.x[i] * math.tan(shock)
            j = 0
            while Y[j, i] < line:
                j += 1
            cells = g.nj - j + (Y[j, i] - line) / (Y[j, i] - Y[j - 1, i])
            assert cells > 3 - 1e-9, (body, mach, i)


def test_refusals():
    cone = Body(Nose("conical", 2.0, 0.5), 3.0)
    with pytest.raises(ValueError, match="refinement"):
        grids.body_grid(cone, 20, 3.0, GAMMA, refinement=3)
    assert grids.body_grid(cone, 20, 3.0, GAMMA, refinement=4).grid.nj == 80
    sliver = Body(Nose("ogive", 3.0, 0.5, 1.0), 0.001)
    with pytest.raises(GridQualityError, match="differ in size"):
        grids.body_grid(sliver, 20, 2.0, GAMMA)
    steep = Body(Nose("conical", 0.5 / math.tan(math.radians(32)), 0.5), 3.0)
    with pytest.raises(ValueError, match="tip half-angles above 30"):
        grids.body_grid(steep, 20, 5.0, GAMMA)
    with pytest.raises(ValueError, match="blunt noses"):
        grids.body_grid(Body(Nose("haack", 3.0, 0.5, 0.0), 5.0), 20, 2.0, GAMMA)
    squat = Body(Nose("conical", 0.5 / math.tan(math.radians(28)), 0.5), 5.0)
    with pytest.raises(ValueError, match="detach"):
        grids.body_grid(squat, 20, 1.3, GAMMA)


def _steep(refinement):
    return lambda x: refinement / (0.05 + 0.2 / (1 + np.exp(-40 * (np.asarray(x) - 1))))


def _largest_ratio(xs):
    d = np.diff(xs)
    return float(max(1.0, np.max(d[1:] / d[:-1]), np.max(d[:-1] / d[1:])))


def test_spacing_slope_limit_across_breakpoints():
    """spacing.test.ts:26-36, with the reference's measured counts: 28
    cells at ratio 2.52 free, 32 at 1.156 limited."""
    bps = [0.0, 0.95, 3.0]
    free = column_positions_array(bps, _steep(1), 1)
    limited = column_positions_array(bps, _steep(1), 1, max_spacing_slope=0.15)
    assert _largest_ratio(free) > 2.5 and _largest_ratio(limited) < 1.16
    assert len(free) - 1 == 28 and len(limited) - 1 == 32
    assert all(b in limited for b in bps)


def test_spacing_gentle_density_untouched_by_the_limit():
    bps = [0.001, 1.0, 4.0]
    for r in (1, 2):
        def density(x, r=r):
            return r / (0.05 + 0.05 * np.asarray(x))
        assert column_positions_array(bps, density, r, max_spacing_slope=0.15) == column_positions_array(bps, density, r)


def test_spacing_limited_family_nests_exactly():
    bps = [0.0, 0.95, 3.0]
    base = column_positions_array(bps, _steep(1), 1, max_spacing_slope=0.15)
    for r in (2, 4, 8):
        fine = column_positions_array(bps, _steep(r), r, max_spacing_slope=0.15)
        assert len(fine) - 1 == r * (len(base) - 1)
        assert fine[::r] == base
