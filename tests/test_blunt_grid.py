"""The blunt-body C grid (bowshock_amd/blunt_grid.py) against the
properties the reference's own tests hold for bluntBodyGrid
(tests/grid/blunt-grid.test.ts): quality gates on random blunt rockets,
nodes on the body, the axis and the end, the outer standoff, exact nesting
of the grids of a family (with and without the spacing's slope limit, on
the generator's boundary and on one given), the two length scales of a
Haack nose against the asymptote of its r ~ x^(3/4) tip, the refusals,
and fixed node fractions. The reference's randomized property (fast-check,
seed 7) is run on numpy-seeded draws of the same distribution.
"""
import math

import numpy as np
import pytest

from bowshock_amd import grids
from bowshock_amd.blunt_grid import (billig_standoff, blunt_body_grid, column_positions_array, grid_on_lines,
                                     outer_cell_height, stretched_fraction, line_strength, with_fixed_fractions)
from bowshock_amd.grids import Body, GridQualityError, Nose, body_length, body_radius

HEMI = Body(Nose("ellipsoid", 0.5, 0.5), 4.0)


def _node(g, i, j):
    return i + j * (g.ni + 1)


def _nests(coarse, fine, r, exact=True):
    worst = 0.0
    for j in range(coarse.nj + 1):
        c = np.arange(coarse.ni + 1) + j * (coarse.ni + 1)
        f = r * np.arange(coarse.ni + 1) + r * j * (fine.ni + 1)
        if exact:
            assert np.array_equal(fine.x[f], coarse.x[c]) and np.array_equal(fine.y[f], coarse.y[c]), (r, j)
        worst = max(worst, np.abs(fine.x[f] - coarse.x[c]).max(), np.abs(fine.y[f] - coarse.y[c]).max())
    return worst


def test_random_blunt_rockets_pass_every_gate():
    """blunt-grid.test.ts:38-50: hemisphere to 3:1 ellipsoid noses of
    radius 0.5, 2 to 15 calibers of cylinder, Mach 1.5 to 5."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        length = rng.integers(500, 1501) / 1000
        cyl = rng.integers(2000, 15001) / 1000
        mach = rng.integers(1500, 5001) / 1000
        g = blunt_body_grid(Body(Nose("ellipsoid", length, 0.5), cyl), 20, mach)
        q = grids.grid_quality(g.grid)
        assert grids.quality_failures(q) == [], (length, cyl, mach)
        assert q["min_wall_angle"] > math.radians(80), (length, cyl, mach)


def test_nodes_on_the_body_the_axis_and_the_end():
    """blunt-grid.test.ts:52-66."""
    b = blunt_body_grid(HEMI, 20, 3.0)
    g = b.grid
    for i in range(g.ni + 1):
        k = _node(g, i, 0)
        assert abs(g.y[k] - body_radius(HEMI, g.x[k])) < 1e-12
    for j in range(g.nj + 1):
        assert g.y[_node(g, 0, j)] == 0.0
        assert abs(g.x[_node(g, g.ni, j)] - body_length(HEMI)) < 1e-12
    assert abs(g.x[_node(g, 0, g.nj)] + b.outer_standoff) < 1e-9
    assert abs(b.outer_standoff - 2 * billig_standoff(3.0, 0.5)) < 1e-12
    assert b.outer_angle == (math.pi / 2 + math.asin(1 / 3.0)) / 2


def test_family_keeps_every_coarser_node():
    """blunt-grid.test.ts:68-88 (1e-12 there; exact here)."""
    coarse = blunt_body_grid(HEMI, 20, 2.0).grid
    for r in (2, 4):
        fine = blunt_body_grid(HEMI, 20, 2.0, refinement=r).grid
        assert fine.ni == r * coarse.ni and fine.nj == r * coarse.nj
        assert _nests(coarse, fine, r) == 0.0


def test_slope_limited_spacing_and_exact_nesting():
    """blunt-grid.test.ts:90-136: a 1.5-caliber ellipsoid with a caliber of
    cylinder at Mach 1.2, where the lines stop turning at the shoulder: the
    cells along the body grow by less than 1.17 (1.1605 in the reference),
    the shoulder is a node, and refinements 2, 4, 8 keep every coarse node
    exactly."""
    body = Body(Nose("ellipsoid", 1.5, 0.5), 1.0)
    b = blunt_body_grid(body, 20, 1.2, gated=False)
    ds = np.hypot(np.diff(b.lines.x), np.diff(b.lines.y))
    worst = np.maximum(ds[1:] / ds[:-1], ds[:-1] / ds[1:]).max()
    assert 1.0 < worst < 1.17
    assert np.any(b.lines.x == body.nose.length)
    for r in (2, 4, 8):
        fine = blunt_body_grid(body, 20, 1.2, refinement=r, gated=False).grid
        assert fine.ni == r * b.grid.ni and fine.nj == r * b.grid.nj
        _nests(b.grid, fine, r)


def test_haack_two_scales_and_ellipsoid_one():
    """blunt-grid.test.ts:138-166: r = c x^(3/4) at a Haack tip crosses
    r = x at (64 / pi^2) (2/3 + C)^2 R^4 / L^3; the series' higher terms
    lower it on the 3-caliber nose by 0.13 % (C = 0) and 0.70 % (C = 1/3),
    the figures the reference states."""
    for c, lowered in ((0.0, 0.0013), (1 / 3, 0.0070)):
        g = blunt_body_grid(Body(Nose("haack", 3.0, 0.5, c), 2.0), 20, 2.0, gated=False)
        asymptotic = (64 / math.pi ** 2) * (2 / 3 + c) ** 2 * 0.5 ** 4 / 3 ** 3
        assert abs(g.tip_scale / asymptotic - 1) < 1e-2
        assert abs((1 - g.tip_scale / asymptotic) - lowered) < 2e-4
        assert g.two_scale and g.scale == 2 * g.tip_scale
    for length in (0.5, 3.0, 6.0):
        g = blunt_body_grid(Body(Nose("ellipsoid", length, 0.5), 4.0), 20, 2.0, gated=False)
        assert not g.two_scale
        assert g.scale == max(0.5, 0.25 / length)


def test_given_outer_boundary_nests_exactly():
    """blunt-grid.test.ts:168-199: a boundary made from the generator's
    own lines at 0.4 of their heights."""
    haack = Body(Nose("haack", 3.0, 0.5, 0.0), 2.0)
    first = blunt_body_grid(haack, 20, 2.0, gated=False)
    L = first.lines
    outer = np.stack([L.x + 0.4 * L.h * L.dx, L.y + 0.4 * L.h * L.dy], axis=1)
    coarse = blunt_body_grid(haack, 20, 2.0, outer_boundary=outer, gated=False).grid
    assert abs(coarse.x[_node(coarse, 0, coarse.nj)] - outer[0, 0]) < 1e-15
    for r in (2, 4):
        fine = blunt_body_grid(haack, 20, 2.0, refinement=r, outer_boundary=outer, gated=False).grid
        assert fine.ni == r * coarse.ni and fine.nj == r * coarse.nj
        _nests(coarse, fine, r)


def test_sharp_tips_served_and_refusals():
    """blunt-grid.test.ts:201-221."""
    sharp = Body(Nose("conical", 2.0, 0.5), 4.0)
    assert blunt_body_grid(sharp, 20, 3.0, gated=False).two_scale
    power = Body(Nose("power", 0.5, 0.5, 0.5), 4.0)   # meets the cylinder at 26.6 deg
    with pytest.raises(ValueError, match="meeting the body at more than"):
        blunt_body_grid(power, 20, 3.0)
    stub = Body(Nose("ellipsoid", 0.5, 0.5), 0.0)
    with pytest.raises(GridQualityError):
        blunt_body_grid(stub, 20, 1.5)
    with pytest.raises(ValueError):
        blunt_body_grid(HEMI, 3, 3.0)
    with pytest.raises(ValueError):
        blunt_body_grid(HEMI, 20, 3.0, refinement=3)
    with pytest.raises(ValueError):
        blunt_body_grid(HEMI, 20, 1.0)
    with pytest.raises(ValueError, match="first on the axis"):
        blunt_body_grid(HEMI, 20, 3.0, outer_boundary=[(-0.2, 0.1), (0.0, 0.5)])


def test_fixed_fractions_reproduce_and_scale():
    """blunt-grid.test.ts:223-238."""
    b = blunt_body_grid(HEMI, 20, 3.0)
    fixed = with_fixed_fractions(b.lines, 20, 0.5)
    same = grid_on_lines(fixed, fixed.h, 20, 0.5)
    assert np.array_equal(same.x, b.grid.x) and np.array_equal(same.y, b.grid.y)
    with pytest.raises(ValueError, match="cells across"):
        grid_on_lines(fixed, fixed.h, 16, 0.5)
    taller = grid_on_lines(fixed, 1.1 * fixed.h, 20, 0.5, gated=False)
    for i in (0, 40, len(fixed) - 1):
        for j in (5, 20):
            k = _node(taller, i, j)
            d = (taller.x[k] - fixed.x[i]) * fixed.dx[i] + (taller.y[k] - fixed.y[i]) * fixed.dy[i]
            assert abs(d / (1.1 * fixed.h[i]) - fixed.fractions[j, i]) < 5e-14


def test_stretching_closed_forms():
    """The node fractions are (e^(b j / nj) - 1) / (e^b - 1) with
    b = min(ln(1 + h / R), 2 ln 10): 0 and 1 at the ends, the ratio of
    neighbouring cell heights e^(b / nj), and the outer cell of a line of
    height h the fraction (e^b - e^(b (nj - 1) / nj)) / (e^b - 1) of it."""
    h, R, nj = 1.3, 0.5, 20
    b = float(line_strength(h, R))
    assert abs(b - math.log1p(h / R)) < 1e-15
    f = stretched_fraction(b, np.arange(nj + 1), nj)
    assert f[0] == 0.0 and abs(f[-1] - 1) < 1e-15
    ratios = np.diff(f)[1:] / np.diff(f)[:-1]
    assert np.allclose(ratios, math.exp(b / nj), rtol=1e-12)
    assert abs(outer_cell_height(h, nj, R) - h * (math.exp(b) - math.exp(b * (nj - 1) / nj)) / math.expm1(b)) < 1e-14
    assert float(line_strength(1e6, R)) == 2 * math.log(10.0)


def test_vectorised_columns_equal_the_scalar_ones():
    """column_positions_array (density on arrays) against the scalar
    restatement of spacing.ts columnPositions used by the H grids, on a
    density with a tip cluster: the same nodes, bit for bit."""
    def density(s):
        return 40.0 / (0.02 + np.sqrt(np.abs(s))) + 3.0 * np.sin(s) ** 2

    bps = [0.0, 0.7, 2.5]
    for r in (1, 2, 4):
        a = column_positions_array(bps, density, r)
        b = grids.column_positions(bps, lambda s: float(density(np.array([s]))[0]), r)
        assert a == b


def test_slope_limit_bounds_the_spacing():
    """With the limit, neighbouring base cells differ in length by about the
    slope at most, the breakpoints stay nodes, and a density whose spacing
    already changes slowly is left as it is."""
    def density(s):
        return np.where(s < 1.0, 60.0, 6.0)  # spacing jumps tenfold at 1

    xs = np.array(column_positions_array([0.0, 1.0, 3.0], density, 1, max_spacing_slope=0.15))
    d = np.diff(xs)
    assert 1.0 in xs.tolist() and xs[0] == 0.0 and xs[-1] == 3.0
    assert np.max(d[1:] / d[:-1]) < 1.2
    smooth = column_positions_array([0.0, 1.0], lambda s: 20.0 + 0 * s, 1, max_spacing_slope=0.15)
    assert smooth == column_positions_array([0.0, 1.0], lambda s: 20.0 + 0 * s, 1)


def test_grid_quality_measures():
    """tests/grid/quality.test.ts: a uniform square grid is perfect; a grid
    sheared by 40 deg with rows growing by 1.5 measures 50 deg and 1.5 and
    fails with the reference's message; a collapsed corner fails rather
    than slipping through on a NaN angle. The fitted grids' limits relax
    the wall angle to 50 deg and the aspect ratio to 13 (quality.ts:64-68)."""
    q = grids.grid_quality(grids.rectangular_grid(4, 4, 0.0, 1.0, 0.0, 1.0))
    assert abs(q["min_angle"] - math.pi / 2) < 1e-12 and abs(q["max_size_ratio"] - 1) < 1e-12
    assert abs(q["max_aspect_ratio"] - 1) < 1e-12 and abs(q["min_area"] - 1 / 16) < 1e-12
    assert grids.quality_failures(q) == []
    shear = math.tan(math.radians(40))

    def node(i, j):
        y = 0.25 * (1.5 ** j - 1)
        return 4 * i * 0.25 + y * shear, y

    q = grids.grid_quality(grids.grid_from_function(3, 3, node))
    assert abs(math.degrees(q["min_angle"]) - 50) < 1e-9 and abs(q["max_size_ratio"] - 1.5) < 1e-9
    assert "differ in size by 1.50 (limit 1.3)" in " ".join(grids.quality_failures(q))
    degenerate = grids.grid_from_function(2, 1, lambda i, j: (0.0, 1.0) if (i == 1 and j == 1) else (float(i), float(j)))
    assert len(grids.quality_failures(grids.grid_quality(degenerate))) > 0
    assert grids.FITTED_QUALITY_LIMITS["min_wall_angle"] == 50 * math.pi / 180
    assert grids.FITTED_QUALITY_LIMITS["max_aspect_ratio"] == 13
    assert grids.FITTED_QUALITY_LIMITS["max_size_ratio"] == grids.QUALITY_LIMITS["max_size_ratio"]


def test_fitted_start_finds_a_planted_shock():
    """bowshock_amd/fitted_start.py on a synthetic captured flow: pressure
    10 below a planted shock height (a smooth function of the line) and 1
    above it, linear across one cell: the heights come back within a cell
    of the planted ones before smoothing, the fitted grid's top nodes lie
    on them, and the start carries the post-shock state below and the
    free stream nowhere inside."""
    from bowshock_amd.fitted_start import _centres, fitted_start, shock_heights

    nj = 20
    b = blunt_body_grid(HEMI, nj, 3.0)
    g, L = b.grid, b.lines
    ni = g.ni
    planted = 0.5 * L.h
    cx, cy = _centres(g)
    prim = np.zeros((nj, ni, 4))
    prim[..., 0], prim[..., 1] = 1.0, 1.0
    for i in range(ni):
        d = (cx[:, i] - L.x[i]) * L.dx[i] + (cy[:, i] - L.y[i]) * L.dy[i]
        hs = 0.5 * (planted[i] + planted[i + 1])
        prim[:, i, 3] = np.where(d < hs, 10.0, 1.0)
    prim = prim.reshape(-1, 4)
    h0 = shock_heights(L, g, prim, smoothing=0)
    X, Y = g.x.reshape(nj + 1, ni + 1)[:, 0], g.y.reshape(nj + 1, ni + 1)[:, 0]
    spacing = np.max(np.diff((X - L.x[0]) * L.dx[0] + (Y - L.y[0]) * L.dy[0]))   # tallest cell on the axis line
    assert np.all(np.abs(h0 - planted) < 0.25 * L.h)
    assert abs(h0[0] - planted[0]) < 2 * spacing
    st = fitted_start(L, g, prim, 0.5)
    top = np.stack([st.grid.x, st.grid.y]).reshape(2, nj + 1, ni + 1)[:, nj]
    along = (top[0] - L.x) * L.dx + (top[1] - L.y) * L.dy
    assert np.allclose(along, st.heights, rtol=0, atol=1e-12)
    assert st.primitives[:, 3].min() >= 1.0 and np.median(st.primitives[:, 3]) == 10.0
