"""Host-side grids and body geometry that feed the Euler path.

Grid generation is preprocessing: it runs once per case on the host, in
double precision, and hands node coordinates to the solver. The functions
restate the reference's generators where a result must match it exactly
(``body_grid`` with ``column_positions`` and the Taylor-Maccoll outer
angle), and add a fixed-count generator for grids of a given size.

Conventions: node (i, j) at ``i + j * (ni + 1)``; in axisymmetric geometry
x is axial and y radial; lengths in body diameters when the base radius is
0.5.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class StructuredGrid:
    """Nodes of an ni x nj cell block (src/grid/structured.ts:12-18)."""

    ni: int
    nj: int
    x: np.ndarray  # (ni + 1) * (nj + 1) float64
    y: np.ndarray


def grid_from_function(ni: int, nj: int, node) -> StructuredGrid:
    """src/grid/structured.ts:57-75 gridFromFunction."""
    if not (isinstance(ni, int) and isinstance(nj, int) and ni >= 1 and nj >= 1):
        raise ValueError("a grid needs at least one cell in each direction")
    x = np.empty((nj + 1) * (ni + 1))
    y = np.empty_like(x)
    for j in range(nj + 1):
        for i in range(ni + 1):
            px, py = node(i, j)
            x[i + j * (ni + 1)] = px
            y[i + j * (ni + 1)] = py
    return StructuredGrid(ni, nj, x, y)


def rectangular_grid(ni, nj, x0, x1, y0, y1) -> StructuredGrid:
    """src/grid/structured.ts:78-87 rectangularGrid (vectorised)."""
    i = np.arange(ni + 1, dtype=np.float64)
    j = np.arange(nj + 1, dtype=np.float64)
    xx = x0 + ((x1 - x0) * i) / ni
    yy = y0 + ((y1 - y0) * j) / nj
    X, Y = np.meshgrid(xx, yy)
    return StructuredGrid(ni, nj, X.ravel().copy(), Y.ravel().copy())


# --------------------------------------------------------------------------
# Nose and body geometry (src/geometry/nose.ts, src/geometry/body.ts)
# --------------------------------------------------------------------------

OGIVE_CONE_LIMIT = 0.001  # nose.ts:105


@dataclass(frozen=True)
class Nose:
    """kind: 'conical', 'ogive', 'ellipsoid', 'power', 'parabolic', 'haack'."""

    kind: str
    length: float
    base_radius: float
    parameter: float = 0.0


@dataclass(frozen=True)
class Body:
    nose: Nose
    cylinder_length: float


def _ogive_arc(p, L, R):
    """nose.ts:115-126 ogiveArc."""
    stretch = R / L if L < R else 1.0
    length = max(L, R)
    rho = math.sqrt(((length * length + R * R) * (((2 - p) * length) ** 2 + (p * R) ** 2))
                    / (4 * (p * R) ** 2))
    xc = length / p
    return stretch, length, rho, xc, math.sqrt(max(0.0, rho * rho - xc * xc))


def nose_radius(nose: Nose, x: float) -> float:
    """nose.ts:76-98 noseRadius."""
    L, R = nose.length, nose.base_radius
    t = min(max(x / L, 0.0), 1.0)
    k = nose.kind
    if k == "conical":
        return R * t
    if k == "ogive":
        p = nose.parameter
        xx = t * L
        if p < OGIVE_CONE_LIMIT:
            return (R * xx) / L
        stretch, _, rho, xc, y0 = _ogive_arc(p, L, R)
        s = xx * stretch
        if s <= 0:
            return 0.0
        return (s * (2 * xc - s)) / (math.sqrt(max(0.0, rho * rho - (xc - s) ** 2)) + y0)
    if k == "ellipsoid":
        return R * math.sqrt(max(0.0, 1 - (1 - t) * (1 - t)))
    if k == "power":
        return R * t ** nose.parameter
    if k == "parabolic":
        kk = nose.parameter
        return (R * (2 * t - kk * t * t)) / (2 - kk)
    if k == "haack":
        theta = math.acos(1 - 2 * t)
        area = theta - math.sin(2 * theta) / 2 + nose.parameter * math.sin(theta) ** 3
        return R * math.sqrt(max(0.0, area) / math.pi)
    raise ValueError(f"unknown nose shape {k}")


def nose_tip_slope(nose: Nose) -> float:
    """nose.ts:145-164 noseTipSlope (inf for a blunt tip)."""
    L, R, k = nose.length, nose.base_radius, nose.kind
    if k == "conical":
        return R / L
    if k == "ogive":
        if nose.parameter < OGIVE_CONE_LIMIT:
            return R / L
        stretch, _, _, xc, y0 = _ogive_arc(nose.parameter, L, R)
        return math.inf if y0 == 0 else (xc / y0) * stretch
    if k == "parabolic":
        return (2 * R) / ((2 - nose.parameter) * L)
    if k == "power":
        return R / L if nose.parameter == 1 else math.inf
    return math.inf


def body_length(body: Body) -> float:
    return body.nose.length + body.cylinder_length


def body_radius(body: Body, x: float) -> float:
    """body.ts:24-27 bodyRadius."""
    if x <= 0:
        return 0.0
    return nose_radius(body.nose, x) if x < body.nose.length else body.nose.base_radius


def body_breakpoints(body: Body):
    """body.ts:29-31 bodyBreakpoints."""
    if body.cylinder_length > 0:
        return [0.0, body.nose.length, body_length(body)]
    return [0.0, body.nose.length]


# --------------------------------------------------------------------------
# Gas dynamics and Taylor-Maccoll (src/validation/gas-dynamics.ts,
# src/validation/taylor-maccoll.ts)
# --------------------------------------------------------------------------

def mach_angle(mach):
    if not mach >= 1:
        raise ValueError("the Mach angle needs M >= 1")
    return math.asin(1 / mach)


def deflection_angle(mach, beta, gamma):
    m2s2 = mach * mach * math.sin(beta) ** 2
    return math.atan((2 * (m2s2 - 1)) / (math.tan(beta) * (mach * mach * (gamma + math.cos(2 * beta)) + 2)))


def normal_shock(mach, gamma):
    if not mach >= 1:
        raise ValueError("a shock needs a supersonic normal Mach number")
    m2 = mach * mach
    return (1 + ((2 * gamma) / (gamma + 1)) * (m2 - 1),
            ((gamma + 1) * m2) / ((gamma - 1) * m2 + 2),
            math.sqrt(((gamma - 1) * m2 + 2) / (2 * gamma * m2 - (gamma - 1))))


def stagnation_pressure_ratio(mach, gamma):
    return (1 + 0.5 * (gamma - 1) * mach * mach) ** (gamma / (gamma - 1))


def _speed_from_mach(m, g):
    return 1 / math.sqrt(1 + 2 / ((g - 1) * m * m))


def _mach_from_speed(s, g):
    return math.sqrt(2 / ((g - 1) * (1 / (s * s) - 1)))


def _tm_derivatives(theta, vr, vt, g):
    a = 0.5 * (g - 1) * (1 - vr * vr - vt * vt)
    vtt = (vr * vt * vt - a * (2 * vr + vt / math.tan(theta))) / (a - vt * vt)
    return vt, vtt


def _rk4(theta, vr, vt, h, g):
    k1r, k1t = _tm_derivatives(theta, vr, vt, g)
    k2r, k2t = _tm_derivatives(theta + h / 2, vr + (h / 2) * k1r, vt + (h / 2) * k1t, g)
    k3r, k3t = _tm_derivatives(theta + h / 2, vr + (h / 2) * k2r, vt + (h / 2) * k2t, g)
    k4r, k4t = _tm_derivatives(theta + h, vr + h * k3r, vt + h * k3t, g)
    return (vr + (h / 6) * (k1r + 2 * k2r + 2 * k3r + k4r),
            vt + (h / 6) * (k1t + 2 * k2t + 2 * k3t + k4t))


def _integrate_to_cone(mach, beta, g, tol):
    """taylor-maccoll.ts:79-145 integrateToCone (adaptive RK4 with step
    doubling and Richardson extrapolation)."""
    mn1 = mach * math.sin(beta)
    _, _, m2n = normal_shock(mn1, g)
    defl = deflection_angle(mach, beta, g)
    mach2 = m2n / math.sin(beta - defl)
    speed = _speed_from_mach(mach2, g)
    vr = speed * math.cos(beta - defl)
    vt = -speed * math.sin(beta - defl)
    theta = beta
    h = -1e-4 * beta
    while theta > 1e-12:
        h = max(h, -0.5 * theta)
        full = _rk4(theta, vr, vt, h, g)
        halfway = _rk4(theta, vr, vt, h / 2, g)
        half = _rk4(theta + h / 2, halfway[0], halfway[1], h / 2, g)
        err = max(abs(half[0] - full[0]), abs(half[1] - full[1])) / 15
        if not err <= tol:
            h /= 2
            if abs(h) < 1e-14 * theta:
                return None
            continue
        nr = half[0] + (half[0] - full[0]) / 15
        nt = half[1] + (half[1] - full[1]) / 15
        if nt >= 0:
            if abs(h) > 1e-12:
                h /= 2
                continue
            f = vt / (vt - nt)
            return theta + f * h, vr + f * (nr - vr)
        theta += h
        vr, vt = nr, nt
        h *= min(2, 0.9 * (tol / max(err, 1e-300)) ** 0.2)
    return None


@dataclass(frozen=True)
class ConeFlow:
    shock_angle: float
    surface_pressure_ratio: float
    surface_mach: float
    surface_pressure_coefficient: float


def taylor_maccoll(mach, cone_angle, gamma, tolerance=1e-13) -> ConeFlow:
    """taylor-maccoll.ts:153-203 taylorMaccoll: bracket the shock angle from
    the Mach angle upward in 0.5 deg steps, then bisect."""
    if not cone_angle > 0:
        raise ValueError("the cone half-angle must be positive")
    mu = mach_angle(mach)

    def cone_at(beta):
        if beta <= mu:
            return 0.0
        r = _integrate_to_cone(mach, beta, gamma, tolerance)
        return math.nan if r is None else r[0]

    lo = hi = mu
    previous = 0.0
    step = 0.5 * (math.pi / 180)
    while True:
        hi += step
        c = cone_at(hi)
        if not hi < math.pi / 2 or math.isnan(c) or c < previous:
            raise ValueError("the shock detaches: no attached conical solution")
        if c >= cone_angle:
            break
        lo = hi
        previous = c
    k = 0
    while k < 60 and hi - lo > 1e-15:
        mid = 0.5 * (lo + hi)
        if cone_at(mid) < cone_angle:
            lo = mid
        else:
            hi = mid
        k += 1
    beta = 0.5 * (lo + hi)
    res = _integrate_to_cone(mach, beta, gamma, tolerance)
    if res is None:
        raise RuntimeError("integration failed to reach the cone")
    pr, _, m2n = normal_shock(mach * math.sin(beta), gamma)
    defl = deflection_angle(mach, beta, gamma)
    post = m2n / math.sin(beta - defl)
    surface_mach = _mach_from_speed(res[1], gamma)
    spr = pr * stagnation_pressure_ratio(post, gamma) / stagnation_pressure_ratio(surface_mach, gamma)
    return ConeFlow(beta, spr, surface_mach, (spr - 1) / (0.5 * gamma * mach * mach))


# --------------------------------------------------------------------------
# Body grids (src/grid/body-grid.ts, src/grid/spacing.ts, src/grid/quality.ts)
# --------------------------------------------------------------------------

QUALITY_LIMITS = dict(min_angle=12 * math.pi / 180, min_wall_angle=60 * math.pi / 180,
                      max_size_ratio=1.3, max_aspect_ratio=10.0)  # quality.ts:24-38
# quality.ts:64-68: the grids of a fitted bow shock, whose lines end on it.
FITTED_QUALITY_LIMITS = dict(QUALITY_LIMITS, min_wall_angle=50 * math.pi / 180, max_aspect_ratio=13.0)
OUTER_ANGLE_FACTOR = 1.2  # body-grid.ts:61
OUTER_CELLS_ABOVE_SHOCK = 3  # body-grid.ts:68
MAX_TIP_HALF_ANGLE = math.pi / 2 - QUALITY_LIMITS["min_wall_angle"]  # body-grid.ts:74


class GridQualityError(ValueError):
    pass


def column_positions(breakpoints, density, refinement):
    """spacing.ts:216-256 columnPositions (without the slope limit): each
    interval's tabulated density is integrated by the trapezoidal rule and
    cut into a whole number (a multiple of the refinement) of equal parts."""
    xs = [breakpoints[0]]
    for k in range(1, len(breakpoints)):
        a, b = breakpoints[k - 1], breakpoints[k]
        base_nodes = (((density(a) + density(b)) / 2) * (b - a)) / refinement
        samples = max(2048, 64 * math.ceil(base_nodes))
        at = [a + ((b - a) * n) / samples for n in range(samples + 1)]
        values = [density(t) for t in at]
        cum = [0.0] * (samples + 1)
        for n in range(1, samples + 1):
            cum[n] = cum[n - 1] + 0.5 * (values[n - 1] + values[n]) * (at[n] - at[n - 1])
        total = cum[samples]
        count = refinement * max(1, _js_round(total / refinement))
        n = 0
        for m in range(1, count):
            target = (total * m) / count
            while cum[n + 1] < target:
                n += 1
            f = (target - cum[n]) / (cum[n + 1] - cum[n])
            xs.append(at[n] + f * (at[n + 1] - at[n]))
        xs.append(b)
    return xs


def _js_round(v):
    """Math.round: halves round towards +infinity."""
    return int(math.floor(v + 0.5))


def grid_quality(grid: StructuredGrid):
    """quality.ts:70-125 gridQuality (vectorised)."""
    ni, nj = grid.ni, grid.nj
    X = grid.x.reshape(nj + 1, ni + 1)
    Y = grid.y.reshape(nj + 1, ni + 1)
    corners = [(X[:-1, :-1], Y[:-1, :-1]), (X[:-1, 1:], Y[:-1, 1:]), (X[1:, 1:], Y[1:, 1:]),
               (X[1:, :-1], Y[1:, :-1])]
    area = np.zeros((nj, ni))
    # NaN (from a degenerate cell's 0 / 0) propagates through np.min and
    # np.max, as through Math.min and Math.max, and then fails the gates.
    min_angle = math.pi
    min_wall = math.pi
    with np.errstate(invalid="ignore", divide="ignore"):
        for k in range(4):
            ax, ay = corners[k]
            bx, by = corners[(k + 1) % 4]
            px, py = corners[(k + 3) % 4]
            area += 0.5 * (ax * by - bx * ay)
            cosang = ((bx - ax) * (px - ax) + (by - ay) * (py - ay)) / (
                np.hypot(bx - ax, by - ay) * np.hypot(px - ax, py - ay))
            ang = np.arccos(np.clip(cosang, -1, 1))
            min_angle = float(np.min([min_angle, ang.min()]))
            min_wall = float(np.min([min_wall, ang[0].min()]))

        def side(a, b):
            return np.hypot(corners[b][0] - corners[a][0], corners[b][1] - corners[a][1])

        w = 0.5 * (side(0, 1) + side(3, 2))
        h = 0.5 * (side(0, 3) + side(1, 2))
        aspect = float(np.max([1.0, np.max(w / h), np.max(h / w)]))
        size = 1.0
        if ni > 1:
            size = float(np.max([size, np.max(np.maximum(w[:, 1:] / w[:, :-1], w[:, :-1] / w[:, 1:]))]))
        if nj > 1:
            size = float(np.max([size, np.max(np.maximum(h[1:] / h[:-1], h[:-1] / h[1:]))]))
    return dict(min_angle=min_angle, min_wall_angle=min_wall, max_size_ratio=size,
                max_aspect_ratio=aspect, min_area=float(area.min()))


def quality_failures(q, limits=QUALITY_LIMITS):
    """quality.ts:128-151 qualityFailures: the limits the grid fails, as the
    reference's messages; written as not(value within limit), so a NaN from
    a degenerate cell fails."""
    def deg(r):
        return f"{math.degrees(r):.1f} deg"

    f = []
    if not q["min_area"] > 0:
        f.append("a cell has non-positive area")
    if not q["min_angle"] >= limits["min_angle"]:
        f.append(f"cell angle {deg(q['min_angle'])} below {deg(limits['min_angle'])}")
    if not q["min_wall_angle"] >= limits["min_wall_angle"]:
        f.append(f"wall cell angle {deg(q['min_wall_angle'])} below {deg(limits['min_wall_angle'])}")
    if not q["max_size_ratio"] <= limits["max_size_ratio"]:
        f.append(f"neighbouring cells differ in size by {q['max_size_ratio']:.2f} "
                 f"(limit {limits['max_size_ratio']:g})")
    if not q["max_aspect_ratio"] <= limits["max_aspect_ratio"]:
        f.append(f"cell aspect ratio {q['max_aspect_ratio']:.1f} above {limits['max_aspect_ratio']:g}")
    return f


@dataclass(frozen=True)
class BodyGrid:
    grid: StructuredGrid
    i_tip: int
    outer_angle: float


def _outer_geometry(body: Body, mach: float, gamma: float, nj_base: int):
    """The outer angle and the base cell count across of body-grid.ts:86-155."""
    tip_slope = nose_tip_slope(body.nose)
    if not math.isfinite(tip_slope):
        raise ValueError("blunt noses (profile vertical at the tip) need a C grid")
    tip_half = math.atan(tip_slope)
    if tip_half > MAX_TIP_HALF_ANGLE:
        raise ValueError(f"tip half-angles above {math.degrees(MAX_TIP_HALF_ANGLE):.1f} deg are not supported")
    shock = taylor_maccoll(mach, tip_half, gamma).shock_angle
    max_outer = (shock + math.pi / 2) / 2
    R = body.nose.base_radius
    x_end = body_length(body)
    x_start = -R
    max_strength = 2 * math.log(QUALITY_LIMITS["max_aspect_ratio"])

    def across_for(angle):
        def across(x):
            wall = body_radius(body, x)
            h = 0.5 * R + (x - x_start) * math.tan(angle) - wall
            return wall, h, min(math.log1p(h / R), max_strength)
        return across

    def fraction_above_shock(angle):
        across = across_for(angle)
        samples = body_breakpoints(body)[1:]
        samples += [(x_end * k) / 1024 for k in range(1, 1024)]
        fewest = math.inf
        for x in samples:
            wall, h, b = across(x)
            eta = math.log1p((math.expm1(b) * (x * math.tan(shock) - wall)) / h) / b
            fewest = min(fewest, 1 - eta)
        return fewest

    base_nj = max(nj_base, math.ceil(OUTER_CELLS_ABOVE_SHOCK / fraction_above_shock(max_outer) - 1e-9))

    def cells_above(angle):
        return base_nj * fraction_above_shock(angle)

    outer = min(OUTER_ANGLE_FACTOR * shock, max_outer)
    if cells_above(outer) < OUTER_CELLS_ABOVE_SHOCK:
        low, high = outer, max_outer
        for _ in range(50):
            mid = (low + high) / 2
            if cells_above(mid) < OUTER_CELLS_ABOVE_SHOCK:
                low = mid
            else:
                high = mid
        outer = high
    return outer, base_nj, across_for(outer), x_start


def body_grid(body: Body, nj: int, mach: float, gamma: float, refinement: int = 1,
              check_quality: bool = True) -> BodyGrid:
    """body-grid.ts:76-181 bodyGrid: H grid around a sharp-nosed body, the
    outer boundary a straight line at 1.2 times the tip's conical shock
    angle, rows stretched towards the body, columns placed by
    ``column_positions`` at the spacing that balances the aspect ratio."""
    if not (isinstance(nj, int) and nj >= 4):
        raise ValueError("a body grid needs at least 4 cells across")
    if refinement < 1 or (refinement & (refinement - 1)):
        raise ValueError("the refinement must be 1, 2, 4, 8, ...")
    outer, base_nj, across, x_start = _outer_geometry(body, mach, gamma, nj)
    njr = base_nj * refinement

    def density(x):
        _, h, b = across(x)
        return 1.0 / ((h * b * math.exp(b / 2)) / (math.expm1(b) * njr))

    xs = column_positions([x_start] + body_breakpoints(body), density, refinement)
    ni = len(xs) - 1
    i_tip = xs.index(0.0)
    grid = _h_grid(xs, njr, across)
    if check_quality:
        fails = quality_failures(grid_quality(grid))
        if fails:
            raise GridQualityError("the generated grid fails its quality checks: " + "; ".join(fails))
    return BodyGrid(grid, i_tip, outer)


def _h_grid(xs, nj, across) -> StructuredGrid:
    ni = len(xs) - 1
    x = np.empty((nj + 1) * (ni + 1))
    y = np.empty_like(x)
    for i, xi in enumerate(xs):
        wall, h, b = across(xi)
        em = math.expm1(b)
        for j in range(nj + 1):
            x[i + j * (ni + 1)] = xi
            y[i + j * (ni + 1)] = wall + (h * math.expm1((b * j) / nj)) / em
    return StructuredGrid(ni, nj, x, y)


def body_grid_fixed(body: Body, ni: int, nj: int, mach: float, gamma: float) -> BodyGrid:
    """An H grid of exactly ni x nj cells with bodyGrid's topology, outer
    boundary and row stretching (body-grid.ts:76-181). The columns follow
    the same target density as bodyGrid; the ni cells are shared among the
    intervals between breakpoints (inflow plane, tip, shoulder, end) in
    proportion to each interval's integrated density, at least one each,
    and placed at equal parts of the cumulative density within each
    interval, so the tip and shoulder lie on grid lines. New: the reference
    derives ni from the geometry instead of taking it."""
    outer, _, across, x_start = _outer_geometry(body, mach, gamma, nj)

    def density(x):
        _, h, b = across(x)
        return (math.expm1(b) * nj) / (h * b * math.exp(b / 2))

    bps = [x_start] + body_breakpoints(body)
    tables = []
    for k in range(1, len(bps)):
        a, b = bps[k - 1], bps[k]
        at = np.linspace(a, b, 4097)
        vals = np.array([density(t) for t in at])
        cum = np.concatenate([[0.0], np.cumsum(0.5 * (vals[1:] + vals[:-1]) * np.diff(at))])
        tables.append((a, b, at, cum))
    totals = np.array([t[3][-1] for t in tables])
    share = totals / totals.sum() * ni
    counts = np.maximum(1, np.floor(share).astype(int))
    while counts.sum() < ni:
        counts[np.argmax(share - counts)] += 1
    while counts.sum() > ni:
        counts[np.argmax(counts - share)] -= 1
    xs = [bps[0]]
    for (a, b, at, cum), count in zip(tables, counts):
        targets = cum[-1] * np.arange(1, count) / count
        xs.extend(np.interp(targets, cum, at).tolist())
        xs.append(b)
    grid = _h_grid(xs, nj, across)
    return BodyGrid(grid, xs.index(0.0), outer)


def cone_cylinder(half_angle_deg: float, cylinder_calibers: float, base_radius: float = 0.5) -> Body:
    """A sharp cone of the given half-angle on a cylinder (lengths in body
    diameters for base radius 0.5)."""
    L = base_radius / math.tan(math.radians(half_angle_deg))
    return Body(Nose("conical", L, base_radius), cylinder_calibers * 2 * base_radius)


def ogive_cylinder(fineness: float, cylinder_calibers: float, base_radius: float = 0.5) -> Body:
    """A tangent ogive (parameter 1) of the given length in calibers on a
    cylinder."""
    return Body(Nose("ogive", fineness * 2 * base_radius, base_radius, 1.0),
                cylinder_calibers * 2 * base_radius)


def free_stream_at(mach: float, gamma: float):
    """verified-drag.ts:225-228 freeStreamAt: rho = p = 1 and the speed
    rounded to 32 bits; returns (state, pressure, dynamic pressure)."""
    u = float(np.float32(mach * math.sqrt(gamma)))
    return (1.0, u, 0.0, 1.0), 1.0, 0.5 * u * u


# --------------------------------------------------------------------------
# Vectorized generator for populations of bodies (the formulas of
# body_grid_fixed above, evaluated on numpy arrays).
# --------------------------------------------------------------------------

def nose_radius_array(nose: Nose, x) -> np.ndarray:
    """noseRadius (nose.ts:76-98) at many stations."""
    L, R = nose.length, nose.base_radius
    t = np.clip(np.asarray(x, dtype=np.float64) / L, 0.0, 1.0)
    k = nose.kind
    if k == "conical" or (k == "ogive" and nose.parameter < OGIVE_CONE_LIMIT):
        return R * t
    if k == "ogive":
        stretch, _, rho, xc, y0 = _ogive_arc(nose.parameter, L, R)
        s = t * L * stretch
        r = (s * (2 * xc - s)) / (np.sqrt(np.maximum(0.0, rho * rho - (xc - s) ** 2)) + y0)
        return np.where(s <= 0, 0.0, r)
    if k == "ellipsoid":
        return R * np.sqrt(np.maximum(0.0, 1 - (1 - t) * (1 - t)))
    if k == "power":
        return R * t ** nose.parameter
    if k == "parabolic":
        kk = nose.parameter
        return (R * (2 * t - kk * t * t)) / (2 - kk)
    if k == "haack":
        theta = np.arccos(1 - 2 * t)
        area = theta - np.sin(2 * theta) / 2 + nose.parameter * np.sin(theta) ** 3
        return R * np.sqrt(np.maximum(0.0, area) / math.pi)
    raise ValueError(f"unknown nose shape {k}")


def body_radius_array(body: Body, x) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    r = nose_radius_array(body.nose, x)
    return np.where(x <= 0, 0.0, np.where(x < body.nose.length, r, body.nose.base_radius))


def _outer_geometry_array(body: Body, shock: float, nj_base: int):
    R = body.nose.base_radius
    x_end = body_length(body)
    x_start = -R
    max_strength = 2 * math.log(QUALITY_LIMITS["max_aspect_ratio"])
    max_outer = (shock + math.pi / 2) / 2
    samples = np.array(body_breakpoints(body)[1:] + [(x_end * k) / 1024 for k in range(1, 1024)])
    wall_s = body_radius_array(body, samples)

    def across_for(angle):
        def across(x):
            wall = body_radius_array(body, x)
            h = 0.5 * R + (np.asarray(x) - x_start) * math.tan(angle) - wall
            return wall, h, np.minimum(np.log1p(h / R), max_strength)
        return across

    def fraction_above_shock(angle):
        h = 0.5 * R + (samples - x_start) * math.tan(angle) - wall_s
        b = np.minimum(np.log1p(h / R), max_strength)
        eta = np.log1p((np.expm1(b) * (samples * math.tan(shock) - wall_s)) / h) / b
        return float(np.min(1 - eta))

    base_nj = max(nj_base, math.ceil(OUTER_CELLS_ABOVE_SHOCK / fraction_above_shock(max_outer) - 1e-9))
    outer = min(OUTER_ANGLE_FACTOR * shock, max_outer)
    if base_nj * fraction_above_shock(outer) < OUTER_CELLS_ABOVE_SHOCK:
        low, high = outer, max_outer
        for _ in range(50):
            mid = (low + high) / 2
            if base_nj * fraction_above_shock(mid) < OUTER_CELLS_ABOVE_SHOCK:
                low = mid
            else:
                high = mid
        outer = high
    return outer, across_for(outer), x_start


def tip_shock_angle(body: Body, mach: float, gamma: float, tolerance: float = 1e-13) -> float:
    """The conical shock angle at the nose tip (body-grid.ts:86-100):
    raises ValueError for a blunt tip, a tip half-angle above the grid's
    limit, or a detached shock."""
    tip_slope = nose_tip_slope(body.nose)
    if not math.isfinite(tip_slope):
        raise ValueError("blunt noses (profile vertical at the tip) need a C grid")
    tip_half = math.atan(tip_slope)
    if tip_half > MAX_TIP_HALF_ANGLE:
        raise ValueError(f"tip half-angles above {math.degrees(MAX_TIP_HALF_ANGLE):.1f} deg are not supported")
    return taylor_maccoll(mach, tip_half, gamma, tolerance).shock_angle


def body_grid_array(body: Body, ni: int, nj: int, shock: float) -> BodyGrid:
    """body_grid_fixed on numpy arrays, given the tip shock angle: the
    same outer boundary, row stretching and column placement."""
    outer, across, x_start = _outer_geometry_array(body, shock, nj)
    bps = [x_start] + body_breakpoints(body)
    tables = []
    for k in range(1, len(bps)):
        a, b = bps[k - 1], bps[k]
        at = np.linspace(a, b, 4097)
        _, h, bb = across(at)
        vals = (np.expm1(bb) * nj) / (h * bb * np.exp(bb / 2))
        cum = np.concatenate([[0.0], np.cumsum(0.5 * (vals[1:] + vals[:-1]) * np.diff(at))])
        tables.append((a, b, at, cum))
    totals = np.array([t[3][-1] for t in tables])
    share = totals / totals.sum() * ni
    counts = np.maximum(1, np.floor(share).astype(int))
    while counts.sum() < ni:
        counts[np.argmax(share - counts)] += 1
    while counts.sum() > ni:
        counts[np.argmax(counts - share)] -= 1
    xs = [bps[0]]
    for (a, b, at, cum), count in zip(tables, counts):
        targets = cum[-1] * np.arange(1, count) / count
        xs.extend(np.interp(targets, cum, at).tolist())
        xs.append(b)
    xa = np.array(xs)
    wall, h, bb = across(xa)
    jj = np.arange(nj + 1)[:, None] / nj
    y = wall[None, :] + (h[None, :] * np.expm1(bb[None, :] * jj)) / np.expm1(bb)[None, :]
    x = np.broadcast_to(xa[None, :], y.shape)
    grid = StructuredGrid(ni, nj, np.ascontiguousarray(x).ravel(), np.ascontiguousarray(y).ravel())
    return BodyGrid(grid, xs.index(0.0), outer)
