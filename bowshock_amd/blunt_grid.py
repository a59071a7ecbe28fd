"""Structured C grids around blunt-nosed axisymmetric bodies, the grids the
fitted shock moves on: src/grid/blunt-grid.ts (bluntBodyGrid,
gridOnLines, withFixedFractions, outerCellHeight, billigStandoff).

Topology: the lines of constant i are straight lines from the body, close
to its normals, fanning out from the nose and vertical along the
cylinder; jMin is the body from the tip to its end, jMax the outer
boundary, iMin the axis ahead of the tip, iMax the outflow plane. The
nodes along each line are stretched towards the body; the lines are
spaced along the body by the density of src/grid/spacing.ts
columnPositions.

The reference evaluates everything one point at a time; here every stage
is evaluated on arrays of arc lengths (the window average of the normal,
the bisection for each line's length to the outer boundary, the
tabulations), with the operations in the reference's order, so the grids
of a family nest exactly as the reference's do: the density of lines and
the node fractions scale by powers of two with the refinement.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from .grids import (QUALITY_LIMITS, Body, GridQualityError, StructuredGrid, grid_quality, nose_radius,
                    nose_radius_array, nose_tip_slope, quality_failures, _js_round)

# blunt-grid.ts:138-145 DEFAULT_TUNING.
DEFAULT_TUNING = dict(standoff_factor=2.0, arc_radius_factor=2.0, tip_scale_factor=2.0, smoothing=0.25,
                      fan_ratio=1.25, max_width_slope=0.5)
MAX_NOSE_RADIUS = 10                     # blunt-grid.ts:148
MAX_SHOULDER_ANGLE = 15 * math.pi / 180  # blunt-grid.ts:155
WIDTH_SAMPLES = 4096                     # blunt-grid.ts:157
UNIFORM_WIDTH_SAMPLES = 1024             # blunt-grid.ts:159
SPACING_SAMPLES = 8192                   # blunt-grid.ts:161
WINDOW_SAMPLES = 64                      # blunt-grid.ts:163
PROFILE_SAMPLES = 20000                  # blunt-grid.ts:165


def billig_standoff(mach: float, radius: float) -> float:
    """blunt-grid.ts:105-107: Billig's correlation of a sphere's shock
    standoff, 0.143 exp(3.24 / M^2) r (J. Spacecraft Rockets 4, 1967)."""
    return 0.143 * math.exp(3.24 / (mach * mach)) * radius


def line_strength(h, radius: float):
    """blunt-grid.ts:548-550: stretching strength ln(1 + h / R), capped at
    2 ln(maximum aspect ratio)."""
    return np.minimum(np.log1p(np.asarray(h, dtype=np.float64) / radius),
                      2 * math.log(QUALITY_LIMITS["max_aspect_ratio"]))


def stretched_fraction(b, j, nj: int):
    """blunt-grid.ts:553-555: (e^(b j / nj) - 1) / (e^b - 1)."""
    b = np.asarray(b, dtype=np.float64)
    return np.expm1((b * j) / nj) / np.expm1(b)


def outer_cell_height(h: float, nj: int, base_radius: float) -> float:
    """blunt-grid.ts:558-560: height of the outermost of nj cells."""
    return float(h * (1 - stretched_fraction(line_strength(h, base_radius), nj - 1, nj)))


@dataclass
class LineSet:
    """The lines of constant i (blunt-grid.ts GridLine), as arrays: body
    point (x, y), unit direction (dx, dy), height h to the outer boundary
    and, when fixed, the node fractions (nj + 1, n)."""

    x: np.ndarray
    y: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    h: np.ndarray
    fractions: Optional[np.ndarray] = None

    def __len__(self):
        return len(self.x)

    def origins_and_directions(self) -> np.ndarray:
        """(n, 4): x, y, dx, dy per line, the layout of bs_set_line_grid."""
        return np.stack([self.x, self.y, self.dx, self.dy], axis=1)

    def with_heights(self, h) -> "LineSet":
        return LineSet(self.x, self.y, self.dx, self.dy, np.asarray(h, dtype=np.float64).copy(), self.fractions)


def with_fixed_fractions(lines: LineSet, nj: int, base_radius: float) -> LineSet:
    """blunt-grid.ts:570-583: each line's node fractions fixed at those of
    its current height."""
    b = line_strength(lines.h, base_radius)
    j = np.arange(nj + 1, dtype=np.float64)[:, None]
    return LineSet(lines.x, lines.y, lines.dx, lines.dy, lines.h, stretched_fraction(b[None, :], j, nj))


def assert_grid_quality(grid: StructuredGrid, limits=QUALITY_LIMITS) -> None:
    """blunt-grid.ts:622-632 assertGridQuality."""
    failures = quality_failures(grid_quality(grid), limits)
    if failures:
        raise GridQualityError("the generated grid fails its quality checks: " + "; ".join(failures))


def grid_on_lines(lines: LineSet, heights, nj: int, base_radius: float, gated: bool = True) -> StructuredGrid:
    """blunt-grid.ts:591-613 gridOnLines: node j of line i at the body point
    plus h_i times the node's fraction (fixed, or stretched for h_i)."""
    h = np.asarray(heights, dtype=np.float64)
    if len(h) != len(lines):
        raise ValueError("one height per line is needed")
    if lines.fractions is not None:
        frac = lines.fractions
        if frac.shape != (nj + 1, len(h)):
            raise ValueError(f"fixed fractions for {frac.shape[0] - 1} cells across, not {nj}")
    else:
        j = np.arange(nj + 1, dtype=np.float64)[:, None]
        frac = stretched_fraction(line_strength(h, base_radius)[None, :], j, nj)
    d = h[None, :] * frac
    x = lines.x[None, :] + d * lines.dx[None, :]
    y = lines.y[None, :] + d * lines.dy[None, :]
    grid = StructuredGrid(len(h) - 1, nj, x.reshape(-1), y.reshape(-1))
    if gated:
        assert_grid_quality(grid)
    return grid


def column_positions_array(breakpoints: Sequence[float], density, refinement: int,
                           max_spacing_slope: Optional[float] = None):
    """spacing.ts:31-71 columnPositions with the density evaluated on
    arrays, and the optional limit on the spacing's slope (spacing.ts:80-118
    limitSpacingSlope, applied to the base grid's spacing)."""
    tables = []
    for k in range(1, len(breakpoints)):
        a, b = float(breakpoints[k - 1]), float(breakpoints[k])
        ends = density(np.array([a, b]))
        base_nodes = (((ends[0] + ends[1]) / 2) * (b - a)) / refinement
        samples = max(2048, 64 * math.ceil(base_nodes))
        at = a + ((b - a) * np.arange(samples + 1, dtype=np.float64)) / samples
        tables.append((b, samples, at, np.array(density(at), dtype=np.float64)))
    if max_spacing_slope is not None:
        if not max_spacing_slope > 0:
            raise ValueError("the spacing slope limit must be positive")
        xs = np.concatenate([t[2] for t in tables]).tolist()
        spacings = np.concatenate([refinement / t[3] for t in tables]).tolist()
        limited = list(spacings)
        for k in range(1, len(limited)):
            limited[k] = min(limited[k], limited[k - 1] + max_spacing_slope * (xs[k] - xs[k - 1]))
        for k in range(len(limited) - 2, -1, -1):
            limited[k] = min(limited[k], limited[k + 1] + max_spacing_slope * (xs[k + 1] - xs[k]))
        k = 0
        for _, samples, _, values in tables:
            lim = np.array(limited[k:k + samples + 1])
            sp = np.array(spacings[k:k + samples + 1])
            values[:] = np.where(lim < sp, refinement / lim, values)
            k += samples + 1
    out = [float(breakpoints[0])]
    for b, samples, at, values in tables:
        cum = np.concatenate([[0.0], np.cumsum(0.5 * (values[:-1] + values[1:]) * (at[1:] - at[:-1]))])
        total = cum[samples]
        count = refinement * max(1, _js_round(total / refinement))
        m = np.arange(1, count, dtype=np.float64)
        target = (total * m) / count
        n = np.searchsorted(cum, target, side="left") - 1
        f = (target - cum[n]) / (cum[n + 1] - cum[n])
        out.extend((at[n] + f * (at[n + 1] - at[n])).tolist())
        out.append(b)
    return out


@dataclass
class BluntGrid:
    """blunt-grid.ts:70-98 BluntGrid."""

    grid: StructuredGrid
    lines: LineSet
    outer_standoff: float
    outer_angle: float
    tip_scale: float
    scale: float
    two_scale: bool


class _Nose:
    """The body's profile by arc length and the outer boundary: the closures
    of bluntBodyGrid (blunt-grid.ts:187-516), on arrays."""

    def __init__(self, body: Body, mach: float, nj_base: int, tuning: dict, outer_boundary):
        nose = body.nose
        self.nose = nose
        self.L, self.R = nose.length, nose.base_radius
        self.tuning = tuning
        N = PROFILE_SAMPLES
        u = np.arange(N + 1, dtype=np.float64) / N
        px = self._profile_x(u)
        pr = nose_radius_array(nose, px)
        self.u_of = u
        self.s_of = np.concatenate([[0.0], np.cumsum(np.hypot(px[1:] - px[:-1], pr[1:] - pr[:-1]))])
        self.s_nose = float(self.s_of[N])
        self.s_end = self.s_nose + body.cylinder_length
        self.sharp = math.isfinite(nose_tip_slope(nose))
        self.ellipsoid = nose.kind == "ellipsoid"

        sx, sy, snx, sny = self.surface(np.array([self.s_nose * (1 - 1e-9)]))
        shoulder = math.atan2(-snx[0], sny[0])
        if not shoulder <= MAX_SHOULDER_ANGLE:
            raise ValueError(f"blunt noses meeting the body at more than {math.degrees(MAX_SHOULDER_ANGLE):.0f} deg "
                             f"are not supported yet, got {math.degrees(shoulder):.1f} deg")
        L, R = self.L, self.R
        x_small = 1e-6 * L
        tip_radius = nose_radius(nose, x_small) ** 2 / (2 * x_small)
        if not tip_radius <= MAX_NOSE_RADIUS * R:
            raise ValueError("noses that are nearly flat at the tip are not supported yet")
        if self.sharp:
            tip_scale = 0.0
        elif nose_radius(nose, L) > L:
            tip_scale = L
        else:
            low, high = 0.0, L
            for _ in range(200):
                mid = 0.5 * (low + high)
                if nose_radius(nose, mid) > mid:
                    low = mid
                else:
                    high = mid
            tip_scale = 0.5 * (low + high)
        self.tip_scale = tip_scale
        if self.ellipsoid:
            scale = max(tip_radius, R)
        elif self.sharp:
            scale = R
        else:
            scale = max(tip_radius, min(R, tuning["tip_scale_factor"] * tip_scale))
        self.scale = scale

        self.outer_angle = (math.pi / 2 + math.asin(1 / mach)) / 2
        self.arc_radius = tuning["arc_radius_factor"] * scale
        self.custom = None
        if outer_boundary is not None:
            cb = np.asarray(outer_boundary, dtype=np.float64).reshape(-1, 2)
            if len(cb) < 2 or cb[0, 1] != 0:
                raise ValueError("the outer boundary needs at least two points, the first on the axis")
            if not np.all(cb[1:, 1] > cb[:-1, 1]):
                raise ValueError("the outer boundary must be ordered outwards from the axis")
            self.custom = cb
            self.outer_standoff = -float(cb[0, 0])
        else:
            self.outer_standoff = tuning["standoff_factor"] * billig_standoff(mach, scale)
        self.centre = self.arc_radius - self.outer_standoff
        self.y_join = self.arc_radius * math.cos(self.outer_angle)
        self.x_join = self.centre - self.arc_radius * math.sin(self.outer_angle)

        self.half_width = self._uniform_half_width() if self.ellipsoid else self._clustered_half_width()
        if self.sharp:
            self.tip_jump = float(self.normal_angle(np.array([self.s_nose * 1e-9]))[0])
            h = float(self.length_along(np.array([0.0]), np.array([-1.0]), np.array([0.0]))[0])
            c = 1 / math.log(tuning["fan_ratio"]) / nj_base
            self.fan_scale = (h * (-self.tip_jump + math.sqrt(self.tip_jump * self.tip_jump + 4 * c * c))) / 2
        else:
            self.tip_jump = 0.0
            self.fan_scale = math.inf

    def _profile_x(self, u):
        return self.L * u * u

    def u_at(self, s):
        N = PROFILE_SAMPLES
        low = np.clip(np.searchsorted(self.s_of, s, side="right") - 1, 0, N - 1)
        f = (s - self.s_of[low]) / (self.s_of[low + 1] - self.s_of[low])
        return self.u_of[low] + np.clip(f, 0.0, 1.0) / N

    def surface(self, s):
        """Body point at arc length s and its outward normal (blunt-grid.ts:213-227)."""
        s = np.asarray(s, dtype=np.float64)
        u = self.u_at(s)
        x = self._profile_x(u)
        y = nose_radius_array(self.nose, x)
        e = 1e-7
        ua = np.maximum(u - e, 0.0)
        ub = np.minimum(u + e, 1.0)
        tx = self._profile_x(ub) - self._profile_x(ua)
        ty = nose_radius_array(self.nose, self._profile_x(ub)) - nose_radius_array(self.nose, self._profile_x(ua))
        t = np.hypot(tx, ty)
        with np.errstate(invalid="ignore", divide="ignore"):
            nx = np.where(u == 0, -1.0, -ty / t)
            ny = np.where(u == 0, 0.0, tx / t)
        on_cyl = s >= self.s_nose
        return (np.where(on_cyl, self.L + (s - self.s_nose), x), np.where(on_cyl, self.R, y),
                np.where(on_cyl, 0.0, nx), np.where(on_cyl, 1.0, ny))

    def outer_x(self, y):
        """Axial position of the outer boundary at height y (blunt-grid.ts:312-333)."""
        y = np.asarray(y, dtype=np.float64)
        if self.custom is None:
            r = self.arc_radius
            arc = self.centre - np.sqrt(np.maximum(r * r - y * y, 0.0))
            return np.where(y <= self.y_join, arc, self.x_join + (y - self.y_join) / math.tan(self.outer_angle))
        cb = self.custom
        low = np.clip(np.searchsorted(cb[:, 1], y, side="right") - 1, 0, len(cb) - 2)
        ax, ay, bx, by = cb[low, 0], cb[low, 1], cb[low + 1, 0], cb[low + 1, 1]
        return ax + ((y - ay) / (by - ay)) * (bx - ax)

    def length_along(self, s, dx, dy):
        """Distance from the body point at s to the outer boundary along
        (dx, dy): doubling from R, then 60 bisections (blunt-grid.ts:336-351)."""
        px, py, _, _ = self.surface(s)

        def outside(d):
            return px + d * dx - self.outer_x(py + d * dy) <= 0

        if np.any(outside(np.zeros_like(px))):
            raise ValueError("the body reaches the outer boundary; the nose is not supported yet")
        high = np.full_like(px, self.R)
        while True:
            more = ~outside(high)
            if not more.any():
                break
            high = np.where(more, high * 2, high)
        low = np.zeros_like(px)
        for _ in range(60):
            mid = (low + high) / 2
            o = outside(mid)
            high = np.where(o, mid, high)
            low = np.where(o, low, mid)
        return high

    def width_at(self, s):
        _, _, nx, ny = self.surface(s)
        return np.minimum(self.tuning["smoothing"] * self.length_along(s, nx, ny), self.s_end - s)

    def _limit(self, ws, widths):
        slope = self.tuning["max_width_slope"]
        w = widths.tolist()
        for k in range(1, len(w)):
            w[k] = min(w[k], w[k - 1] + slope * (ws[k] - ws[k - 1]))
        for k in range(len(w) - 2, -1, -1):
            w[k] = min(w[k], w[k + 1] + slope * (ws[k + 1] - ws[k]))
        return np.array(w)

    def _uniform_half_width(self):
        """blunt-grid.ts:363-380."""
        n = UNIFORM_WIDTH_SAMPLES
        step = self.s_end / n
        k = np.arange(n + 1, dtype=np.float64)
        ws = k * step
        widths = self.width_at(ws)
        w = widths.tolist()
        for m in range(1, n + 1):
            w[m] = min(w[m], w[m - 1] + self.tuning["max_width_slope"] * step)
        for m in range(n - 1, -1, -1):
            w[m] = min(w[m], w[m + 1] + self.tuning["max_width_slope"] * step)
        w = np.array(w)

        def half_width(s):
            t = np.clip(s / step, 0, n)
            m = np.minimum(np.floor(t), n - 1).astype(np.int64)
            return w[m] + (t - m) * (w[m + 1] - w[m])

        return half_width

    def _clustered_half_width(self):
        """blunt-grid.ts:386-418."""
        n = WIDTH_SAMPLES
        k = np.arange(n + 1, dtype=np.float64)
        ws = (self.s_end * k * k) / (n * n)
        w = self._limit(ws.tolist(), self.width_at(ws))
        s_end = self.s_end

        def half_width(s):
            t = np.clip(np.sqrt(np.maximum(s, 0) / s_end) * n, 0, n)
            m = np.minimum(np.floor(t), n - 1).astype(np.int64)
            f = (np.clip(s, 0, s_end) - ws[m]) / (ws[m + 1] - ws[m])
            return w[m] + np.clip(f, 0, 1) * (w[m + 1] - w[m])

        return half_width

    def normal_angle(self, s):
        """Angle of the body normal to the upstream axis, odd about the tip
        and pi / 2 beyond the nose (blunt-grid.ts:423-428)."""
        s = np.asarray(s, dtype=np.float64)
        a = np.abs(s)
        ang = np.full_like(a, math.pi / 2)
        nose = a < self.s_nose
        if nose.any():
            _, _, nx, ny = self.surface(a[nose])
            ang[nose] = np.arctan2(ny, -nx)
        return np.where(s < 0, -ang, ang)

    def fan(self, s):
        return self.tip_jump * -np.expm1(-s / self.fan_scale)

    def line_angle(self, s):
        """The window average of the normal's angle less the tip's jump,
        plus the fan (blunt-grid.ts:470-482), odd about the tip."""
        s = np.asarray(s, dtype=np.float64)
        a = np.abs(s)
        w = self.half_width(a)
        k = np.arange(WINDOW_SAMPLES, dtype=np.float64)[None, :]
        t = a[:, None] - w[:, None] + (2 * w[:, None] * (k + 0.5)) / WINDOW_SAMPLES
        vals = self.normal_angle(t.reshape(-1)).reshape(t.shape) - np.sign(t) * self.tip_jump
        total = np.zeros_like(a)
        for m in range(WINDOW_SAMPLES):
            total = total + vals[:, m]
        ang = total / WINDOW_SAMPLES + self.fan(a)
        ang = np.where(a >= self.s_end, math.pi / 2, ang)
        ang = np.where(a == 0, 0.0, ang)
        return np.where(s < 0, -ang, ang)

    def turn_rate(self, s):
        """blunt-grid.ts:484-494."""
        e = np.maximum(np.minimum(self.half_width(s) / 4, np.maximum(self.fan_scale / 4, s / 8)),
                       self.s_end / PROFILE_SAMPLES)
        return (self.line_angle(s + e) - self.line_angle(s - e)) / (2 * e)

    def line(self, s):
        angle = self.line_angle(s)
        dx = -np.cos(angle)
        dy = np.sin(angle)
        return dx, dy, self.length_along(s, dx, dy)

    def spacing(self, s, nj: int):
        """blunt-grid.ts:508-513: the wall cells as long as the geometric mean
        of the wall and outer cell heights, allowing for the lines' spread."""
        s = np.asarray(s, dtype=np.float64)
        _, _, h = self.line(s)
        b = line_strength(h, self.R)
        wall_height = (h * b) / (np.expm1(b) * nj)
        return (wall_height * np.exp(b / 2)) / np.sqrt(1 + self.turn_rate(s) * h)

    def tabulated_density(self, nj: int):
        """blunt-grid.ts:519-536."""
        n = SPACING_SAMPLES
        k = np.arange(n + 1, dtype=np.float64)
        ds = (self.s_end * k * k) / (n * n)
        dens = 1 / self.spacing(ds, nj)
        s_end = self.s_end

        def density(s):
            s = np.asarray(s, dtype=np.float64)
            t = np.clip(np.sqrt(np.maximum(s, 0) / s_end) * n, 0, n)
            m = np.minimum(np.floor(t), n - 1).astype(np.int64)
            f = (np.clip(s, 0, s_end) - ds[m]) / (ds[m + 1] - ds[m])
            return dens[m] + np.clip(f, 0, 1) * (dens[m + 1] - dens[m])

        return density


def blunt_body_grid(body: Body, nj: int, mach: float, refinement: int = 1, outer_boundary=None,
                    gated: bool = True, tuning: Optional[dict] = None) -> BluntGrid:
    """blunt-grid.ts:167-540 bluntBodyGrid: the C grid around a blunt (or
    sharp, as the limit of a two-scale) nose for a supersonic free stream.
    `outer_boundary`: points (x, y) ordered outwards from the axis, the
    first on it, in place of the generator's arc and line."""
    tun = dict(DEFAULT_TUNING, **(tuning or {}))
    if not (isinstance(nj, int) and nj >= 4):
        raise ValueError("a body grid needs at least 4 cells across")
    if not (isinstance(refinement, int) and refinement >= 1 and refinement & (refinement - 1) == 0):
        raise ValueError("the refinement must be 1, 2, 4, 8, ...")
    if not mach > 1:
        raise ValueError("blunt-body grids are for supersonic free streams")
    if not (body.nose.length > 0 and body.nose.base_radius > 0 and body.cylinder_length >= 0):
        raise ValueError("the body needs a positive nose length and radius")
    p = _Nose(body, mach, nj, tun, outer_boundary)
    nj_fine = nj * refinement
    breakpoints = [0.0, p.s_nose, p.s_end] if body.cylinder_length > 0 else [0.0, p.s_nose]
    if p.ellipsoid:
        ss = column_positions_array(breakpoints, lambda s: 1 / p.spacing(s, nj_fine), refinement,
                                    max_spacing_slope=0.15)
    else:
        ss = column_positions_array(breakpoints, p.tabulated_density(nj_fine), refinement)
    ss = np.array(ss)
    x, y, _, _ = p.surface(ss)
    dx, dy, h = p.line(ss)
    lines = LineSet(x, y, dx, dy, h)
    grid = grid_on_lines(lines, h, nj_fine, p.R, gated)
    return BluntGrid(grid, lines, p.outer_standoff, p.outer_angle, p.tip_scale, p.scale,
                     bool(p.sharp or p.scale < p.R))
