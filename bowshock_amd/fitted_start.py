"""The start of a fitted bow shock from a captured solution on the same
lines (src/analysis/fitted-flow.ts bluntStart and shockLayerStart,
simplified): the captured shock's height on each line where the pressure
crosses half its rise, smoothed; the lines' node fractions fixed at those
heights (withFixedFractions); the flow of each fitted cell taken from the
captured column at the same distance along its line.

The reference pushes the captured shock out by a margin to build a second
captured grid and families of fitted grids on it; here the fitted grid is
built on the captured grid's own lines, the form the HIP engine's moving
grid (bs_set_line_grid) takes directly.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .blunt_grid import LineSet, grid_on_lines, with_fixed_fractions
from .grids import StructuredGrid


@dataclass
class FittedStart:
    lines: LineSet          # fixed fractions at `heights`
    heights: np.ndarray     # (ni + 1,)
    grid: StructuredGrid    # the fitted grid at those heights
    primitives: np.ndarray  # (ni nj, 4) start on it


def _centres(grid: StructuredGrid):
    ni, nj = grid.ni, grid.nj
    X = grid.x.reshape(nj + 1, ni + 1)
    Y = grid.y.reshape(nj + 1, ni + 1)
    # Cell centroids of the quadrilaterals (area-weighted triangles).
    x0, x1, x2, x3 = X[:-1, :-1], X[:-1, 1:], X[1:, 1:], X[1:, :-1]
    y0, y1, y2, y3 = Y[:-1, :-1], Y[:-1, 1:], Y[1:, 1:], Y[1:, :-1]
    a1 = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    a2 = 0.5 * ((x2 - x0) * (y3 - y0) - (x3 - x0) * (y2 - y0))
    cx = (a1 * (x0 + x1 + x2) + a2 * (x0 + x2 + x3)) / (3 * (a1 + a2))
    cy = (a1 * (y0 + y1 + y2) + a2 * (y0 + y2 + y3)) / (3 * (a1 + a2))
    return cx, cy


def shock_heights(lines: LineSet, grid: StructuredGrid, primitives: np.ndarray, smoothing: int = 2) -> np.ndarray:
    """Distance along each line to the captured shock: in the column of
    cells beside the line, from the outer boundary inwards, the first
    crossing of half the column's pressure rise above the free stream's
    (p = 1), interpolated between cell centres; then `smoothing` passes of
    1/4, 1/2, 1/4 (halves at the ends), as fitted-flow.ts smooths the
    boundary it builds."""
    ni, nj = grid.ni, grid.nj
    prim = np.asarray(primitives, dtype=np.float64).reshape(nj, ni, 4)
    cx, cy = _centres(grid)
    h = np.zeros(ni + 1)
    for i in range(ni + 1):
        col = min(i, ni - 1)
        d = (cx[:, col] - lines.x[i]) * lines.dx[i] + (cy[:, col] - lines.y[i]) * lines.dy[i]
        p = prim[:, col, 3]
        mid = 1 + 0.5 * (p.max() - 1)
        j = nj - 1
        while j > 0 and p[j] < mid:
            j -= 1
        h[i] = d[-1] if j == nj - 1 else d[j] + (p[j] - mid) / (p[j] - p[j + 1]) * (d[j + 1] - d[j])
    for _ in range(smoothing):
        c = h.copy()
        h[0] = 0.5 * (c[0] + c[1])
        h[1:-1] = 0.25 * c[:-2] + 0.5 * c[1:-1] + 0.25 * c[2:]
        h[-1] = 0.5 * (c[-1] + c[-2])
    return h


def fitted_start(lines: LineSet, grid: StructuredGrid, primitives: np.ndarray, base_radius: float,
                 smoothing: int = 2) -> FittedStart:
    """The fitted grid on the captured grid's lines and its starting flow."""
    ni, nj = grid.ni, grid.nj
    heights = shock_heights(lines, grid, primitives, smoothing)
    fixed = with_fixed_fractions(lines.with_heights(heights), nj, base_radius)
    fg = grid_on_lines(fixed, heights, nj, base_radius, gated=False)
    prim = np.asarray(primitives, dtype=np.float64).reshape(nj, ni, 4)
    cx, cy = _centres(grid)
    fcx, fcy = _centres(fg)
    start = np.zeros((nj, ni, 4))
    for i in range(ni):
        d_cap = (cx[:, i] - lines.x[i]) * lines.dx[i] + (cy[:, i] - lines.y[i]) * lines.dy[i]
        d_fit = (fcx[:, i] - lines.x[i]) * lines.dx[i] + (fcy[:, i] - lines.y[i]) * lines.dy[i]
        for k in range(4):
            start[:, i, k] = np.interp(d_fit, d_cap, prim[:, i, k])
    return FittedStart(fixed, heights, fg, start.reshape(-1, 4))
