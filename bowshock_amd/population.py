"""Seeded case populations for batched solves: a Mach sweep of one nose
and a population of nose-plus-cylinder bodies at several Mach numbers.

The draws use mulberry32 (docs/research/refusals/population.ts:64-72),
seed 20261020 as the reference's population (population.ts:158-159). The
five nose kinds map to profiles as the reference does (noseShape,
docs/research/refusals/evaluate.ts:139-151): conical, tangent ogive
(ogive, parameter 1), parabolic (K = 1), ellipsoid and Von Karman
(Haack, parameter 0). Per rocket, in this order: kind = floor(5 r),
fineness U[1, 5] calibers, diameter U[24, 100] mm, cylinder U[6, 20]
calibers. Lengths are in diameters (base radius 0.5); the diameter only
rescales an inviscid solution and is carried as an input.

Grid sizes (the reference sizes its grids from the geometry instead,
body-grid.ts:76-181): nj = ni / 2, and
- Mach sweep: the nose on a 4-caliber cylinder,
  ni = 40 + 120 (f - 1) / 4 rounded to even (40 ... 160);
- population: L = f + cylinder, ni = 160 + 480 (L - 7) / 18 rounded to
  even (160 ... 640).

A case is refused, with its reason, where the H grid cannot carry it:
blunt tips (ellipsoid, Von Karman; the reference uses a C grid,
src/grid/blunt-grid.ts), tip half-angles above 30 degrees
(body-grid.ts:74), and Mach numbers at which the tip shock detaches.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import grids

SEED = 20261020
GAMMA = 1.4
NOSE_KINDS = ("conical", "tangent-ogive", "parabolic", "ellipsoid", "von-karman")
SWEEP_MACHS = tuple(round(1.2 + 0.1 * k, 10) for k in range(39))   # 1.2 : 0.1 : 5.0
POPULATION_MACHS = (1.5, 2.0, 3.0, 4.0)


def mulberry32(seed: int):
    """population.ts:64-72: 32-bit state, uniform doubles in [0, 1)."""
    a = seed & 0xFFFFFFFF

    def imul(x, y):
        return (x * y) & 0xFFFFFFFF

    def rand():
        nonlocal a
        a = (a + 0x6D2B79F5) & 0xFFFFFFFF
        t = a
        t = imul(t ^ (t >> 15), t | 1)
        t ^= (t + imul(t ^ (t >> 7), t | 61)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296

    return rand


def nose_of(kind: str, fineness: float) -> grids.Nose:
    """noseShape (evaluate.ts:139-151) at base radius 0.5."""
    if kind == "conical":
        return grids.Nose("conical", fineness, 0.5)
    if kind == "tangent-ogive":
        return grids.Nose("ogive", fineness, 0.5, 1.0)
    if kind == "parabolic":
        return grids.Nose("parabolic", fineness, 0.5, 1.0)
    if kind == "ellipsoid":
        return grids.Nose("ellipsoid", fineness, 0.5)
    if kind == "von-karman":
        return grids.Nose("haack", fineness, 0.5, 0.0)
    raise ValueError(f"unknown nose kind {kind}")


@dataclass(frozen=True)
class Rocket:
    index: int
    kind: str
    fineness: float
    diameter_mm: float
    cylinder: float


def rockets(n: int, seed: int = SEED) -> List[Rocket]:
    r = mulberry32(seed)
    out = []
    for k in range(n):
        kind = NOSE_KINDS[min(4, int(5 * r()))]
        fineness = 1 + 4 * r()
        diameter = 24 + 76 * r()
        cylinder = 6 + 14 * r()
        out.append(Rocket(k, kind, fineness, diameter, cylinder))
    return out


def _even(v: float) -> int:
    return max(2, 2 * int(round(v / 2)))


@dataclass
class Case:
    rocket: Rocket
    mach: float
    ni: int
    nj: int
    cylinder: float
    refused: Optional[str] = None
    grid: Optional[grids.StructuredGrid] = None
    i_tip: int = 0

    @property
    def body(self) -> grids.Body:
        return grids.Body(nose_of(self.rocket.kind, self.rocket.fineness), self.cylinder)

    @property
    def cells(self) -> int:
        return self.ni * self.nj

    def features(self) -> np.ndarray:
        """[one-hot(5), fineness, diameter (m), Mach]."""
        v = np.zeros(8)
        v[NOSE_KINDS.index(self.rocket.kind)] = 1.0
        v[5] = self.rocket.fineness
        v[6] = self.rocket.diameter_mm / 1000.0
        v[7] = self.mach
        return v


def sweep_cases(seed: int = SEED) -> List[Case]:
    """One nose (the first draw whose tip an H grid carries) on a
    4-caliber cylinder at Mach 1.2, 1.3, ..., 5.0."""
    r = rockets(64, seed)
    pick = next(x for x in r if x.kind in ("conical", "tangent-ogive", "parabolic")
                and math.atan(grids.nose_tip_slope(nose_of(x.kind, x.fineness))) <= grids.MAX_TIP_HALF_ANGLE)
    ni = _even(40 + 120 * (pick.fineness - 1) / 4)
    return [Case(pick, m, ni, ni // 2, 4.0) for m in SWEEP_MACHS]


def population_cases(n_rockets: int = 2000, machs: Sequence[float] = POPULATION_MACHS, seed: int = SEED) -> List[Case]:
    out = []
    for rk in rockets(n_rockets, seed):
        L = rk.fineness + rk.cylinder
        ni = _even(160 + 480 * (L - 7) / 18)
        for m in machs:
            out.append(Case(rk, m, ni, ni // 2, rk.cylinder))
    return out


def _grid_job(args) -> Tuple[Optional[str], Optional[Tuple[np.ndarray, np.ndarray]], int]:
    body, ni, nj, mach, tol = args
    try:
        shock = grids.tip_shock_angle(body, mach, GAMMA, tol)
    except ValueError as e:
        return str(e), None, 0
    bg = grids.body_grid_array(body, ni, nj, shock)
    return None, (bg.grid.x, bg.grid.y), bg.i_tip


def build_grids(cases: List[Case], workers: int = 0, tolerance: float = 1e-10) -> List[Case]:
    """Generate every case's grid (or its refusal) in a process pool; the
    Taylor-Maccoll tolerance only moves the outer boundary's angle."""
    jobs = [(c.body, c.ni, c.nj, c.mach, tolerance) for c in cases]
    workers = workers or min(16, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 8:
        with ProcessPoolExecutor(workers) as ex:
            results = list(ex.map(_grid_job, jobs, chunksize=8))
    else:
        results = [_grid_job(j) for j in jobs]
    for c, (refusal, xy, i_tip) in zip(cases, results):
        c.refused = refusal
        if xy is not None:
            c.grid = grids.StructuredGrid(c.ni, c.nj, xy[0], xy[1])
            c.i_tip = i_tip
    return cases


def refusal_counts(cases: Sequence[Case]) -> dict:
    counts: dict = {}
    for c in cases:
        key = c.refused or "accepted"
        if "blunt" in key:
            key = "blunt tip (needs a C grid)"
        elif "tip half-angle" in key:
            key = "tip half-angle above 30 deg"
        elif "detach" in key:
            key = "detached tip shock"
        counts[key] = counts.get(key, 0) + 1
    return counts
