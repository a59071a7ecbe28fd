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

Blunt tips (ellipsoid, Von Karman) take the C grid of
src/grid/blunt-grid.ts (bowshock_amd/blunt_grid.py), the axis ahead of
the tip a wall: 20 cells across refined by the power of two nearest
ni / 80 (2, 4 or 8, so nj = 40 ... 160; ni follows from the generator's
density of lines). A case is refused, with its reason, where its grid
cannot carry it: a C grid that fails the generator's quality gates, tip
half-angles above 30 degrees (body-grid.ts:74), and Mach numbers at which
an H grid's tip shock detaches.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import grids

SEED = 20261020
GAMMA = 1.4
NOSE_KINDS = ("conical", "tangent-ogive", "parabolic", "ellipsoid", "von-karman")
BLUNT_KINDS = ("ellipsoid", "von-karman")
C_GRID_BASE_NJ = 20
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
    topology: str = "H"   # "C": the blunt-body grid, iMin the axis ahead of the tip
    refinement: int = 1   # a C grid's refinement of its family's base grid

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


def c_grid_refinement(ni: int) -> int:
    """The power of two nearest ni / 80, within 1 ... 8."""
    return int(2 ** min(3, max(0, round(math.log2(ni / 80)))))


def boundaries(case: Case, free) -> dict:
    """The case's boundary conditions: an H grid's iMin is the inflow
    ahead of the tip, a C grid's the axis (a wall; blunt-body.test.ts:35)."""
    inflow = {"kind": "inflow", "state": free}
    return {"iMin": inflow if case.topology == "H" else {"kind": "wall"}, "iMax": {"kind": "outflow"},
            "jMin": {"kind": "wall"}, "jMax": inflow}


def _grid_job(args):
    body, ni, nj, mach, tol, blunt = args[:6]
    refinement = args[6] if len(args) > 6 else 0
    if blunt:
        from .blunt_grid import blunt_body_grid

        r = refinement or c_grid_refinement(ni)
        try:
            bg = blunt_body_grid(body, C_GRID_BASE_NJ, mach, refinement=r)
        except ValueError as e:   # GridQualityError included
            return "C grid: " + str(e), None, 0, "C", r
        return None, (bg.grid.x, bg.grid.y, bg.grid.ni, bg.grid.nj), 0, "C", r
    try:
        shock = grids.tip_shock_angle(body, mach, GAMMA, tol)
    except ValueError as e:
        return str(e), None, 0, "H", 1
    bg = grids.body_grid_array(body, ni, nj, shock)
    return None, (bg.grid.x, bg.grid.y, ni, nj), bg.i_tip, "H", 1


def build_grids(cases: List[Case], workers: int = 0, tolerance: float = 1e-10) -> List[Case]:
    """Generate every case's grid (or its refusal) in a process pool; the
    Taylor-Maccoll tolerance only moves the outer boundary's angle. Blunt
    tips get C grids, whose size the generator sets: ni and nj change."""
    jobs = [(c.body, c.ni, c.nj, c.mach, tolerance, c.rocket.kind in BLUNT_KINDS) for c in cases]
    workers = workers or min(16, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 8:
        # The C grids take about a second each, the H grids milliseconds:
        # the C jobs first, in small chunks, so the pool stays balanced.
        order = sorted(range(len(jobs)), key=lambda k: not jobs[k][5])
        with ProcessPoolExecutor(workers) as ex:
            done = list(ex.map(_grid_job, [jobs[k] for k in order], chunksize=2))
        results = [None] * len(jobs)
        for k, r in zip(order, done):
            results[k] = r
    else:
        results = [_grid_job(j) for j in jobs]
    for c, (refusal, xy, i_tip, topology, refinement) in zip(cases, results):
        c.refused = refusal
        c.topology = topology
        c.refinement = refinement
        if xy is not None:
            c.ni, c.nj = int(xy[2]), int(xy[3])
            c.grid = grids.StructuredGrid(c.ni, c.nj, xy[0], xy[1])
            c.i_tip = i_tip
    return cases


def coarser_cases(cases: Sequence[Case], workers: int = 0, tolerance: float = 1e-10) -> List[Optional[Case]]:
    """Each accepted case's grid half as fine, for the start of a solve
    from a coarser grid (solveTransonic, transonic-flow.ts:157-168): a C
    grid's family member of half the refinement (its grids nest), an H
    grid with half the columns and rows over the same body and outer
    boundary. None where the case has no grid, its C grid is its family's
    base grid, or the coarser grid is refused."""
    out: List[Optional[Case]] = []
    jobs, where = [], []
    for k, c in enumerate(cases):
        if c.grid is None or (c.topology == "C" and c.refinement < 2):
            out.append(None)
            continue
        blunt = c.topology == "C"
        ni, nj = (c.ni, c.nj) if blunt else (max(2, c.ni // 2), max(2, (c.nj + 1) // 2))
        out.append(Case(c.rocket, c.mach, ni, nj, c.cylinder))
        jobs.append((c.body, ni, nj, c.mach, tolerance, blunt, c.refinement // 2 if blunt else 0))
        where.append(k)
    workers = workers or min(16, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 8:
        with ProcessPoolExecutor(workers) as ex:
            results = list(ex.map(_grid_job, jobs, chunksize=2))
    else:
        results = [_grid_job(j) for j in jobs]
    for k, (refusal, xy, i_tip, topology, refinement) in zip(where, results):
        c = out[k]
        if xy is None:
            out[k] = None
            continue
        c.topology, c.refinement, c.i_tip = topology, refinement, i_tip
        c.ni, c.nj = int(xy[2]), int(xy[3])
        c.grid = grids.StructuredGrid(c.ni, c.nj, xy[0], xy[1])
    return out


def refusal_counts(cases: Sequence[Case]) -> dict:
    counts: dict = {}
    for c in cases:
        key = c.refused or ("accepted" if c.topology == "H" else "accepted (C grid)")
        if key.startswith("C grid"):
            key = "C grid fails its quality gates" if "quality" in key else "C grid refused"
        elif "blunt" in key:
            key = "blunt tip (needs a C grid)"
        elif "tip half-angle" in key:
            key = "tip half-angle above 30 deg"
        elif "detach" in key:
            key = "detached tip shock"
        counts[key] = counts.get(key, 0) + 1
    return counts
