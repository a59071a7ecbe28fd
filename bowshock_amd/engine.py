"""GpuEuler2D over the HIP engine: the reference's GPU engine interface
(src/engine-gpu/euler2d-gpu.ts, class GpuEuler2D) for Python callers.

Method names, argument meaning and errors follow the reference: a config
object with the fields of Euler2DConfig (euler2d.ts:58-81), boundaries per
side as ``{"kind": ..., "state": (rho, u, v, p)}``, primitive states as
(rho, u, v, p), conserved state read back as FP32 arrays indexed
``i + j * ni`` (GpuState2D, euler2d-gpu.ts:116-125). Calls are synchronous;
a call made while another is running on the same solver raises BusyError,
as the reference's ExclusiveRunner does (gpu-common.ts:122-140).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, Mapping, Optional

import numpy as np

from . import _native as N

SIDES = ("iMin", "iMax", "jMin", "jMax")
BC_KINDS = {"inflow": 0, "outflow": 1, "wall": 2, "periodic": 3, "faceStates": 4, "farField": 5}
LIMITERS = {"firstOrder": 0, "minmod": 1, "vanLeer": 2, "vanAlbada": 3, "monotonizedCentral": 4,
            "unlimitedCentral": 5}
RECONSTRUCTIONS = {"perVariable": 0, "shared": 1, "sharedScaled": 2}
FLUXES = {"hllc": 0, "rusanov": 1}
INTEGRATORS = {"forwardEuler": 0, "ssprk2": 1, "ssprk3": 2}
KERNELS = {"stage": 0, "implicitPrepare": 1, "implicitRelax": 2, "implicitUpdate": 3}
RELAXATION_FORMS = {"fluxDifference": N.BS_RELAX_FLUX_DIFFERENCE, "linearized": N.BS_RELAX_LINEARIZED}


class RangeError(ValueError):
    """Invalid configuration or argument (the reference's RangeError)."""


class NonPhysicalStateError(RuntimeError):
    """src/solver/errors.ts:2-7."""


class GpuMemoryLimitError(MemoryError):
    """src/engine-gpu/gpu-common.ts:7."""


class GpuExecutionError(RuntimeError):
    """src/engine-gpu/gpu-common.ts:10."""


class GpuPrecisionError(RuntimeError):
    """src/engine-gpu/gpu-common.ts:17."""


class BusyError(RuntimeError):
    """A call overlapped another on the same solver."""


class UnsupportedError(NotImplementedError):
    pass


_ERRORS = {N.BS_ERR_RANGE: RangeError, N.BS_ERR_NONPHYSICAL: NonPhysicalStateError,
           N.BS_ERR_MEMORY: GpuMemoryLimitError, N.BS_ERR_EXECUTION: GpuExecutionError,
           N.BS_ERR_PRECISION: GpuPrecisionError, N.BS_ERR_BUSY: BusyError,
           N.BS_ERR_UNSUPPORTED: UnsupportedError}


def check(rc: int):
    if rc != N.BS_OK:
        raise _ERRORS.get(rc, RuntimeError)(N.last_error())


@dataclass
class GpuState2D:
    """euler2d-gpu.ts:116-125."""

    time: float
    steps: int
    rho: np.ndarray
    mu: np.ndarray
    mv: np.ndarray
    energy: np.ndarray


@dataclass
class Euler2DConfig:
    """euler2d.ts:58-81 / GpuEuler2DConfig euler2d-gpu.ts:82-92."""

    grid: object  # .ni, .nj, .x, .y
    geometry: str = "axisymmetric"
    gamma: float = 1.4
    flux: str = "hllc"
    limiter: str = "minmod"
    reconstruction: str = "perVariable"
    timeIntegrator: str = "ssprk3"
    cfl: float = 0.5
    timeStepping: str = "local"
    boundaries: Mapping[str, Mapping] = field(default_factory=dict)


class Context:
    """A HIP device (the reference's GPUDevice)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(N.lib().bs_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if self._h:
            N.lib().bs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _c_config(cfg: Euler2DConfig, keep: list) -> N.Config:
    c = N.Config()
    g = cfg.grid
    x = np.ascontiguousarray(g.x, dtype=np.float64)
    y = np.ascontiguousarray(g.y, dtype=np.float64)
    keep += [x, y]
    c.ni, c.nj = g.ni, g.nj
    c.x, c.y = N.dptr(x), N.dptr(y)
    c.geometry = 1 if cfg.geometry == "axisymmetric" else 0
    c.gamma = cfg.gamma
    if cfg.flux not in FLUXES:
        raise RangeError(f"flux {cfg.flux} is not available on the GPU")
    c.flux = FLUXES[cfg.flux]
    c.limiter = LIMITERS[cfg.limiter]
    c.reconstruction = RECONSTRUCTIONS[cfg.reconstruction]
    c.integrator = INTEGRATORS[cfg.timeIntegrator]
    c.cfl = cfg.cfl
    c.time_stepping = 1 if cfg.timeStepping == "local" else 0
    for k, side in enumerate(SIDES):
        bc = cfg.boundaries[side]
        c.boundaries[k].kind = BC_KINDS[bc["kind"]]
        st = bc.get("state")
        if st is None and bc["kind"] == "faceStates" and bc.get("states") is not None:
            st = bc["states"][0]
        st = st if st is not None else (0.0, 0.0, 0.0, 0.0)
        c.boundaries[k].rho, c.boundaries[k].u, c.boundaries[k].v, c.boundaries[k].p = map(float, st)
    return c


def compute_metrics(grid, geometry="axisymmetric"):
    """GridMetrics (structured.ts:28-50) in FP64, from the library."""
    ni, nj = grid.ni, grid.nj
    fi, fj, n = (ni + 1) * nj, ni * (nj + 1), ni * nj
    out = {k: np.zeros(m) for k, m in (("iNx", fi), ("iNy", fi), ("iArea", fi), ("jNx", fj), ("jNy", fj),
                                        ("jArea", fj), ("volume", n), ("planarArea", n), ("centroidX", n),
                                        ("centroidY", n))}
    x = np.ascontiguousarray(grid.x, dtype=np.float64)
    y = np.ascontiguousarray(grid.y, dtype=np.float64)
    check(N.lib().bs_compute_metrics(ni, nj, N.dptr(x), N.dptr(y), 1 if geometry == "axisymmetric" else 0,
                                     *[N.dptr(v) for v in out.values()]))
    return out


class GpuEuler2D:
    """The HIP twin of the reference's GpuEuler2D. precision="fp64" gives
    the precision extension: state, metrics and stencil arithmetic in FP64
    (in place of the reference's two-float engine, euler2d-two-float.ts);
    its implicit steps relax in FP32 with the linearized neighbour term."""

    def __init__(self, ctx: Context, config: Euler2DConfig, precision: str = "fp32"):
        if precision not in ("fp32", "fp64"):
            raise RangeError(f"unknown precision {precision!r}")
        self.config = config
        self.ctx = ctx
        self.precision = precision
        keep: list = []
        c = _c_config(config, keep)
        h = C.c_void_p()
        create = N.lib().bs_solver_create_fp64 if precision == "fp64" else N.lib().bs_solver_create
        check(create(ctx._h, C.byref(c), C.byref(h)))
        self._h = h
        self.ni, self.nj = config.grid.ni, config.grid.nj
        self.metrics = compute_metrics(config.grid, config.geometry)
        self.last_report = N.Report()
        for side in SIDES:
            bc = config.boundaries[side]
            if bc["kind"] == "faceStates" and bc.get("states") is not None:
                self.setFaceStates(side, bc["states"])

    @staticmethod
    def create(ctx: Context, config: Euler2DConfig, precision: str = "fp32") -> "GpuEuler2D":
        """euler2d-gpu.ts:217-291 (FP32); euler2d-two-float.ts:151 (FP64)."""
        return GpuEuler2D(ctx, config, precision)

    @property
    def n(self):
        return self.ni * self.nj

    def destroy(self):
        """euler2d-gpu.ts:659."""
        if getattr(self, "_h", None):
            N.lib().bs_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def setCellPrimitives(self, cell):
        """euler2d-gpu.ts:439-459: cell(i, j) -> (rho, u, v, p), or an
        (n, 4) array of primitives in cell order."""
        if callable(cell):
            prim = np.array([cell(c % self.ni, c // self.ni) for c in range(self.n)], dtype=np.float64)
        else:
            prim = np.ascontiguousarray(np.broadcast_to(np.asarray(cell, dtype=np.float64), (self.n, 4)))
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        check(N.lib().bs_set_primitives(self._h, N.dptr(prim)))

    def setInitialCondition(self, field: Callable):
        """euler2d-gpu.ts:412-432: field(x, y) at the cell centroids."""
        cx, cy = self.metrics["centroidX"], self.metrics["centroidY"]
        self.setCellPrimitives(np.array([field(a, b) for a, b in zip(cx, cy)], dtype=np.float64))

    def setState(self, rho, mu, mv, energy):
        """Conserved arrays; float64 arrays go through the FP64 entry
        (exact on an FP64 solver, rounded once on an FP32 one)."""
        sizes = {np.asarray(a).size for a in (rho, mu, mv, energy)}
        if sizes != {self.n}:
            raise RangeError(f"the state needs {self.n} values per array, got {sorted(sizes)}")
        if any(np.asarray(a).dtype == np.float64 for a in (rho, mu, mv, energy)):
            arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (rho, mu, mv, energy)]
            check(N.lib().bs_set_state_f64(self._h, *[N.dptr(a) for a in arrs], self.n))
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in (rho, mu, mv, energy)]
        check(N.lib().bs_set_state(self._h, *[N.fptr(a) for a in arrs], self.n))

    def iterate(self, count: int):
        """euler2d-gpu.ts:462-473."""
        check(N.lib().bs_iterate(self._h, int(count), C.byref(self.last_report)))

    def advanceTo(self, tEnd: float, maxSteps: int = 1_000_000):
        """euler2d-gpu.ts:511-527."""
        check(N.lib().bs_advance_to(self._h, float(tEnd), int(maxSteps), C.byref(self.last_report)))

    def densityResidualNorm(self) -> float:
        """euler2d-gpu.ts:530-546 (FP64 reduction on the GPU)."""
        out = C.c_double()
        check(N.lib().bs_residual_norm(self._h, C.byref(out)))
        return out.value

    def checkPrecision(self) -> float:
        """TransonicGridSolver.checkPrecision (transonic-drag.ts:56-66,
        gpu-transonic.ts:108-131): raises GpuPrecisionError when the
        residual differs from a double-precision one of the same state by
        more than 1e-8 of the largest; returns the relative difference
        (0 for an FP64 solver, whose residual is the double one)."""
        out = C.c_double()
        check(N.lib().bs_check_precision(self._h, C.byref(out)))
        return out.value

    def residualVector(self) -> np.ndarray:
        """euler2d.ts:462-477: (n, 4)."""
        r = np.zeros(4 * self.n)
        check(N.lib().bs_residual_vector(self._h, N.dptr(r)))
        return r.reshape(self.n, 4)

    def readState(self) -> GpuState2D:
        """euler2d-gpu.ts:549-572: float32 arrays, or float64 ones from an
        FP64 solver (readConserved, euler2d-two-float.ts:486-501)."""
        rep = N.Report()
        if self.precision == "fp64":
            a = [np.zeros(self.n, dtype=np.float64) for _ in range(4)]
            check(N.lib().bs_read_state_f64(self._h, *[N.dptr(v) for v in a], C.byref(rep)))
        else:
            a = [np.zeros(self.n, dtype=np.float32) for _ in range(4)]
            check(N.lib().bs_read_state(self._h, *[N.fptr(v) for v in a], C.byref(rep)))
        return GpuState2D(rep.time, int(rep.steps), *a)

    def setFaceStates(self, side: str, states):
        """euler2d-gpu.ts:575-600."""
        a = np.ascontiguousarray(states, dtype=np.float64).reshape(-1)
        check(N.lib().bs_set_face_states(self._h, SIDES.index(side), N.dptr(a), len(a) // 4))

    def setFaceFluxes(self, side: str, fluxes):
        """euler2d-gpu.ts:606-630."""
        k = SIDES.index(side)
        if fluxes is None:
            check(N.lib().bs_set_face_fluxes(self._h, k, None, 0))
            return
        a = np.ascontiguousarray(fluxes, dtype=np.float64).reshape(-1)
        check(N.lib().bs_set_face_fluxes(self._h, k, N.dptr(a), len(a)))

    def wallFluxes(self, side: str = "jMin") -> np.ndarray:
        """Euler2D.boundaryFluxes (euler2d.ts:499-525) evaluated on the GPU
        in FP64: (faces, 9) rows of x, y, nx, ny, area, flux[4]."""
        k = SIDES.index(side)
        m = self.nj if k < 2 else self.ni
        out = np.zeros(9 * m)
        check(N.lib().bs_wall_fluxes(self._h, k, N.dptr(out), m))
        return out.reshape(m, 9)

    def forebodyDrag(self, pressure: float, dynamicPressure: float, baseRadius: float, side: str = "jMin") -> float:
        """forces.ts:19-29 forebodyPressureDrag over the side's faces, FP64 on
        the GPU."""
        out = C.c_double()
        check(N.lib().bs_forebody_drag(self._h, SIDES.index(side), pressure, dynamicPressure, baseRadius,
                                       C.byref(out)))
        return out.value

    def setFittedShock(self, freeStream):
        """GpuEuler2DConfig.fittedShock (euler2d-gpu.ts:85-91): jMax, a
        faceStates side, becomes a fitted shock with this free stream ahead
        of it (None turns it off)."""
        if freeStream is None:
            check(N.lib().bs_set_fitted_shock(self._h, None))
            return
        w = np.ascontiguousarray(freeStream, dtype=np.float64)
        check(N.lib().bs_set_fitted_shock(self._h, N.dptr(w)))

    def readShock(self):
        """euler2d-gpu.ts:630-656: speed, Mach number and signal per jMax face."""
        out = [np.zeros(self.ni) for _ in range(3)]
        check(N.lib().bs_read_shock(self._h, *[N.dptr(a) for a in out], self.ni))
        return {"speed": out[0], "mach": out[1], "signal": out[2]}

    def setLineGrid(self, lines, fractions, firstNode: str = "fixed"):
        """GpuLineGrid.create (line-grid.ts:63-110) on this solver: lines
        as (ni + 1, 4) rows x, y, dx, dy (unit directions), fractions as
        (nj + 1, ni + 1) (row j, line i); firstNode "symmetry" or "fixed"."""
        ln = np.ascontiguousarray(lines, dtype=np.float64).reshape(self.ni + 1, 4)
        fr = np.ascontiguousarray(fractions, dtype=np.float64).reshape(self.nj + 1, self.ni + 1)
        check(N.lib().bs_set_line_grid(self._h, N.dptr(ln), N.dptr(fr), 1 if firstNode == "symmetry" else 0))

    def setHeights(self, heights):
        """GpuLineGrid.setHeights: the nodes at these heights, metrics on the GPU."""
        h = self._heights(heights)
        check(N.lib().bs_set_heights(self._h, N.dptr(h), h.size))

    def moveGrid(self, heights):
        """euler2d-gpu.ts:400-405: the lines to these heights, the flow remapped."""
        h = self._heights(heights)
        check(N.lib().bs_move_grid(self._h, N.dptr(h), h.size))

    def iterateFitted(self, moves: int, stepsPerMove: int = 1):
        """euler2d-gpu.ts:481-505."""
        check(N.lib().bs_iterate_fitted(self._h, int(moves), int(stepsPerMove), C.byref(self.last_report)))

    def readHeights(self) -> np.ndarray:
        """GpuLineGrid.readHeights (line-grid.ts:257-270)."""
        h = np.zeros(self.ni + 1)
        check(N.lib().bs_read_heights(self._h, N.dptr(h), h.size))
        return h

    def _heights(self, heights) -> np.ndarray:
        """One height per line (line-grid.ts:301: RangeError otherwise)."""
        h = np.ascontiguousarray(heights, dtype=np.float64).reshape(-1)
        if h.size != self.ni + 1:
            raise RangeError(f"expected {self.ni + 1} heights, got {h.size}")
        return h

    def setRelaxation(self, form: str):
        """The relaxation's neighbour term: "fluxDifference" (the
        reference's, euler2d.ts:981-1005; default of FP32 solvers) or
        "linearized" (the two-float engine's, euler2d-two-float.wgsl:
        574-598; the only form of FP64 solvers)."""
        check(N.lib().bs_set_relaxation(self._h, RELAXATION_FORMS[form]))

    def setImplicitCfl(self, cfl: float):
        """euler2d-gpu.ts:668-681."""
        check(N.lib().bs_set_implicit_cfl(self._h, float(cfl)))

    def iterateImplicit(self, count: int, cfl: float, sweeps: int, omega: float = 1.5, growth: float = 1.2):
        """euler2d-gpu.ts:687-723."""
        check(N.lib().bs_iterate_implicit(self._h, int(count), float(cfl), int(sweeps), float(omega), float(growth),
                                          C.byref(self.last_report)))

    def iterateAnderson(self, anderson: "AndersonMixing", count: int, cfl: float, sweeps: int, omega: float = 1.5,
                        growth: float = 1.2):
        """Implicit steps accelerated by `anderson` (size 4 ni nj), the inner
        loop of solveTransonic (transonic-flow.ts:184-208) on the device."""
        check(N.lib().bs_iterate_anderson(self._h, anderson._h, int(count), float(cfl), int(sweeps), float(omega),
                                          float(growth), C.byref(self.last_report)))

    def enableGraphs(self, on: bool = True):
        """Run steps as captured HIP graphs (default on)."""
        check(N.lib().bs_enable_graphs(self._h, 1 if on else 0))

    def enableTiming(self, on=True):
        """HIP-event timing of the solver's kernels: True (or "launch") times
        every launch, without graphs; "region" times whole batches of
        explicit local steps, which launch only the stage kernel, with graphs
        on; False stops."""
        mode = 2 if on == "region" else (1 if on else 0)
        check(N.lib().bs_enable_timing(self._h, mode))

    def kernelTiming(self, kernel: str = "stage"):
        """(ms per launch, launches) since the last call."""
        ms = C.c_double()
        L = C.c_int64()
        check(N.lib().bs_kernel_timing(self._h, KERNELS[kernel], C.byref(ms), C.byref(L)))
        return ms.value, L.value


class AndersonMixing:
    """anderson.ts:20-123 AndersonMixing with its history on the GPU."""

    def __init__(self, ctx: Context, size: int, depth: int):
        h = C.c_void_p()
        check(N.lib().bs_anderson_create(ctx._h, int(size), int(depth), C.byref(h)))
        self._h = h
        self.size, self.depth = int(size), int(depth)

    def reset(self):
        check(N.lib().bs_anderson_reset(self._h))

    def next(self, x, f) -> np.ndarray:
        """The next iterate from x and its correction f (both unchanged)."""
        xa = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        fa = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        if xa.size != self.size or fa.size != self.size:
            raise RangeError("wrong vector size")
        out = np.empty(self.size)
        check(N.lib().bs_anderson_next(self._h, N.dptr(xa), N.dptr(fa), N.dptr(out)))
        return out

    def state(self):
        """(columns held, columns used, restarts, gamma oldest first)."""
        cols, used, restarts = C.c_int32(), C.c_int32(), C.c_int32()
        gamma = np.zeros(self.depth)
        check(N.lib().bs_anderson_state(self._h, C.byref(cols), C.byref(used), C.byref(restarts), N.dptr(gamma)))
        return cols.value, used.value, restarts.value, gamma

    def destroy(self):
        if getattr(self, "_h", None):
            N.lib().bs_anderson_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def gpu_primitives(state: GpuState2D, gamma: float) -> np.ndarray:
    """gpuPrimitive2D (euler2d-gpu.ts:136-146): FP64 primitives of a state
    read back, pressure with the FP32 gamma for an FP32 state (the gamma
    itself for an FP64 state); (n, 4)."""
    g32 = float(gamma) if state.rho.dtype == np.float64 else float(np.float32(gamma))
    rho = state.rho.astype(np.float64)
    mu = state.mu.astype(np.float64)
    mv = state.mv.astype(np.float64)
    e = state.energy.astype(np.float64)
    p = (g32 - 1) * (e - 0.5 * (mu * mu + mv * mv) / rho)
    return np.stack([rho, mu / rho, mv / rho, p], axis=1)
