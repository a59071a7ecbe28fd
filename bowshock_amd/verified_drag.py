"""The forebody pressure drag of a body with its measured discretisation
error, solved on the HIP engine: analysis/verified-drag.ts
verifiedForebodyDrag (lines 230-319) over the HIP GpuEuler2D.

Each grid of the nested bodyGrid family (20 cells across at refinement 1,
doubled per level) is solved from the free stream with the reference's
scheme (HLLC, van Albada shared-scaled reconstruction, SSP-RK2 at CFL 0.8,
local steps) in batches of 500 r steps until the drag itself settles
(SettlingValue), then checked: the flow must leave the grid supersonic
and the outer boundary must still see the free stream. The levels feed
refine_until_verified. The drag is integrated in FP64 on the GPU from the
FP32 state, or from the FP64 state of the FP64 solver.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import grids
from .engine import Context, Euler2DConfig, GpuEuler2D, gpu_primitives
from .verification import SettlingValue, VerifiedValue, refine_until_verified

# verified-drag.ts:102-200
GPU_NOISE = 2e-6
CPU_NOISE = 1e-12
DRAG_FLOOR = 1e-3
ORDER_TOLERANCE = 0.5
OUTER_TOLERANCE = 1e-5
BASE_NJ = 20
BATCH = 500
MAX_BATCHES = 60


class IterationStalled(RuntimeError):
    pass


class InvalidCase(RuntimeError):
    """The boundary conditions do not represent the case: the outflow is
    not supersonic, or the bow shock reaches the outer boundary."""


@dataclass
class LevelReport:
    refinement: int
    ni: int
    nj: int
    iterations: int
    drag: float


@dataclass
class VerifiedDrag:
    kind: str                      # 'solved', 'not-converged', 'invalid'
    result: Optional[VerifiedValue] = None
    reports: List[LevelReport] = field(default_factory=list)
    refinement: int = 0
    reason: str = ""


def _primitives(gpu: GpuEuler2D, gamma: float) -> np.ndarray:
    """(nj, ni, 4) primitives of the current state (gpuPrimitive2D,
    euler2d-gpu.ts:136-146)."""
    return gpu_primitives(gpu.readState(), gamma).reshape(gpu.nj, gpu.ni, 4)


def outflow_mach(prim: np.ndarray, gamma: float) -> float:
    """verified-drag.ts:335-343: smallest axial Mach number in the last
    column of cells."""
    w = prim[:, -1, :]
    return float(np.min(w[:, 1] / np.sqrt(gamma * w[:, 3] / w[:, 0])))


def outer_deviation(prim: np.ndarray, free, gamma: float) -> float:
    """verified-drag.ts:350-369: largest deviation from the free stream in
    the row along the outer boundary."""
    rho_f, u_f, v_f, p_f = free
    a = math.sqrt(gamma * p_f / rho_f)
    w = prim[-1, :, :]
    return float(max(np.max(np.abs(w[:, 0] / rho_f - 1)), np.max(np.abs(w[:, 3] / p_f - 1)),
                     np.max(np.abs(w[:, 1] - u_f)) / a, np.max(np.abs(w[:, 2] - v_f)) / a))


def verified_forebody_drag(ctx: Context, body: grids.Body, *, mach: float, gamma: float = 1.4,
                           tolerance: float = 1e-4, max_refinement: int = 4, noise: float = GPU_NOISE,
                           floor: float = DRAG_FLOOR, order_tolerance: float = ORDER_TOLERANCE,
                           on_level: Optional[Callable[[LevelReport], None]] = None,
                           precision: str = "fp32") -> VerifiedDrag:
    """verified-drag.ts:230-319 on the HIP engine. precision="fp64" solves
    on the FP64 solver, which settles at the CPU engine's noise (1e-12)."""
    free, p_inf, q_inf = grids.free_stream_at(mach, gamma)
    reports: List[LevelReport] = []
    last = [1]

    def solve_at(r: int) -> float:
        last[0] = r
        bg = grids.body_grid(body, BASE_NJ, mach, gamma, refinement=r)
        bcs = {"iMin": {"kind": "inflow", "state": free}, "iMax": {"kind": "outflow"},
               "jMin": {"kind": "wall"}, "jMax": {"kind": "inflow", "state": free}}
        cfg = Euler2DConfig(grid=bg.grid, boundaries=bcs, geometry="axisymmetric", gamma=gamma, flux="hllc",
                            limiter="vanAlbada", reconstruction="sharedScaled", timeIntegrator="ssprk2", cfl=0.8,
                            timeStepping="local")
        gpu = GpuEuler2D.create(ctx, cfg, precision)
        try:
            gpu.setCellPrimitives(np.array(free))
            settling = SettlingValue(noise)
            for batch in range(1, MAX_BATCHES + 1):
                gpu.iterate(BATCH * r)
                drag = gpu.forebodyDrag(p_inf, q_inf, body.nose.base_radius)
                if not math.isfinite(drag):
                    raise IterationStalled("the solution is not finite")
                if settling.add(drag):
                    prim = _primitives(gpu, gamma)
                    out = outflow_mach(prim, gamma)
                    if not out > 1:
                        raise InvalidCase(f"the flow leaves the grid at an axial Mach number of {out:.3f}, where "
                                          "the outflow condition does not hold")
                    dev = outer_deviation(prim, free, gamma)
                    if not dev <= OUTER_TOLERANCE:
                        raise InvalidCase(f"the flow along the outer boundary differs from the free stream by "
                                          f"{dev:.1e}: the bow shock reaches it")
                    rep = LevelReport(r, bg.grid.ni, bg.grid.nj, batch * BATCH * r, drag)
                    reports.append(rep)
                    if on_level:
                        on_level(rep)
                    return drag
            raise IterationStalled(f"the drag still changes by more than {noise} of its value after "
                                   f"{MAX_BATCHES * BATCH * r} iterations")
        finally:
            gpu.destroy()

    try:
        res = refine_until_verified(solve_at, tolerance=tolerance, floor=floor, max_refinement=max_refinement,
                                    order_tolerance=order_tolerance, noise=noise)
        return VerifiedDrag("solved", res, reports)
    except IterationStalled as e:
        return VerifiedDrag("not-converged", None, reports, last[0], str(e))
    except InvalidCase as e:
        return VerifiedDrag("invalid", None, reports, last[0], str(e))
