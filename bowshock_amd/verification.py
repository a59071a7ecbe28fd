"""Solution verification: when an iterated value has settled, and the
discretisation error of a value from a nested family of grids.

Restates src/verification/iterative-convergence.ts (SettlingValue) and
src/verification/grid-convergence.ts (gridConvergence, refineUntilVerified)
of the reference: the settling test on the value's own geometric rate, and
the grid convergence index of Celik et al. (2008) with a refinement ratio
of 2, the observed order capped at the formal order 2, agreement of the
two finest triples' orders required, an oscillation bounded by the range
of the three finest values, and a finest change within the noise bounded
by the order the coarser grids show.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

# iterative-convergence.ts:241-243
_MEASURABLE_CHANGE = 10
_RATE_POINTS = 8


class SettlingValue:
    """iterative-convergence.ts:54-119: whether a value iterated towards a
    limit has stopped changing by more than `noise` (relative) — the last
    change within the noise, and what geometric convergence at the
    measured rate still leaves within it, on two successive checks."""

    def __init__(self, noise: float, known_rate: float = math.nan):
        self.noise = noise
        self.known_rate = known_rate
        self.values: List[float] = []
        self.measurable: List[tuple] = []
        self.passes = 0

    @property
    def measured_rate(self) -> float:
        """exp of the slope of log(change) over the last 8 measurable
        changes against their check index (least squares); NaN if fewer
        than two."""
        pts = self.measurable[-_RATE_POINTS:]
        if len(pts) < 2:
            return math.nan
        mx = sum(x for x, _ in pts) / len(pts)
        my = sum(y for _, y in pts) / len(pts)
        sxy = sum((x - mx) * (y - my) for x, y in pts)
        sxx = sum((x - mx) ** 2 for x, _ in pts)
        return math.exp(sxy / sxx)

    @property
    def rate(self) -> float:
        usable = [r for r in (self.measured_rate, self.known_rate) if r < 1]
        return max(usable) if usable else math.nan

    @property
    def rate_to_pass_on(self) -> float:
        m = self.measured_rate
        return m if m < 1 else self.known_rate

    def add(self, value: float) -> bool:
        n = len(self.values)
        change = abs(value - self.values[-1]) if n else math.nan
        self.values.append(value)
        allowed = self.noise * abs(value)
        if change > _MEASURABLE_CHANGE * allowed:
            self.measurable.append((n, math.log(change)))
        rate = self.rate
        remaining = 0.0
        if rate < 1:
            k = max(1, math.ceil(math.log(0.5) / math.log(rate)))
            if n < k:
                remaining = math.inf
            else:
                f = rate ** k
                remaining = abs(value - self.values[n - k]) * f / (1 - f)
        passed = change <= allowed and remaining <= allowed
        self.passes = self.passes + 1 if passed else 0
        return self.passes >= 2


# grid-convergence.ts:52-60
FORMAL_ORDER = 2
SAFETY_FACTOR = 1.25
RATIO = 2
_MEASURABLE_GRID = 10

_OSCILLATES = "the values oscillate under refinement"
_TOO_SMALL = "the changes between grids are too small to measure"


@dataclass
class ConvergenceEstimate:
    """kind: 'monotone' (value, extrapolated, order, uncertainty),
    'oscillatory' (value, uncertainty) or 'unresolved' (value, reason)."""
    kind: str
    value: float
    extrapolated: float = math.nan
    order: float = math.nan
    uncertainty: float = math.nan
    reason: str = ""


def _unmeasurable(values: Sequence[float], noise: float) -> float:
    return _MEASURABLE_GRID * noise * max(abs(v) for v in values)


def _observed_order(coarse: float, medium: float, fine: float, noise: float):
    e21 = medium - fine
    e32 = coarse - medium
    lim = _unmeasurable([coarse, medium, fine], noise)
    if abs(e21) <= lim or abs(e32) <= lim:
        return _TOO_SMALL
    if e32 / e21 < 0:
        return _OSCILLATES
    if abs(e32 / e21) <= 1:
        return "the changes do not shrink under refinement"
    return math.log(abs(e32 / e21)) / math.log(RATIO)


def _within_noise(values: Sequence[float], coarse_order: float, noise: float) -> ConvergenceEstimate:
    value = values[3]
    e21 = values[2] - value
    predicted = (values[1] - values[2]) / RATIO ** coarse_order
    if not abs(e21 - predicted) <= _unmeasurable(values, noise):
        return ConvergenceEstimate("unresolved", value, reason=(
            f"the finest change ({e21:.2e}) is below the noise but not the one the coarser grids predict "
            f"({predicted:.2e})"))
    gain = RATIO ** min(coarse_order, FORMAL_ORDER) - 1
    n = noise * max(abs(v) for v in values)
    bound = max(abs(e21), abs(predicted)) + 2 * n
    return ConvergenceEstimate("monotone", value, extrapolated=value - predicted / gain, order=coarse_order,
                               uncertainty=SAFETY_FACTOR * bound / gain)


def grid_convergence(values: Sequence[float], order_tolerance: float, noise: float) -> ConvergenceEstimate:
    """grid-convergence.ts:109-155 gridConvergence: values on grids refined
    by 2, coarsest first."""
    if not all(math.isfinite(v) for v in values):
        raise ValueError("the values must be finite")
    n = len(values)
    value = values[-1]
    if n < 4:
        return ConvergenceEstimate("unresolved", value, reason="four grids are needed to confirm the order")
    finest = values[n - 3:]
    fine = _observed_order(finest[0], finest[1], value, noise)
    if fine == _OSCILLATES:
        return ConvergenceEstimate("oscillatory", value, uncertainty=max(finest) - min(finest))
    coarse = _observed_order(values[n - 4], values[n - 3], values[n - 2], noise)
    if fine == _TOO_SMALL and not isinstance(coarse, str):
        return _within_noise(values[n - 4:], coarse, noise)
    for triple in (fine, coarse):
        if isinstance(triple, str):
            return ConvergenceEstimate("unresolved", value, reason=triple)
    order = fine
    if not abs(order - coarse) <= order_tolerance:
        return ConvergenceEstimate("unresolved", value, reason=(
            f"the observed orders of the two finest triples differ ({coarse:.2f} and {order:.2f})"))
    p = min(order, FORMAL_ORDER)
    e21 = values[n - 2] - value
    gain = RATIO ** p - 1
    return ConvergenceEstimate("monotone", value, extrapolated=value - e21 / gain, order=order,
                               uncertainty=SAFETY_FACTOR * abs(e21) / gain)


@dataclass
class VerifiedValue:
    estimate: ConvergenceEstimate
    uncertainty: Optional[float]
    verified: bool
    levels: List[tuple] = field(default_factory=list)   # (refinement, value), coarsest first
    stopped: Optional[tuple] = None                      # (refinement, reason)


def refine_until_verified(solve_at: Callable[[int], float], *, tolerance: float, floor: float, max_refinement: int,
                          order_tolerance: float, noise: float,
                          stop_on: Optional[Callable[[BaseException], Optional[str]]] = None) -> VerifiedValue:
    """grid-convergence.ts:253-283 refineUntilVerified: solve on grids
    refined by 2 until the error band (the estimate's uncertainty plus
    `floor` times the value) is known and within `tolerance`, or the finest
    refinement is reached. A failure on a grid after the first stops the
    refinement with the earlier result when `stop_on` gives a reason."""
    levels: List[tuple] = []
    result: Optional[VerifiedValue] = None
    r = 1
    while r <= max_refinement:
        try:
            value = solve_at(r)
        except Exception as e:  # noqa: BLE001 - the caller's stop rule decides
            reason = None if result is None or stop_on is None else stop_on(e)
            if result is None or reason is None:
                raise
            return VerifiedValue(result.estimate, result.uncertainty, result.verified, list(result.levels),
                                 (r, reason))
        levels.append((r, value))
        est = grid_convergence([v for _, v in levels], order_tolerance, noise)
        unc = None if est.kind == "unresolved" else est.uncertainty + floor * abs(est.value)
        verified = unc is not None and unc <= tolerance
        result = VerifiedValue(est, unc, verified, list(levels))
        if verified:
            break
        r *= 2
    if result is None:
        raise ValueError("the finest refinement must be at least 1")
    return result
