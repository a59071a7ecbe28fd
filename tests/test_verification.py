"""Solution verification (CPU): the settling test and the grid convergence
estimate, on the known answers of the reference's own tests
(tests/verification/iterative-convergence.test.ts,
tests/verification/grid-convergence.test.ts)."""
import math

import pytest

from bowshock_amd.verification import SettlingValue, grid_convergence, refine_until_verified
from bowshock_amd.verified_drag import GPU_NOISE


def park_miller(seed):
    """The reference tests' generator: seed * 16807 mod 2^31 - 1."""
    state = seed

    def rand():
        nonlocal state
        state = (state * 16807) % 2147483647
        return (state - 1) / 2147483646

    return rand


def settle(values, limit, noise, known_rate=math.nan, max_checks=10_000):
    s = SettlingValue(noise, known_rate)
    for check in range(max_checks):
        v = values(check)
        if s.add(v):
            return check, abs(v - limit)
    return math.inf, math.inf


def test_fast_geometric_convergence_settles_on_the_second_quiet_check():
    check, error = settle(lambda k: 1 + 0.1 * 0.1 ** k, 1, 1e-6)
    assert check == 7 and error < 1e-7


def test_slow_convergence_with_noise_settles_near_its_limit():
    """iterative-convergence.test.ts:70-94: a start 1e4 noise away, every
    value off by up to half the noise, settles within about 1.6 noise of
    its limit at rates up to 0.9, within 100 checks."""
    noise = 1e-6
    for rate, bound in ((0.5, 1.0), (0.7, 1.5), (0.9, 2.0)):
        rand = park_miller(7)
        worst, latest = 0.0, 0
        for _ in range(100):
            jitter = [0.5 * noise * (2 * rand() - 1) for _ in range(300)]
            s = SettlingValue(noise)
            for k in range(300):
                err = 1e4 * rate ** k
                if s.add(1 + err * noise + jitter[k]):
                    worst = max(worst, err)
                    latest = max(latest, k)
                    break
            else:
                pytest.fail(f"rate {rate} did not settle")
        assert latest < 100 and worst < bound, (rate, worst, latest)


def test_rates_known_and_measured():
    s = SettlingValue(1e-6, 0.1)
    for v in (1.1, 1.05, 1.025):
        s.add(v)
    assert s.rate == pytest.approx(0.5, abs=1e-12)
    assert s.rate_to_pass_on == pytest.approx(0.5, abs=1e-12)
    assert SettlingValue(1e-6, 0.3).rate_to_pass_on == 0.3


def test_a_converged_value_settles_at_the_second_change():
    assert settle(lambda k: 2, 2, 1e-6)[0] == 2
    rand = park_miller(5)
    for _ in range(100):
        assert settle(lambda k: 2 * (1 + 0.5e-6 * (2 * rand() - 1)), 2, 1e-6)[0] == 2


OPTS = dict(order_tolerance=0.3, noise=1e-12)


def family(phi, grids=4):
    return [phi(2.0 ** -k) for k in range(grids)]


def test_second_order_family():
    e = grid_convergence(family(lambda h: 0.3 + 0.04 * h * h), **OPTS)
    assert e.kind == "monotone"
    assert e.order == pytest.approx(2, abs=1e-12)
    assert e.extrapolated == pytest.approx(0.3, abs=1e-14)
    assert e.uncertainty / abs(e.value - 0.3) == pytest.approx(1.25, abs=1e-10)


def test_first_order_family_and_the_order_cap():
    e = grid_convergence(family(lambda h: -1 + 0.2 * h), **OPTS)
    assert e.kind == "monotone" and e.order == pytest.approx(1, abs=1e-12)
    assert e.uncertainty / abs(e.value + 1) == pytest.approx(1.25, abs=1e-10)
    e = grid_convergence(family(lambda h: 2 + 0.5 * h ** 3), **OPTS)
    assert e.kind == "monotone" and e.order == pytest.approx(3, abs=1e-12)
    assert e.uncertainty > 1.25 * abs(e.value - 2)


def test_coarse_grids_outside_the_asymptotic_range():
    phi = lambda h: 1 + 0.1 * h + 5 * h ** 4  # noqa: E731
    e = grid_convergence(family(phi), **OPTS)
    assert e.kind == "unresolved" and "orders of the two finest triples differ" in e.reason
    assert grid_convergence(family(phi, 9), **OPTS).kind == "monotone"


def test_oscillation_bound():
    e = grid_convergence([1, 1.1, 1.05, 1.07], **OPTS)
    assert e.kind == "oscillatory" and e.uncertainty == pytest.approx(0.05, abs=1e-12)


def test_finest_change_within_the_noise():
    noise = 1e-6
    exact = family(lambda h: 1 + 1e-4 * h * h)
    e = grid_convergence(exact, order_tolerance=0.3, noise=noise)
    assert e.kind == "monotone" and e.order == pytest.approx(2, abs=1e-6)
    assert e.uncertainty == pytest.approx((1.25 * (1e-4 * (0.25 ** 2 - 0.125 ** 2) + 2e-6)) / 3, abs=1e-9)
    rand = park_miller(3)
    for _ in range(200):
        noisy = [v * (1 + noise * (2 * rand() - 1)) for v in exact]
        r = grid_convergence(noisy, order_tolerance=0.3, noise=noise)
        assert r.kind == "monotone" and abs(r.value - 1) <= r.uncertainty
    e = grid_convergence([1.0005, 1.0001, 1, 1], order_tolerance=0.3, noise=1e-6)
    assert e.kind == "unresolved" and "below the noise but not the one the coarser grids predict" in e.reason


@pytest.mark.parametrize("values,reason", [([1, 1.001, 1.003, 1.007], "do not shrink"),
                                           ([1, 1, 1, 1], "too small to measure"),
                                           ([1.2, 1.1, 1.05], "four grids")])
def test_no_band(values, reason):
    e = grid_convergence(values, **OPTS)
    assert e.kind == "unresolved" and reason in e.reason


FAMILIES = [
    ("M 3, 20 deg", 2.79089952842, [2.78351098509, 2.78653016256, 2.78857768215, 2.78967376953]),
    ("M 2, 15 deg", 1.56629307899, [1.56674616602, 1.56616704408, 1.5661418598, 1.56619348995]),
    ("M 1.3, 10 deg", 1.16466347855, [1.16511515578, 1.16473539747, 1.16466553813, 1.16465784725]),
    ("M 1.41, 24 deg", 1.79216769814, [1.73233844783, 1.77139994048, 1.77685553456, 1.78443542194]),
]


@pytest.mark.parametrize("name,exact,values", FAMILIES)
def test_measured_cone_families(name, exact, values):
    res = refine_until_verified(lambda r: values[int(math.log2(r))], tolerance=math.inf, floor=2e-4,
                                max_refinement=8, order_tolerance=0.5, noise=1e-12)
    if res.uncertainty is not None:
        assert abs(res.estimate.value - exact) <= res.uncertainty


def test_a_finer_grid_that_fails_leaves_the_coarser_result():
    values = [1.04, 1.01, 1.0025]

    def solve_at(r):
        if r == 8:
            raise ValueError("too large for this GPU")
        return values[int(math.log2(r))]

    opts = dict(order_tolerance=0.5, noise=1e-12, floor=0, tolerance=0)
    coarser = refine_until_verified(solve_at, max_refinement=4, **opts)
    stopped = refine_until_verified(solve_at, max_refinement=8,
                                    stop_on=lambda e: str(e) if isinstance(e, ValueError) else None, **opts)
    assert stopped.stopped == (8, "too large for this GPU")
    assert stopped.levels == coarser.levels and stopped.uncertainty == coarser.uncertainty
    with pytest.raises(ValueError, match="too large"):
        refine_until_verified(solve_at, max_refinement=8, **opts)

    def refuse(r):
        raise ValueError("refused")

    with pytest.raises(ValueError, match="refused"):
        refine_until_verified(refuse, max_refinement=8, stop_on=lambda e: "any", **opts)


def test_bands_hold_with_32_bit_noise():
    rand = park_miller(1)
    for _, exact, values in FAMILIES:
        for _ in range(50):
            noisy = [v * (1 + GPU_NOISE * (2 * rand() - 1)) for v in values]
            res = refine_until_verified(lambda r: noisy[int(math.log2(r))], tolerance=math.inf, floor=2e-4,
                                        max_refinement=8, order_tolerance=0.5, noise=GPU_NOISE)
            if res.uncertainty is not None:
                assert abs(res.estimate.value - exact) <= res.uncertainty
