"""The FP64 reciprocal, quotient, square root and reciprocal square root of
the device code (bs_numerics.hpp: bs_rcp / bs_div / bs_sqrt / bs_rsqrt for
double) start from the hardware estimates v_rcp_f64 / v_rsq_f64 and refine
them by Newton-Raphson and Goldschmidt iterations. This checks the
iteration scheme itself on the CPU: every fused multiply-add is evaluated
exactly (rational arithmetic, one rounding), the starting estimates carry
a relative error of up to 2^-22 (coarser than the hardware's), and the
results must lie within the stated bound of the correctly rounded ones
(reciprocal, quotient and square root within 1 ulp, reciprocal square
root within 2)."""
import math
import random
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np

EST_ERR = 2.0 ** -22
SAMPLES = 400


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def mul(a, b):
    return float(Fraction(a) * Fraction(b))


def ulps(x, ref):
    return abs(int(np.float64(x).view(np.int64)) - int(np.float64(ref).view(np.int64)))


def estimate(exact, rng):
    return float(exact * (1.0 + rng.uniform(-EST_ERR, EST_ERR)))


def rcp(x, r):
    r = fma(r, fma(-x, r, 1.0), r)
    return fma(r, fma(-x, r, 1.0), r)


def div(a, b, r0):
    r = rcp(b, r0)
    q = mul(a, r)
    return fma(r, fma(-b, q, a), q)


def sqrt(x, y):
    # The device forms g = 0 * inf = NaN for x = 0 and then selects x.
    if x == 0.0:
        return x
    g, h = mul(x, y), mul(0.5, y)
    r = fma(-g, h, 0.5)
    g = fma(g, r, g)
    h = fma(h, r, h)
    g = fma(fma(-g, g, x), h, g)
    return fma(fma(-g, g, x), h, g)


def rsqrt(x, y):
    y = fma(mul(0.5, y), fma(mul(-x, y), y, 1.0), y)
    return fma(mul(0.5, y), fma(mul(-x, y), y, 1.0), y)


def operands(rng):
    # The magnitudes of a nondimensional or SI state: densities, pressures,
    # energies, areas, volumes and wave speeds.
    return [10.0 ** rng.uniform(-8, 8) for _ in range(SAMPLES)]


def test_reciprocal_and_quotient_within_one_ulp():
    rng = random.Random(7)
    worst = 0
    for x in operands(rng):
        r = rcp(x, estimate(1.0 / x, rng))
        worst = max(worst, ulps(r, float(Fraction(1) / Fraction(x))))
        a = 10.0 ** rng.uniform(-8, 8) * rng.choice((-1.0, 1.0))
        q = div(a, x, estimate(1.0 / x, rng))
        worst = max(worst, ulps(q, float(Fraction(a) / Fraction(x))))
    assert worst <= 1


def test_square_root_within_one_ulp_and_zero_exact():
    rng = random.Random(11)
    worst = 0
    for x in operands(rng):
        s = sqrt(x, estimate(1.0 / math.sqrt(x), rng))
        worst = max(worst, ulps(s, math.sqrt(x)))  # math.sqrt is correctly rounded
    assert worst <= 1
    assert sqrt(0.0, math.inf) == 0.0


def test_reciprocal_square_root_within_two_ulps():
    getcontext().prec = 60
    rng = random.Random(13)
    worst = 0
    for x in operands(rng):
        ref = float(1 / Decimal(x).sqrt())
        worst = max(worst, ulps(rsqrt(x, estimate(ref, rng)), ref))
    assert worst <= 2


def test_one_iteration_fewer_is_not_enough():
    # The bound depends on the number of steps: with one Newton step the
    # reciprocal keeps an error of about 2^-44, hundreds of ulps.
    x = 3.7
    r0 = float((1.0 / x) * (1.0 + EST_ERR))
    r1 = fma(r0, fma(-x, r0, 1.0), r0)
    assert ulps(r1, float(Fraction(1) / Fraction(x))) > 16
    assert ulps(rcp(x, r0), float(Fraction(1) / Fraction(x))) <= 1
