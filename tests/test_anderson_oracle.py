"""The Anderson oracle (oracle/anderson.py, a restatement of
src/solver/anderson.ts) pinned by what the mathematics fixes: on a linear
fixed-point iteration Anderson mixing is GMRES on its residuals, so with a
history as deep as the dimension it reaches the solution in at most
dimension + 1 steps; gamma is the least-squares fit of f by dF (numpy's
lstsq); a column that repeats another is dropped, leaving the fit of the
others; reset forgets the history."""
import numpy as np
import pytest

from anderson import AndersonMixing


def linear_problem(n, seed):
    rng = np.random.default_rng(seed)
    a = np.eye(n) + 0.3 * rng.normal(size=(n, n))
    b = rng.normal(size=n)
    return a, b


def test_finite_termination_on_a_linear_iteration():
    n = 6
    a, b = linear_problem(n, 3)
    exact = np.linalg.solve(a, b)
    omega = 0.2

    def f(x):
        return omega * (b - a @ x)

    mix = AndersonMixing(n, depth=n + 2)
    x = np.zeros(n)
    for _ in range(n + 2):
        x = mix.next(x, f(x))
    assert np.abs(x - exact).max() < 1e-9
    plain = np.zeros(n)
    for _ in range(n + 2):
        plain = plain + f(plain)
    assert np.abs(plain - exact).max() > 1e-3


def test_gamma_is_the_least_squares_fit():
    rng = np.random.default_rng(5)
    n, depth = 40, 4
    mix = AndersonMixing(n, depth)
    xs = [rng.normal(size=n) for _ in range(7)]
    fs = [rng.normal(size=n) for _ in range(7)]
    for x, f in zip(xs, fs):
        out = mix.next(x, f)
    df = np.stack([fs[k + 1] - fs[k] for k in range(2, 6)], axis=1)
    dx = np.stack([xs[k + 1] - xs[k] for k in range(2, 6)], axis=1)
    gamma, *_ = np.linalg.lstsq(df, fs[-1], rcond=None)
    assert np.allclose(mix.gamma, gamma, rtol=1e-10, atol=1e-12)
    assert np.allclose(out, xs[-1] + fs[-1] - (dx + df) @ gamma, rtol=1e-10, atol=1e-12)


def test_a_repeated_column_is_dropped_with_the_older_ones():
    """Two equal f differences: the columns from the oldest through the
    first copy are dropped, leaving the fit by the newer ones."""
    rng = np.random.default_rng(9)
    n = 30
    mix = AndersonMixing(n, 4)
    base = rng.normal(size=n)
    step = rng.normal(size=n)
    fs = [base, base + step, base + 2 * step, base + 2 * step + rng.normal(size=n)]
    xs = [rng.normal(size=n) for _ in fs]
    for x, f in zip(xs, fs):
        mix.next(x, f)
    # Columns: f1-f0 = step, f2-f1 = step (dependent), f3-f2 = random.
    assert mix.gamma[0] == 0.0
    newer = np.stack([fs[2] - fs[1], fs[3] - fs[2]], axis=1)
    g, *_ = np.linalg.lstsq(newer, fs[-1], rcond=None)
    assert np.allclose(mix.gamma[1:], g, rtol=1e-10, atol=1e-12)


def test_reset_forgets_and_sizes_are_checked():
    mix = AndersonMixing(3, 2)
    mix.next(np.ones(3), np.ones(3))
    mix.next(2 * np.ones(3), np.zeros(3) + 0.5)
    mix.reset()
    out = mix.next(np.zeros(3), np.ones(3))
    assert np.array_equal(out, np.ones(3))
    with pytest.raises(ValueError):
        mix.next(np.zeros(2), np.zeros(2))
    with pytest.raises(ValueError):
        AndersonMixing(3, 0)
