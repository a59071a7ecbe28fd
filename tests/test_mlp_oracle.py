"""The FP64 network oracle and the warm-start field transforms (CPU).

The oracle's backward pass is pinned by central finite differences of its
forward pass, its Adam step by the closed form of a first step, and the
resampling by exactness on bilinear fields and identity at equal sizes."""
import numpy as np
import pytest

import mlp_oracle as M
from bowshock_amd import warmstart as W


def _net(seed=0, inputs=5, hidden=7, depth=2, outputs=6):
    rng = np.random.default_rng(seed)
    n = sum(a * b + b for a, b in M.shapes(inputs, hidden, depth, outputs))
    return M.unpack(rng.normal(0, 0.6, n), inputs, hidden, depth, outputs), rng


def test_backward_matches_finite_differences():
    layers, rng = _net()
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 6))
    loss, grads = M.loss_and_grads(layers, x, t)
    flat = M.pack(layers)
    g = M.pack(grads)
    h = 1e-6
    for k in rng.choice(len(flat), 40, replace=False):
        p, q = flat.copy(), flat.copy()
        p[k] += h
        q[k] -= h
        lp = M.loss_and_grads(M.unpack(p, 5, 7, 2, 6), x, t)[0]
        lq = M.loss_and_grads(M.unpack(q, 5, 7, 2, 6), x, t)[0]
        assert g[k] == pytest.approx((lp - lq) / (2 * h), rel=1e-6, abs=1e-9)
    assert loss == pytest.approx(np.mean((M.forward(layers, x)[0] - t) ** 2))


def test_a_planted_sign_error_fails_the_difference_check():
    """The check above catches a flipped activation derivative."""
    layers, rng = _net(1)
    x = rng.normal(size=(4, 5))
    t = rng.normal(size=(4, 6))
    _, grads = M.loss_and_grads(layers, x, t)
    # Gradient of the first layer's bias with SiLU' replaced by -SiLU'.
    y, cache = M.forward(layers, x)
    dz = 2 * (y - t) / y.size
    for k in range(len(layers) - 1, 0, -1):
        z = cache[k - 1][1]
        s = 1 / (1 + np.exp(-z))
        sign = -1.0 if k == 1 else 1.0
        dz = (dz @ layers[k][0].T) * sign * (s * (1 + z * (1 - s)))
    planted = dz.sum(axis=0)
    flat = M.pack(layers)
    k0 = 5 * 7  # first bias of the first layer
    h = 1e-6
    p, q = flat.copy(), flat.copy()
    p[k0] += h
    q[k0] -= h
    fd = (M.loss_and_grads(M.unpack(p, 5, 7, 2, 6), x, t)[0] - M.loss_and_grads(M.unpack(q, 5, 7, 2, 6), x, t)[0]) / (2 * h)
    assert fd == pytest.approx(grads[0][1][0], rel=1e-6)
    assert planted[0] != pytest.approx(fd, rel=1e-3)


def test_adam_first_step_moves_by_the_learning_rate():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    q, m, v = M.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(q, p - 0.01 * np.sign(g), rtol=0, atol=1e-7)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_resample_is_exact_on_bilinear_fields_and_identity_at_equal_size():
    ni, nj = 33, 17
    ic = (np.arange(ni) + 0.5) / ni
    jc = (np.arange(nj) + 0.5) / nj
    f = np.stack([1 + 2 * ic[None, :] + 3 * jc[:, None] + 0.5 * ic[None, :] * jc[:, None]] * 4)
    assert np.allclose(W.resample(f, ni, nj), f, atol=1e-14)
    g = W.resample(f, 80, 40)
    ig = (np.arange(80) + 0.5) / 80
    jg = (np.arange(40) + 0.5) / 40
    # Inside the hull of the source centres the interpolant is exact.
    inner_i = (ig >= ic[0]) & (ig <= ic[-1])
    inner_j = (jg >= jc[0]) & (jg <= jc[-1])
    exact = 1 + 2 * ig[None, :] + 3 * jg[:, None] + 0.5 * ig[None, :] * jg[:, None]
    assert np.allclose(g[0][np.ix_(inner_j, inner_i)], exact[np.ix_(inner_j, inner_i)], atol=1e-12)


def test_target_round_trip_and_clamp():
    ni, nj = 80, 40
    rng = np.random.default_rng(3)
    prim = np.stack([np.full((nj, ni), 1.7), rng.normal(size=(nj, ni)), rng.normal(size=(nj, ni)),
                     np.full((nj, ni), 3.0)])
    t = W.to_target(prim, u_inf=2.0)
    back = W.from_target(t, 2.0, ni, nj)
    assert np.allclose(back, prim.reshape(4, -1).T)
    t2 = t.copy()
    t2[:3200] = -1.0  # negative densities
    assert W.from_target(t2, 2.0, ni, nj)[:, 0].min() == pytest.approx(0.05)


def test_snap_to_free_stream():
    """Coarse cells within the tolerance of the free stream (normalized:
    rho 1, u 1, v 0, p 1 / u_inf^2) become it exactly; any variable outside
    keeps the cell as predicted."""
    from bowshock_amd import warmstart as W

    n = W.COARSE_NI * W.COARSE_NJ
    u = np.array([2.0, 3.0])
    t = np.concatenate([np.full(n, 1.005), np.full(n, 0.99), np.full(n, 0.003), np.full(n, 0.25 * 1.01)])
    t = np.stack([t, t])
    t[0, 7] = 1.5
    s = W.snap_to_free_stream(t, u, 0.02).reshape(2, 4, n)
    assert np.all(s[0, 0, :7] == 1.0) and np.all(s[0, 2, :7] == 0.0) and np.all(s[0, 3, :7] == 0.25)
    assert s[0, 0, 7] == 1.5 and s[0, 3, 7] == t[0, 3 * n + 7]
    # At Mach 3 (u_inf = 3) the pressure 0.2525 is far from 1/9: nothing snaps.
    assert np.array_equal(s[1].ravel(), t[1])
