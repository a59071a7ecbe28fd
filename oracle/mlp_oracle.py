"""FP64 NumPy restatement of the warm-start network of include/bowshock.h
(bs_mlp_*): forward pass, mean-squared-error loss, backward pass and one
Adam step.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the BF16
MFMA kernels. The product never imports it. The reference has no learned
components, so nothing here follows a reference file; the backward pass is
pinned by central finite differences of the forward pass
(tests/test_mlp_oracle.py).

Network: inputs -> hidden (SiLU) x depth -> outputs (linear),
Z_l = A_{l-1} W_l + b_l, A_l = SiLU(Z_l) = Z_l sigmoid(Z_l); loss
mean((Y - T)^2) over the batch and the outputs. Parameters in the library's
order: per layer W [in][out] row-major, then b [out].
"""
from __future__ import annotations

import numpy as np


def shapes(inputs, hidden, depth, outputs):
    dims = [inputs] + [hidden] * depth + [outputs]
    return [(dims[k], dims[k + 1]) for k in range(depth + 1)]


def unpack(params, inputs, hidden, depth, outputs):
    out, o = [], 0
    for a, b in shapes(inputs, hidden, depth, outputs):
        W = np.asarray(params[o:o + a * b], dtype=np.float64).reshape(a, b)
        o += a * b
        bias = np.asarray(params[o:o + b], dtype=np.float64)
        o += b
        out.append((W, bias))
    assert o == len(params)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(layers, x):
    """Returns the output and the per-layer (input, pre-activation)."""
    a = np.asarray(x, dtype=np.float64)
    cache = []
    for k, (W, b) in enumerate(layers):
        z = a @ W + b
        cache.append((a, z))
        a = z if k == len(layers) - 1 else z * _sig(z)
    return a, cache


def loss_and_grads(layers, x, t):
    y, cache = forward(layers, x)
    t = np.asarray(t, dtype=np.float64)
    d = y - t
    loss = float(np.mean(d * d))
    dz = 2.0 * d / d.size
    grads = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        a, _ = cache[k]
        W, _ = layers[k]
        grads[k] = (a.T @ dz, dz.sum(axis=0))
        if k > 0:
            z_prev = cache[k - 1][1]
            s = _sig(z_prev)
            dz = (dz @ W.T) * (s * (1.0 + z_prev * (1.0 - s)))
    return loss, grads


def adam(params, grads, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new params, m, v."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
