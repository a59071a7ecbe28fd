"""The BF16 MFMA network kernels (bs_mlp_*) against the FP64 oracle: the
forward pass and the gradients agree with an FP64 evaluation that rounds to
BF16 where the kernels do (GEMM operands, activations, backward
gradients) to FP32-accumulation tolerance, and with the plain FP64
network to BF16 tolerance; one Adam step matches the oracle's update of
the same gradients; training reduces the loss."""
import numpy as np
import pytest

import mlp_oracle as M
from bowshock_amd.warmstart import Mlp

pytestmark = pytest.mark.gpu

INPUTS, HIDDEN, DEPTH, OUTPUTS, BATCH = 9, 128, 2, 192, 128
# Shapes that reach every GEMM path: the first runs on the 64 x 64 tile;
# in the second, the output layer's weight gradient (256 x 16384, 256 tiles
# of 128 x 128, transposed A) and its input gradient (K = 16384, split
# along K) take the 128 x 128 tile.
SHAPES = [(9, 128, 2, 192, 128), (9, 256, 1, 16384, 128)]


def bf16(a):
    """Round to BF16 (nearest, ties to even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _setup(ctx, seed=0, shape=SHAPES[0]):
    inputs, hidden, depth, outputs, batch = shape
    rng = np.random.default_rng(seed)
    net = Mlp(ctx, inputs, hidden, depth, outputs, max_batch=batch)
    net.init_params(seed)
    params = net.get_params().astype(np.float64)
    # Non-zero biases so the bias path is exercised.
    layers = M.unpack(params, inputs, hidden, depth, outputs)
    layers = [(W, rng.normal(0, 0.1, b.shape)) for W, b in layers]
    params = M.pack(layers).astype(np.float32).astype(np.float64)
    net.set_params(params)
    x = rng.normal(size=(300, inputs))
    t = rng.normal(size=(300, outputs))
    net.load_data(x, t)
    return net, M.unpack(params, inputs, hidden, depth, outputs), x, t, rng


def _emulated(layers, x, t):
    """The kernels' arithmetic: BF16 operands, FP32-or-better accumulation."""
    a = bf16(np.asarray(x, dtype=np.float32))
    zs, acts = [], [a]
    for k, (W, b) in enumerate(layers):
        z = a @ bf16(W) + np.asarray(b, dtype=np.float32)
        zs.append(z)
        if k < len(layers) - 1:
            a = bf16(z / (1 + np.exp(-z)))
            acts.append(a)
    y = zs[-1]
    d = y - t
    dz = 2 * d / d.size
    grads = [None] * len(layers)
    dzb = bf16(dz)
    for k in range(len(layers) - 1, -1, -1):
        grads[k] = (acts[k].T @ dzb, dz.sum(axis=0))
        if k > 0:
            z = zs[k - 1]
            s = 1 / (1 + np.exp(-z))
            dz = (dzb @ bf16(layers[k][0]).T) * (s * (1 + z * (1 - s)))
            dzb = bf16(dz)
    return y, float(np.mean(d * d)), grads


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", SHAPES)
def test_forward_matches_oracle(ctx, shape):
    net, layers, x, t, _ = _setup(ctx, shape=shape)
    y = net.predict(x[:200]).astype(np.float64)  # 200 rows: two blocks, the second padded
    y_emu = _emulated(layers, x[:200], t[:200])[0]
    y64 = M.forward(layers, x[:200])[0]
    assert np.max(np.abs(y - y_emu)) < 2e-4 * np.max(np.abs(y_emu)) + 1e-5
    assert _rel(y, y64) < 2e-2


@pytest.mark.parametrize("shape", SHAPES)
def test_gradients_match_oracle(ctx, shape):
    net, layers, x, t, rng = _setup(ctx, 1, shape)
    idx = rng.choice(300, BATCH, replace=False)
    loss, g = net.gradients(idx)
    _, loss_emu, g_emu = _emulated(layers, x[idx], t[idx])
    loss64, g64 = M.loss_and_grads(layers, x[idx], t[idx])
    assert loss == pytest.approx(loss_emu, rel=1e-4)
    assert loss == pytest.approx(loss64, rel=1e-2)
    flat_emu = M.pack(g_emu)
    flat64 = M.pack(g64)
    assert _rel(g, flat_emu) < 2e-3
    assert _rel(g, flat64) < 3e-2
    # Layer by layer, so an error in one GEMM shape cannot hide in the norm.
    o = 0
    for (W, b), (W64, b64) in zip(g_emu, g64):
        gw = g[o:o + W.size].reshape(W.shape)
        o += W.size
        gb = g[o:o + b.size]
        o += b.size
        assert _rel(gw, W) < 5e-3 and _rel(gb, b) < 5e-3


def test_adam_step_matches_oracle(ctx):
    net, layers, x, t, rng = _setup(ctx, 2)
    idx = rng.choice(300, BATCH, replace=False)
    p0 = net.get_params().astype(np.float64)
    _, g = net.gradients(idx)
    net.train_batch(idx, 1e-3)
    p1 = net.get_params().astype(np.float64)
    q, _, _ = M.adam(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.max(np.abs(p1 - q)) < 1e-6


def test_training_reduces_the_loss(ctx):
    net, layers, x, _, rng = _setup(ctx, 3)
    teacher = M.unpack(rng.normal(0, 0.3, net.count), INPUTS, HIDDEN, DEPTH, OUTPUTS)
    t = M.forward(teacher, x)[0]
    net.load_data(x, t)
    first = None
    for s in range(300):
        loss = net.train_batch(rng.integers(0, 300, BATCH), 2e-3)
        first = first if first is not None else loss
    assert loss < 0.2 * first


def test_gemm_timing(ctx):
    net = Mlp(ctx, 9, 512, 4, 12800, max_batch=256)
    ms, flops = net.time_gemm(4, reps=20)
    assert ms > 0 and flops == 2.0 * 256 * 12800 * 512
