"""Learned warm start for batched solves (bs_mlp_* in include/bowshock.h).

No counterpart in the reference, which starts every solve from the free
stream (verified-drag.ts:225-240) or from a coarser grid's solution
(transonic-flow.ts:157-168). A multilayer perceptron maps a case's
descriptors to its converged primitive field sampled on a fixed
computational grid (80 x 40 in (i / ni, j / nj)); a new case starts from
the prediction, interpolated to its own grid and clamped to physical
values, unless that start has a larger density residual than the free
stream.

Inputs: one-hot nose kind (5), fineness, diameter (m), Mach number and
cylinder length (calibers). The cylinder length is an input because the
population varies it and the computational coordinates stretch with it.
Outputs: rho / rho_inf, u / u_inf, v / u_inf and p / (rho_inf u_inf^2) on
the coarse grid, standardized per output with the training set's mean and
deviation.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .engine import Context, RangeError, check

COARSE_NI, COARSE_NJ = 80, 40


class Mlp:
    """The device network: BF16 MFMA GEMMs, FP32 master weights, Adam."""

    def __init__(self, ctx: Context, inputs: int, hidden: int = 512, depth: int = 4, outputs: int = 12800,
                 max_batch: int = 256):
        h = C.c_void_p()
        check(N.lib().bs_mlp_create(ctx._h, inputs, hidden, depth, outputs, max_batch, C.byref(h)))
        self._h = h
        self.inputs, self.hidden, self.depth, self.outputs, self.max_batch = inputs, hidden, depth, outputs, max_batch
        self.count = int(N.lib().bs_mlp_param_count(h))

    def destroy(self):
        if getattr(self, "_h", None):
            N.lib().bs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def init_params(self, seed: int = 0):
        """Uniform weights of variance 1 / fan-in (SiLU layers at gain 2),
        zero biases."""
        rng = np.random.default_rng(seed)
        dims = [self.inputs] + [self.hidden] * self.depth + [self.outputs]
        parts = []
        for k in range(self.depth + 1):
            a, b = dims[k], dims[k + 1]
            gain = 2.0 if k < self.depth else 1.0
            lim = math.sqrt(3.0 * gain / a)
            parts += [rng.uniform(-lim, lim, size=a * b), np.zeros(b)]
        self.set_params(np.concatenate(parts))

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.count:
            raise RangeError(f"expected {self.count} parameters")
        check(N.lib().bs_mlp_set_params(self._h, N.fptr(p), p.size))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.count, dtype=np.float32)
        check(N.lib().bs_mlp_get_params(self._h, N.fptr(p), p.size))
        return p

    def load_data(self, x, y):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._y = np.ascontiguousarray(y, dtype=np.float32)
        if self._x.shape[1] != self.inputs or self._y.shape[1] != self.outputs or len(self._x) != len(self._y):
            raise RangeError("data shapes do not match the network")
        check(N.lib().bs_mlp_load_data(self._h, N.fptr(self._x), N.fptr(self._y), len(self._x)))

    def gradients(self, idx):
        i = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.empty(self.count, dtype=np.float32)
        loss = C.c_double()
        check(N.lib().bs_mlp_gradients(self._h, i.ctypes.data_as(C.POINTER(C.c_int32)), len(i), N.fptr(g),
                                       C.byref(loss)))
        return loss.value, g

    def train_batch(self, idx, lr: float) -> float:
        i = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_double()
        check(N.lib().bs_mlp_train_batch(self._h, i.ctypes.data_as(C.POINTER(C.c_int32)), len(i), float(lr),
                                         C.byref(loss)))
        return loss.value

    def predict(self, x) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((len(xx), self.outputs), dtype=np.float32)
        check(N.lib().bs_mlp_predict(self._h, N.fptr(xx), len(xx), N.fptr(y)))
        return y

    def time_gemm(self, layer: int, reps: int = 50):
        ms, flops = C.c_double(), C.c_double()
        check(N.lib().bs_mlp_time_gemm(self._h, layer, reps, C.byref(ms), C.byref(flops)))
        return ms.value, flops.value


# ---------------------------------------------------------------------------
# Fields on the computational grid
# ---------------------------------------------------------------------------

def _axis_weights(n_src: int, n_dst: int):
    """Linear interpolation between cell centres (k + 1/2) / n of two
    uniform 1-D grids over [0, 1], clamped at the ends."""
    s = (np.arange(n_dst) + 0.5) / n_dst * n_src - 0.5
    s = np.clip(s, 0.0, n_src - 1.0)
    i0 = np.minimum(np.floor(s).astype(int), max(n_src - 2, 0))
    w = s - i0
    i1 = np.minimum(i0 + 1, n_src - 1)
    return i0, i1, w


def resample(field: np.ndarray, ni: int, nj: int) -> np.ndarray:
    """Bilinear resampling of a (channels, nj_src, ni_src) cell field to
    (channels, nj, ni), both grids uniform in (i / ni, j / nj)."""
    c, sj, si = field.shape
    i0, i1, wi = _axis_weights(si, ni)
    j0, j1, wj = _axis_weights(sj, nj)
    rows = field[:, j0, :] * (1 - wj)[None, :, None] + field[:, j1, :] * wj[None, :, None]
    return rows[:, :, i0] * (1 - wi)[None, None, :] + rows[:, :, i1] * wi[None, None, :]


def primitives_from_state(state, ni: int, nj: int, gamma: float = 1.4) -> np.ndarray:
    """(4, nj, ni) primitives from a GpuState2D (pressure with the FP32
    gamma for an FP32 state, gpuPrimitive2D, euler2d-gpu.ts:136-146; the
    gamma itself for an FP64 state)."""
    g = float(gamma) if state.rho.dtype == np.float64 else float(np.float32(gamma))
    rho = state.rho.astype(np.float64)
    u = state.mu / rho
    v = state.mv / rho
    p = (g - 1) * (state.energy - 0.5 * (state.mu.astype(np.float64) ** 2 + state.mv.astype(np.float64) ** 2) / rho)
    return np.stack([rho, u, v, p]).reshape(4, nj, ni)


def to_target(prim: np.ndarray, u_inf: float) -> np.ndarray:
    """Free-stream-normalized coarse field, flattened (rho, u, v, p)."""
    coarse = resample(prim, COARSE_NI, COARSE_NJ)
    scale = np.array([1.0, u_inf, u_inf, u_inf * u_inf])[:, None, None]
    return (coarse / scale).ravel()


def snap_to_free_stream(targets: np.ndarray, u_inf, tol: float) -> np.ndarray:
    """Predicted coarse fields (count, 4 * COARSE_NJ * COARSE_NI) with every
    coarse cell whose four normalized primitives all lie within `tol` of
    the free stream's (rho = 1, u = 1, v = 0, p = 1 / u_inf^2 after the
    scaling of to_target) set to the free stream exactly: the region ahead
    of the shock then starts as a free-stream start does."""
    t = np.array(targets, dtype=np.float64).reshape(len(targets), 4, -1)
    u = np.asarray(u_inf, dtype=np.float64).reshape(-1, 1)
    free = [np.ones_like(u), np.ones_like(u), np.zeros_like(u), 1.0 / (u * u)]
    near = np.ones(t[:, 0].shape, dtype=bool)
    for q in range(4):
        ref = free[q]
        scale = np.maximum(np.abs(ref), 1.0 / (u * u) if q == 3 else 1.0)
        near &= np.abs(t[:, q] - ref) <= tol * scale
    for q in range(4):
        t[:, q] = np.where(near, free[q], t[:, q])
    return t.reshape(len(targets), -1)


def from_target(t: np.ndarray, u_inf: float, ni: int, nj: int, floor: float = 0.05) -> np.ndarray:
    """(ni nj, 4) primitives on a case's grid from a predicted coarse field,
    density and pressure clamped to at least `floor` of the free stream's
    (rho_inf = p_inf = 1)."""
    coarse = t.reshape(4, COARSE_NJ, COARSE_NI) * np.array([1.0, u_inf, u_inf, u_inf * u_inf])[:, None, None]
    fine = resample(coarse, ni, nj)
    fine[0] = np.maximum(fine[0], floor)
    fine[3] = np.maximum(fine[3], floor)
    return np.ascontiguousarray(fine.reshape(4, -1).T)


@dataclass
class Standardizer:
    mean: np.ndarray
    std: np.ndarray

    @staticmethod
    def fit(a: np.ndarray) -> "Standardizer":
        m = a.mean(axis=0)
        s = a.std(axis=0)
        return Standardizer(m, np.where(s > 1e-8, s, 1.0))

    def apply(self, a):
        return (a - self.mean) / self.std

    def invert(self, a):
        return a * self.std + self.mean


@dataclass
class WarmStart:
    net: Mlp
    xs: Standardizer
    ys: Standardizer

    def predict_targets(self, features: np.ndarray) -> np.ndarray:
        return self.ys.invert(self.net.predict(self.xs.apply(features)).astype(np.float64))


def train(ctx: Context, features: np.ndarray, targets: np.ndarray, *, steps: int = 3000, batch: int = 256,
          lr: float = 1e-3, seed: int = 0, hidden: int = 512, depth: int = 4, log=None) -> WarmStart:
    """Adam on mini-batches drawn with a seeded generator, the learning
    rate decayed by a cosine to 5% of its start."""
    xs = Standardizer.fit(features)
    ys = Standardizer.fit(targets)
    net = Mlp(ctx, features.shape[1], hidden, depth, targets.shape[1], max_batch=batch)
    net.init_params(seed)
    net.load_data(xs.apply(features), ys.apply(targets))
    rng = np.random.default_rng(seed + 1)
    n = len(features)
    for s in range(steps):
        idx = rng.integers(0, n, size=batch)
        rate = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * s / steps)))
        loss = net.train_batch(idx, rate)
        if log and (s % 500 == 0 or s == steps - 1):
            log(s, loss)
    return WarmStart(net, xs, ys)


def time_gemms(ctx: Context, batch: int = 256, reps: int = 50, inputs: int = 9, hidden: int = 512, depth: int = 4,
               outputs: int = 12800):
    """Device time per forward GEMM of each layer of the warm-start network
    (bs_mlp_time_gemm, HIP events around `reps` back-to-back launches):
    [(layer, microseconds, TFLOP/s)]."""
    net = Mlp(ctx, inputs, hidden, depth, outputs, max_batch=batch)
    net.init_params(0)
    out = []
    for layer in range(depth + 1):
        ms, flops = net.time_gemm(layer, reps)
        out.append((layer, ms * 1e3, flops / (ms * 1e-3) / 1e12))
    net.destroy()
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="Time the warm-start network's forward GEMMs on the GPU.")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--reps", type=int, default=50)
    args = ap.parse_args()
    for layer, us, tflops in time_gemms(Context(0), args.batch, args.reps):
        print(f"batch {args.batch} layer {layer}: {us:.2f} us, {tflops:.1f} TFLOP/s", flush=True)
