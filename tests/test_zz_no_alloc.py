"""The allocation-free warm start (bs_batch_set_coarse_noalloc) against
bs_batch_set_coarse, through the C ABI: the states it writes are bitwise
those of the allocating entry point on the same batches and fields, on
two calls in a row (the second with smaller coarse fields on a smaller
batch, after the scratch holds the first call's contents), on a batch
whose fields pass through the scratch in more than one group, and after a
call refused for a non-physical field (whose flag the next call clears);
a field larger than the scratch is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import cases
from bowshock_amd import _native as N
from bowshock_amd.batch import GpuBatch2D
from bowshock_amd.engine import Euler2DConfig

pytestmark = pytest.mark.gpu

SCRATCH = 8 << 20


def _configs(shapes, mach=2.5):
    out = []
    for ni, nj in shapes:
        grid, free, p_inf, q_inf, bcs = cases.cone_case(ni, nj, mach=mach)
        out.append(Euler2DConfig(grid=grid, boundaries=bcs, geometry="axisymmetric", gamma=1.4, flux="hllc",
                                 limiter="minmod", reconstruction="perVariable", timeIntegrator="ssprk3", cfl=0.5,
                                 timeStepping="local"))
    return out


def _fields(count, cni, cnj, seed):
    rng = np.random.default_rng(seed)
    f = np.empty((count, 4, cnj, cni))
    f[:, 0] = rng.uniform(0.5, 3.0, (count, cnj, cni))
    f[:, 1] = rng.uniform(0.2, 1.0, (count, cnj, cni))
    f[:, 2] = rng.uniform(-0.2, 0.2, (count, cnj, cni))
    f[:, 3] = rng.uniform(0.02, 4.0, (count, cnj, cni))  # some below the floor
    scale = np.stack([np.ones(count), rng.uniform(1.5, 4.0, count), rng.uniform(1.5, 4.0, count),
                      rng.uniform(2.0, 9.0, count)], axis=1)
    return np.ascontiguousarray(f.reshape(count, -1)), np.ascontiguousarray(scale)


def _set(entry, batch, fields, scale, cni, cnj, floor=0.05):
    return getattr(N.lib(), entry)(batch._h, N.dptr(fields), cni, cnj, N.dptr(scale), floor)


def _states(batch):
    out = []
    for k in range(batch.count):
        s = batch.readState(k)
        out.append(np.stack([s.rho, s.mu, s.mv, s.energy]))
    return out


def _pair(ctx, shapes):
    cfgs = _configs(shapes)
    return GpuBatch2D(ctx, cfgs), GpuBatch2D(ctx, cfgs)


def _same(a, b):
    sa, sb = _states(a), _states(b)
    assert len(sa) == len(sb)
    for k, (x, y) in enumerate(zip(sa, sb)):
        assert np.array_equal(x, y), k


def test_two_calls_in_a_row_equal_the_allocating_entry(ctx):
    ref, mine = _pair(ctx, [(40, 20), (72, 36), (24, 12), (130, 64), (56, 28)])
    f1, s1 = _fields(5, 80, 40, seed=1)
    assert _set("bs_batch_set_coarse", ref, f1, s1, 80, 40) == N.BS_OK
    assert _set("bs_batch_set_coarse_noalloc", mine, f1, s1, 80, 40) == N.BS_OK
    _same(ref, mine)
    first = _states(mine)
    # Second call: a smaller batch and smaller fields; the scratch still
    # holds the first call's.
    ref2, mine2 = _pair(ctx, [(32, 16), (48, 24)])
    f2, s2 = _fields(2, 20, 10, seed=2)
    assert _set("bs_batch_set_coarse", ref2, f2, s2, 20, 10) == N.BS_OK
    assert _set("bs_batch_set_coarse_noalloc", mine2, f2, s2, 20, 10) == N.BS_OK
    _same(ref2, mine2)
    # The first batch kept its state; the fields were applied (the start
    # is not the free stream).
    for x, y in zip(first, _states(mine)):
        assert np.array_equal(x, y)
    assert np.ptp(first[0][0]) > 0.5
    for b in (ref, mine, ref2, mine2):
        b.destroy()


def test_fields_through_the_scratch_in_groups(ctx):
    cni, cnj = 80, 40
    per = 4 * cni * cnj * 8
    count = (SCRATCH - 256) // per + 9  # one group full, a second with 9 cases
    ref, mine = _pair(ctx, [(8 + 2 * (k % 5), 4 + (k % 3)) for k in range(count)])
    f, s = _fields(count, cni, cnj, seed=3)
    assert _set("bs_batch_set_coarse", ref, f, s, cni, cnj) == N.BS_OK
    assert _set("bs_batch_set_coarse_noalloc", mine, f, s, cni, cnj) == N.BS_OK
    _same(ref, mine)
    ref.destroy()
    mine.destroy()


def test_a_refused_field_then_a_good_one(ctx):
    ref, mine = _pair(ctx, [(40, 20), (24, 12)])
    f, s = _fields(2, 16, 8, seed=4)
    bad = f.copy()
    bad[1, 5] = np.nan
    assert _set("bs_batch_set_coarse_noalloc", mine, bad, s, 16, 8) == N.BS_ERR_NONPHYSICAL
    assert _set("bs_batch_set_coarse", ref, bad, s, 16, 8) == N.BS_ERR_NONPHYSICAL
    assert _set("bs_batch_set_coarse", ref, f, s, 16, 8) == N.BS_OK
    assert _set("bs_batch_set_coarse_noalloc", mine, f, s, 16, 8) == N.BS_OK
    _same(ref, mine)
    ref.destroy()
    mine.destroy()


def test_a_field_larger_than_the_scratch_is_refused(ctx):
    (b,) = [GpuBatch2D(ctx, _configs([(24, 12)]))]
    cni, cnj = 1024, 257  # 4 * 1024 * 257 doubles = 8.4 MB
    f = np.ones((1, 4 * cni * cnj))
    s = np.ones((1, 4))
    assert _set("bs_batch_set_coarse_noalloc", b, f, s, cni, cnj) == N.BS_ERR_RANGE
    assert "scratch holds at most" in N.last_error()
    b.destroy()
