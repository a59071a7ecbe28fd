"""Bounds of every global index the stencil tile forms (tile_body and
load_cell in bowshock_amd/csrc/bs_kernels.hpp), replayed on the CPU for
every thread of every tile: the state, metric, time-step and boundary
arrays it reads and the arrays it writes, on grids whose last tiles are
partial in both directions, with walls, face states and periodic sides,
for the FP32 tile (32 x 8, also the batch's large FP32 tile), the FP64
tile (32 x 8, with the FP64 stage's far-face metric reads; also 64 x 4),
the batch's small tile (32 x 4) and a 64 x 8 tile. The index arithmetic is
the kernel's, restated; each access is checked against its array's length. Among the grids is the
412 x 80 blunt-body C grid (walls on iMin and jMin) of the fitted
workload."""
import numpy as np
import pytest

IMIN, IMAX, JMIN, JMAX = range(4)
INFLOW, OUTFLOW, WALL, PERIODIC, FACE_STATES, FAR_FIELD = range(6)


def load_cell_indices(i, j, ni, nj, kinds):
    """(array, index) pairs load_cell reads for cell or ghost (i, j)."""
    out = []
    if 0 <= i < ni and 0 <= j < nj:
        return [("state", i + j * ni)]
    if i < 0:
        side, t, mi, mj, wi, wj = IMIN, j, -1 - i, j, i + ni, j
        out.append(("iface", j * (ni + 1)))
    elif i >= ni:
        side, t, mi, mj, wi, wj = IMAX, j, 2 * ni - 1 - i, j, i - ni, j
        out.append(("iface", ni + j * (ni + 1)))
    elif j < 0:
        side, t, mi, mj, wi, wj = JMIN, i, i, -1 - j, i, j + nj
        out.append(("jface", i))
    else:
        side, t, mi, mj, wi, wj = JMAX, i, i, 2 * nj - 1 - j, i, j - nj
        out.append(("jface", i + nj * ni))
    kind = kinds[side]
    if kind == FACE_STATES:
        out.append((f"face{side}", 4 * t + 3))
        return out
    if kind == INFLOW:
        return out
    if kind == FAR_FIELD:
        ii = 0 if side == IMIN else ni - 1 if side == IMAX else i
        jj = 0 if side == JMIN else nj - 1 if side == JMAX else j
        out.append(("state", ii + jj * ni))
        return out
    c = wi + wj * ni if kind == PERIODIC else mi + mj * ni
    out.append(("state", c))
    return out


def tile_accesses(ni, nj, TX, TY, kinds, fp64_stage):
    """Every (array, index) the tile kernel touches over the whole grid."""
    tiles_x, tiles_y = -(-ni // TX), -(-nj // TY)
    acc = []
    for by in range(tiles_y):
        for bx in range(tiles_x):
            i0, j0 = bx * TX, by * TY
            inner = i0 >= 2 and j0 >= 2 and i0 + TX + 2 <= ni and j0 + TY + 2 <= nj
            for tid in range(TX * TY):
                li, lj = tid % TX, tid // TX
                gi, gj = i0 + li, j0 + lj
                valid = gi < ni and gj < nj
                if gi <= ni and gj < nj:
                    acc.append(("iface", gi + gj * (ni + 1)))
                if gi < ni and gj <= nj:
                    acc.append(("jface", gi + gj * ni))
                if valid:
                    acc += [("cell", gi + gj * ni)]

                def fetch(i, j):
                    if inner:
                        assert 0 <= i < ni and 0 <= j < nj
                        return [("state", i + j * ni)]
                    return load_cell_indices(i, j, ni, nj, kinds)

                if not (gi >= ni and gj >= nj) and gi <= ni + 1 and gj <= nj + 1:
                    acc += fetch(gi, gj)
            for h in range(4 * (TX + TY)):
                if h < 4 * TX:
                    r = h // TX
                    wi, wj = 2 + h % TX, (r if r < 2 else TY + r)
                else:
                    r = (h - 4 * TX) // TY
                    wj, wi = 2 + (h - 4 * TX) % TY, (r if r < 2 else TX + r)
                hi, hj = i0 + wi - 2, j0 + wj - 2
                oi, oj = hi < 0 or hi >= ni, hj < 0 or hj >= nj
                if (oi and oj) or hi > ni + 1 or hj > nj + 1:
                    continue
                acc += fetch(hi, hj)
            # Faces: each thread its low i- and j-face, the edge waves the
            # tile's high faces; boundary faces may read an imposed flux.
            faces = [(0, li, lj) for li in range(TX) for lj in range(TY)]
            faces += [(1, li, lj) for li in range(TX) for lj in range(TY)]
            faces += [(1, fx, TY) for fx in range(TX)] + [(0, TX, fy) for fy in range(TY)]
            for d, fx, fy in faces:
                gi, gj = i0 + fx, j0 + fy
                if d == 0:
                    if not inner and (gi > ni or gj >= nj):
                        continue
                    acc.append(("iface", gi + gj * (ni + 1)))
                    side = IMIN if gi == 0 else IMAX if gi == ni else -1
                    if not inner and side >= 0 and kinds[side] == FACE_STATES:
                        acc.append((f"face{side}", 4 * gj + 3))
                else:
                    if not inner and (gi >= ni or gj > nj):
                        continue
                    acc.append(("jface", gi + gj * ni))
                    side = JMIN if gj == 0 else JMAX if gj == nj else -1
                    if not inner and side >= 0 and kinds[side] == FACE_STATES:
                        acc.append((f"face{side}", 4 * gi + 3))
            if fp64_stage:
                # The FP64 stage reads the time step's far faces' metrics.
                for tid in range(TX * TY):
                    gi, gj = i0 + tid % TX, j0 + tid // TX
                    if gi < ni and gj < nj:
                        acc.append(("iface", gi + 1 + gj * (ni + 1)))
                        acc.append(("jface", gi + (gj + 1) * ni))
            acc.append(("partial", bx + by * tiles_x))
    return acc, tiles_x * tiles_y


GRIDS = [
    # (ni, nj, kinds iMin, iMax, jMin, jMax)
    (412, 80, (WALL, OUTFLOW, WALL, INFLOW)),        # the fitted workload's C grid
    (412, 80, (WALL, OUTFLOW, WALL, FACE_STATES)),   # the same with a fitted-shock jMax
    (130, 63, (INFLOW, OUTFLOW, WALL, FAR_FIELD)),
    (37, 21, (PERIODIC, PERIODIC, WALL, INFLOW)),
    (2, 2, (FACE_STATES, OUTFLOW, WALL, INFLOW)),
    (66, 9, (INFLOW, OUTFLOW, PERIODIC, PERIODIC)),
]


@pytest.mark.parametrize("ni,nj,kinds", GRIDS)
@pytest.mark.parametrize("TX,TY,fp64", [(32, 8, False), (64, 8, False), (32, 8, True), (64, 4, True), (32, 4, False)])
def test_every_index_of_the_tile_is_in_bounds(ni, nj, kinds, TX, TY, fp64):
    acc, blocks = tile_accesses(ni, nj, TX, TY, kinds, fp64)
    sizes = {"state": ni * nj, "cell": ni * nj, "iface": (ni + 1) * nj, "jface": ni * (nj + 1),
             "partial": blocks + 64}
    for side in range(4):
        sizes[f"face{side}"] = 4 * (nj if side < 2 else ni)
    arr = {}
    for name, k in acc:
        arr.setdefault(name, []).append(k)
    for name, ks in arr.items():
        ks = np.array(ks)
        assert ks.min() >= 0 and ks.max() < sizes[name], (name, ks.min(), ks.max(), sizes[name])
