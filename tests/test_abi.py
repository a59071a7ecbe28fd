"""The C ABI library (CPU): it loads, exports every function that
include/bowshock.h declares, and its host-side pieces (metrics, argument
checks) behave without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cases
import pyoracle as O
from bowshock_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bowshock.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bs_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(N.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.SIGNATURES), "the ctypes table and the header disagree"


def test_version_and_last_error_need_no_gpu():
    assert b"gfx950" in N.lib().bs_version()
    assert N.lib().bs_last_error() is not None


def test_metrics_from_the_library_match_the_oracle():
    """bs_compute_metrics is the host FP64 metric computation the solver
    uploads from (structured.ts:94-168); stateless, no GPU."""
    from bowshock_amd.engine import compute_metrics

    g = cases.distorted_grid(13, 9)
    for geometry in ("planar", "axisymmetric"):
        a = compute_metrics(g, geometry)
        b = O.compute_metrics(g, geometry == "axisymmetric")
        pairs = {"iNx": "inx", "iNy": "iny", "iArea": "iarea", "jNx": "jnx", "jNy": "jny", "jArea": "jarea",
                 "volume": "volume", "planarArea": "planar", "centroidX": "cx", "centroidY": "cy"}
        for k, v in pairs.items():
            assert np.array_equal(a[k], b[v]), k


def test_metrics_reject_bad_grids():
    from bowshock_amd.engine import RangeError, compute_metrics
    from bowshock_amd.grids import StructuredGrid

    g = cases.distorted_grid(4, 4)
    with pytest.raises(RangeError):
        compute_metrics(StructuredGrid(4, 4, g.x, g.y - 1.0), "axisymmetric")  # below the axis
    flipped = StructuredGrid(4, 4, g.x[::-1].copy(), g.y.copy())
    with pytest.raises(RangeError):
        compute_metrics(flipped, "planar")  # inverted cells


def test_context_without_a_device_is_an_error_not_a_fallback():
    """On a machine with no GPU the context fails loudly; nothing runs on
    the CPU instead."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "from bowshock_amd.engine import Context, GpuExecutionError, RangeError\n"
            "try:\n    Context(0)\n    print('device')\n"
            "except (GpuExecutionError, RangeError) as e:\n    print('error', e)\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    assert out.startswith("device") or out.startswith("error")


def test_solver_create_rejects_null_arguments():
    h = C.c_void_p()
    rc = N.lib().bs_solver_create(None, None, C.byref(h))
    assert rc == N.BS_ERR_RANGE
    assert N.last_error()


def test_wrappers_refuse_arrays_of_the_wrong_length_before_the_library():
    """A short array would be read past its end by the C entry points, so
    the facades refuse it (line-grid.ts:301 throws RangeError for heights
    of another length); the C entry points check the count again."""
    from bowshock_amd.batch import GpuBatch2D
    from bowshock_amd.engine import GpuEuler2D, RangeError

    g = object.__new__(GpuEuler2D)
    g.ni, g.nj = 4, 3
    g._h = None
    with pytest.raises(RangeError, match="5 heights"):
        g.setHeights(np.ones(4))
    with pytest.raises(RangeError, match="5 heights"):
        g.moveGrid(np.ones(6))
    with pytest.raises(RangeError, match="12 values"):
        g.setState(np.ones(12), np.ones(12), np.ones(11), np.ones(12))
    b = object.__new__(GpuBatch2D)
    b.count = 3
    b._h = None
    with pytest.raises(RangeError, match="one flag per case"):
        b.setActive([1, 0])


def test_entry_points_refuse_a_wrong_count_without_a_solver():
    """The count arguments are checked after the handle: a null handle is
    refused first, and no entry point dereferences the arrays."""
    h = np.ones(4)
    assert N.lib().bs_set_heights(None, N.dptr(h), 4) == N.BS_ERR_RANGE
    assert N.lib().bs_batch_set_active(None, None, 0) == N.BS_ERR_RANGE
