"""wr_prepare_cells_noalloc: the prepare phase of one outer step on the
caller's device array of CollapseState records with no allocation by the
library (its partials in a fixed device array).  Through the C ABI, two calls
in a row -- the second on a smaller grid -- must give wr_prepare's
(dt_grid, limiting cell, status) and prepared cells bit for bit."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASK = np.ones(216, bool)
MASK[156:160] = False


class _Hip:
    """The HIP runtime the library links (for the test's own device buffers)."""

    def __init__(self, wr):
        # the runtime instance the library was linked against (by its soname,
        # already loaded): another copy in the process (a framework's bundled
        # one) would see a different device state
        wr.library()
        self.lib = ctypes.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | ctypes.RTLD_GLOBAL)
        L = self.lib
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.lib.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), src, arr.nbytes, 2) == 0


def _noalloc(wr, cells, completed, grid_time, step, perturb=True):
    lib = wr.library()
    hip = _Hip(wr)
    cells = np.ascontiguousarray(cells)
    d_cells, d_dt = hip.alloc(cells.nbytes), hip.alloc(8)
    d_cell, d_fail = hip.alloc(4), hip.alloc(4)
    try:
        hip.h2d(d_cells, cells)
        r = lib.wr_prepare_cells_noalloc(d_cells, len(cells), completed, grid_time, step,
                                         int(perturb), d_dt, d_cell, d_fail, None)
        assert r == 1, lib.wr_last_error()
        assert hip.lib.hipDeviceSynchronize() == 0
        got = np.zeros(len(cells), dtype=wr.CELL_DTYPE)
        dt, cell, fail = np.zeros(1), np.zeros(1, np.int32), np.zeros(1, np.int32)
        hip.d2h(got, d_cells)
        hip.d2h(dt, d_dt)
        hip.d2h(cell, d_cell)
        hip.d2h(fail, d_fail)
    finally:
        for p in (d_cells, d_dt, d_cell, d_fail):
            hip.lib.hipFree(p)
    return float(dt[0]), int(cell[0]), int(fail[0]), got


def _with_context(wr, cells, completed, grid_time, step, perturb=True):
    with wr.Context(len(cells)) as ctx:
        ctx.upload(cells)
        r, dt, cell = ctx.prepare(completed, grid_time, step, perturb)
        out = np.zeros(len(cells), dtype=wr.CELL_DTYPE)
        ctx.download(out)
    return dt, cell, r, out


def test_noalloc_prepare_matches_wr_prepare(wr):
    # first call: 70,000 cells after 19 steps of a run, at step 20 (perturbation)
    n1 = 70_000
    init = wr.make_collapse_states(n1)
    with wr.Context(n1) as ctx:
        ctx.upload(init)
        info = ctx.run(perturb=True, max_steps=20)
        cells1 = np.zeros(n1, dtype=wr.CELL_DTYPE)
        ctx.download(cells1)
    t1 = info["grid_time"]
    print(f"no-alloc call 1: {n1} cells, step 20", flush=True)
    a = _noalloc(wr, cells1, 20, t1, 20)
    b = _with_context(wr, cells1, 20, t1, 20)
    assert a[2] == b[2] == 1
    assert a[0] == b[0] and a[1] == b[1] and a[1] >= 0
    assert np.array_equal(a[3].view(np.uint8).reshape(n1, 216)[:, MASK],
                          b[3].view(np.uint8).reshape(n1, 216)[:, MASK])
    # second call, smaller grid, step 0 (no perturbation): partials of the
    # first call are left in the scratch array
    n2 = 1_000
    cells2 = wr.make_recipe_states(n2, 1, 1.0, 100.0, 8000.0)
    print(f"no-alloc call 2: {n2} cells, step 0", flush=True)
    a = _noalloc(wr, cells2, 0, 0.0, 0)
    b = _with_context(wr, cells2, 0, 0.0, 0)
    assert a[2] == b[2] == 1
    assert a[0] == b[0] and a[1] == b[1]
    assert np.array_equal(a[3].view(np.uint8).reshape(n2, 216)[:, MASK],
                          b[3].view(np.uint8).reshape(n2, 216)[:, MASK])
    # a record whose time is not the grid time: BAD_INPUTS, no candidate
    bad = cells2.copy()
    bad["time"][7] = 1.0
    a = _noalloc(wr, bad, 0, 0.0, 0)
    b = _with_context(wr, bad, 0, 0.0, 0)
    assert a[2] == b[2] == -1 and a[0] == b[0] and a[1] == -1


def test_noalloc_prepare_rejects_oversized_grids(wr):
    lib = wr.library()
    one = ctypes.c_void_p(1)
    assert lib.wr_prepare_cells_noalloc(one, (1 << 24) + 1, 0, 0.0, 0, 1, one, one, one,
                                        None) == wr.runtime.ERR_ARGUMENT
    assert b"2^24" in lib.wr_last_error() or b"16777216" in lib.wr_last_error()
