"""The kernel-level ABI (wr_prepare / wr_advance, the reference's two kernels
prepare_grid_timestep_kernel and advance_collapse_gridwide_kernel,
reproducer.cpp:7696-7785) driven step by step from the host, as the
reference's host loop drives them (7859-7959), at the full 64^3 size.

After every prepare the grid-wide (dt, cell) minimum the library returns --
a block reduction over 1,024 partial blocks on the device -- must equal the
host minimum of the oracle's prepare over all 262,144 cells (the lowest cell
index holding the minimum, as `candidate < dt_grid` in 7888-7893 keeps the
first), and the prepared cells (perturbed at steps 20 and 40) must equal the
oracle's bit for bit.  The dt_grid sequence of the step-driven run must equal
wr_run's device-resident loop, and the two final grids must be identical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 45
MASK = np.ones(216, bool)
MASK[156:160] = False  # reserved padding


def _bytes(c):
    return c.view(np.uint8).reshape(len(c), 216)[:, MASK]


def test_grid_minimum_through_kernel_abi_64cube(wr, oracle_wr):
    n = 64 ** 3
    init = wr.make_collapse_states(n)
    host = init.copy()
    cells = np.zeros(n, dtype=wr.CELL_DTYPE)
    dts, limiters = [], []
    with wr.Context(n) as ctx:
        ctx.upload(init)
        grid_time, completed = 0.0, 0
        for step in range(STEPS):
            r, dt, cell = ctx.prepare(completed, grid_time, step, perturb=True)
            assert r == 1
            hdt, hcell, hf = oracle_wr.prepare_grid(host, completed, grid_time, step, True)
            assert hf == 1
            assert dt == hdt and cell == hcell, (step, dt, cell, hdt, hcell)
            ctx.download(cells)
            same = np.all(_bytes(cells) == _bytes(host), axis=1)
            assert same.all(), f"step {step}: {np.count_nonzero(~same)} prepared cells differ"
            nxt = grid_time + dt
            r, integrated = ctx.advance(completed, nxt, dt)
            assert r == 1 and integrated > 0
            # the host copy follows the device cells through the advance
            ctx.download(host)
            dts.append(dt)
            limiters.append(cell)
            grid_time, completed = nxt, completed + 1
        stepped = host.copy()
    print(f"64^3 kernel-ABI steps: {STEPS}, limiting cells {sorted(set(limiters))[:8]}, "
          f"dt {dts[0]:.6g} .. {dts[-1]:.6g}")
    with wr.Context(n) as ctx:
        ctx.upload(init)
        info = ctx.run(perturb=True, max_steps=STEPS)
        assert info["status"] == 1 and info["completed_global_steps"] == STEPS
        hist = ctx.dt_history()
        ctx.download(cells)
    assert np.array_equal(hist, np.array(dts))
    assert np.array_equal(_bytes(cells), _bytes(stepped))


def test_run_cells_loop_index_starts_at_zero(wr, oracle_wr):
    """run_cells_cuda starts its loop index (which the perturbation schedule
    reads) at 0 whatever completed count the caller passes (7859); the count
    only selects which cells take part.  Cells at completed = 7 (time 0)
    through wr_run_cells must equal the oracle's loop from the same count."""
    n = 27
    cells = wr.make_collapse_states(n)
    cells["completed_steps"] = 7
    ref = cells.copy()
    r, done = wr.run_cells(cells, perturb=True, completed_global_steps=7)
    rr, rdone = oracle_wr.run_gridwide(ref, True, 1000, completed0=7)
    assert (r, done) == (rr, rdone) == (1, rdone)
    assert np.array_equal(_bytes(cells), _bytes(ref))
    # resume semantics of wr_run: the loop index continues from the count
    with wr.Context(n) as ctx:
        c2 = wr.make_collapse_states(n)
        ctx.upload(c2)
        a = ctx.run(perturb=True, max_steps=30, completed=0, grid_time=0.0, first_step=0)
        assert a["completed_global_steps"] == 30
        b = ctx.run(perturb=True, max_steps=30, completed=30, grid_time=a["grid_time"])
        ctx.download(c2)
    one = wr.make_collapse_states(n)
    oracle_wr.run_gridwide(one, True, 60)
    assert b["completed_global_steps"] == 60
    assert np.array_equal(_bytes(c2), _bytes(one))
