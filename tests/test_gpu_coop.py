"""Wave-cooperative advance (wr_set_coop): the most expensive cells of each
outer step advanced by whole 64-lane waves -- the generated cooperative
network programs, the wave-level LU and solves -- must give the final states
and all seven ROS2S counters of the single-lane path and of the oracle, bit
for bit: with every cell cooperative, with a mix, on the failure path, and on
the full 64^3 run."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASK = np.ones(216, bool)
MASK[156:160] = False


def _bytes(c):
    return c.view(np.uint8).reshape(len(c), 216)[:, MASK]


def _run(wr, init, steps, coop, waves=None, max_integ=0):
    n = len(init)
    with wr.Context(n) as ctx:
        ctx.set_coop(coop, waves)
        if max_integ:
            ctx.set_integrator_max_steps(max_integ)
        ctx.upload(init)
        t0 = time.perf_counter()
        info = ctx.run(perturb=True, max_steps=steps)
        el = time.perf_counter() - t0
        cells = np.zeros(n, dtype=wr.CELL_DTYPE)
        ctx.download(cells)
        dts = ctx.dt_history()
    return info, cells, dts, el


@pytest.mark.parametrize("n,coop,waves", [(8, 8, 8), (27, 27, 4), (64, 5, 3)])
def test_coop_cells_bit_exact_against_oracle(wr, oracle_wr, n, coop, waves):
    print(f"coop case: {n} cells, coop {coop}, waves {waves}", flush=True)
    init = wr.make_collapse_states(n)
    info, cells, dts, _ = _run(wr, init, 60, coop, waves)
    assert info["status"] == 1 and info["completed_global_steps"] == 60
    ref = init.copy()
    status = oracle_wr.replay_cells(ref, np.arange(n), dts, perturb=True)
    assert np.all(status == 1)
    assert np.array_equal(_bytes(cells), _bytes(ref))


def test_coop_recipe_cells_bit_exact(wr, oracle_wr):
    """log-normal recipe cells (T 100-8000 K): every temperature branch of the
    rate fits through the cooperative programs."""
    n = 64
    init = wr.make_recipe_states(n, 1, 1.0, 100.0, 8000.0)
    info, cells, dts, _ = _run(wr, init, 20, 64, 16)
    assert info["status"] == 1
    ref = init.copy()
    assert np.all(oracle_wr.replay_cells(ref, np.arange(n), dts, perturb=True) == 1)
    assert np.array_equal(_bytes(cells), _bytes(ref))


def test_coop_failure_path(wr):
    """Integrator step limit 12: the status, the step of the failure, the
    failing cell and its state (density-scaled composition, e from eos_rt,
    counters) are those of the single-lane path.  Which other cells still
    advance in the failing step depends on timing -- a lane stops taking cells
    once a failure is recorded, as each thread of the reference's kernel
    returns early once the failure flag is set -- so the cells are compared
    where both runs advanced them the same number of steps."""
    n = 27
    init = wr.make_collapse_states(n)
    a = _run(wr, init, 1000, 27, 27, max_integ=12)
    b = _run(wr, init, 1000, 0, 0, max_integ=12)
    assert a[0]["status"] == b[0]["status"] != 1
    assert a[0]["failed_step"] == b[0]["failed_step"] and a[0]["failed_cell"] == b[0]["failed_cell"]
    f = a[0]["failed_cell"]
    assert np.array_equal(_bytes(a[1][f:f + 1]), _bytes(b[1][f:f + 1]))
    same = a[1]["completed_steps"] == b[1]["completed_steps"]
    assert same.sum() >= 1
    assert np.array_equal(_bytes(a[1][same]), _bytes(b[1][same]))


def test_coop_64cube_bit_identical_to_single_lane(wr):
    """The full 64^3 x 1000-step run: final states, dt history and step count
    with cooperative waves equal those of the single-lane path."""
    n = 64 ** 3
    init = wr.make_collapse_states(n)
    base = _run(wr, init, 1000, 0, 0)
    for coop, waves in ((4, 4), (16, 16)):
        got = _run(wr, init, 1000, coop, waves)
        print(f"64^3 x 1000: single-lane {base[3]:.3f} s, coop {coop}/{waves} {got[3]:.3f} s "
              f"({base[3] / got[3]:.2f}x)", flush=True)
        assert got[0]["completed_global_steps"] == base[0]["completed_global_steps"]
        assert np.array_equal(got[2], base[2])
        assert np.array_equal(_bytes(got[1]), _bytes(base[1]))
