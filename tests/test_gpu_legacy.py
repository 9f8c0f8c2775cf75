"""The legacy-local driver policy on the GPU (wr_set_policy): the reference's
CPU driver (reproducer.cpp:8628-8648, advance_collapse_step 7646-7685), in
which every cell steps with its own free-fall dt, checked bit for bit against
the oracle's restatement of that loop built with the device libm (the
restatement itself is checked against the reference's compiled loop in
tests/test_legacy.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _equal(a, b):
    """Byte equality of two cell arrays outside the 4 padding bytes."""
    mask = np.ones(216, bool)
    mask[156:160] = False
    return np.array_equal(a.view(np.uint8).reshape(len(a), 216)[:, mask],
                          b.view(np.uint8).reshape(len(b), 216)[:, mask])


def _run_gpu(wr, cells, perturb, **launch):
    with wr.Context(len(cells)) as ctx:
        ctx.set_policy("legacy-local")
        if launch:
            ctx.set_queue_launch(launch["threads"], 0, launch.get("solo", 0))
        if "integ_max_steps" in launch:
            ctx.set_integrator_max_steps(launch["integ_max_steps"])
        ctx.upload(cells)
        info = ctx.run(perturb=perturb)
        ctx.download(cells)
        assert len(ctx.dt_history()) == 0
    return info


def test_legacy_local_grid1_known_answer(wr, oracle_wr):
    """One cell: 438 completed steps (the reference's grid-1 answer; for one
    cell the two policies coincide), equal to the oracle bit for bit."""
    cells = wr.make_collapse_states(1)
    info = _run_gpu(wr, cells, True)
    ocells = oracle_wr.make_cells(1, wr.CELL_DTYPE)
    orr, odone, ofailed = oracle_wr.run_legacy(ocells, True)
    assert info["status"] == orr == 1 and ofailed == -1
    assert info["completed_global_steps"] == odone == 438
    assert info["stop_reason"] == 3  # every cell stopped (no cell integrated)
    assert _equal(cells, ocells)


@pytest.mark.parametrize("n,perturb", [(64, True), (64, False), (125, True)])
def test_legacy_local_matches_oracle(wr, oracle_wr, n, perturb):
    cells = wr.make_collapse_states(n)
    info = _run_gpu(wr, cells, perturb)
    ocells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
    orr, odone, ofailed = oracle_wr.run_legacy(ocells, perturb)
    assert info["status"] == orr == 1 and ofailed == -1 and info["failed_cell"] == -1
    assert info["completed_global_steps"] == odone
    assert _equal(cells, ocells)
    # each cell keeps its own time: the perturbed cells end at different times
    t = cells["time"]
    assert np.all(t > 0)
    if perturb:
        assert len(np.unique(t)) > 1


def test_legacy_local_failure_restores_later_cells(wr, oracle_wr):
    """TOO_MANY_STEPS forced from the host (integrator limit 12): on 27 cells
    cell 3 fails in step 227.  All 27 advance in that step on the GPU; the
    cells after cell 3 are put back to their state of the step's start, as the
    reference's sequential loop never reached them, and the grid matches the
    oracle bit for bit (the failed cell with its scaled composition, e from
    eos_rt and counters added)."""
    n = 27
    cells = wr.make_collapse_states(n)
    info = _run_gpu(wr, cells, True, threads=256, integ_max_steps=12)
    oracle_wr._set_integrator_max_steps(12)
    try:
        ocells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
        orr, odone, ofailed = oracle_wr.run_legacy(ocells, True)
    finally:
        oracle_wr._set_integrator_max_steps(0)
    assert info["status"] == orr == -4
    assert info["failed_cell"] == ofailed == 3
    assert info["completed_global_steps"] == odone == 227
    assert info["stop_reason"] == 4 and info["failed_phase"] == 2
    assert _equal(cells, ocells)
    # cells after the failing one did not take step 227
    assert np.all(cells["completed_steps"][4:] <= 227)


def test_legacy_local_launch_shape_independent(wr):
    """16^3 cells: the results do not depend on the queue kernel's launch
    shape (which lane takes which cell)."""
    n = 16 ** 3
    a = wr.make_collapse_states(n)
    b = a.copy()
    ia = _run_gpu(wr, a, True, threads=256, solo=64)
    ib = _run_gpu(wr, b, True, threads=64, solo=0)
    assert ia == ib
    assert _equal(a, b)


def test_legacy_local_rejects_fp32_factorisation(wr):
    cells = wr.make_collapse_states(8)
    with wr.Context(8) as ctx:
        ctx.set_policy("legacy-local")
        ctx.set_lu_precision(1, 2)
        ctx.upload(cells)
        with pytest.raises(RuntimeError, match="legacy-local"):
            ctx.run(perturb=True)
