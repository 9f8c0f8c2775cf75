"""Jacobian reuse on the GPU (wr_set_jacobian_reuse): the queue-fed kernel's
lagged-J instance against the oracle's restatement (built with the device
libm), bit for bit, and its results against the exact-J run through the
reference's comparator at 16^3."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu(wr, cells, age, max_steps=1000, threads=256, solo=64):
    with wr.Context(len(cells)) as ctx:
        ctx.set_queue_launch(threads, 0, solo)
        ctx.set_jacobian_reuse(age)
        ctx.upload(cells)
        info = ctx.run(perturb=True, max_steps=max_steps)
        ctx.download(cells)
    return info


@pytest.mark.parametrize("n,age", [(27, 3), (64, 8)])
def test_jacobian_reuse_bit_exact(wr, oracle_wr, n, age):
    cells = wr.make_collapse_states(n)
    info = _gpu(wr, cells, age)
    oracle_wr._set_jacobian_reuse(age)
    try:
        ocells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
        orr, odone = oracle_wr.run_gridwide(ocells, True)
    finally:
        oracle_wr._set_jacobian_reuse(1)
    assert info["status"] == orr == 1
    assert info["completed_global_steps"] == odone
    ocells["reserved_"] = 0
    cells["reserved_"] = 0
    assert cells.tobytes() == ocells.tobytes()
    assert int(cells["stats"]["jacobian_calls"].sum()) < int(cells["stats"]["accepted_steps"].sum())


def test_jacobian_reuse_16cube(wr, tmp_path):
    """4096 cells, 300 outer steps, J lagged up to 4 accepted steps: launch
    shapes agree bit for bit, and the grid passes the reference comparator
    against the exact-J run."""
    n, g = 16 ** 3, 16
    exact = wr.make_collapse_states(n)
    _gpu(wr, exact, 1, 300)
    a = wr.make_collapse_states(n)
    b = a.copy()
    ia = _gpu(wr, a, 4, 300)
    ib = _gpu(wr, b, 4, 300, threads=64, solo=0)
    assert ia == ib and a.tobytes() == b.tobytes()
    path = os.fspath(tmp_path / "exact.bin")
    wr.write_final_states(exact, g, path)
    s = wr.compare_final_states(a, g, path)
    assert s["pass"], s
    assert int(a["stats"]["jacobian_calls"].sum()) < 0.8 * int(exact["stats"]["jacobian_calls"].sum())


def test_jacobian_reuse_rejects_other_modes(wr):
    cells = wr.make_collapse_states(8)
    with wr.Context(8) as ctx:
        ctx.set_jacobian_reuse(4)
        ctx.upload(cells)
        ctx.set_lu_precision(1, 2)
        with pytest.raises(RuntimeError, match="Jacobian reuse"):
            ctx.run(perturb=True, max_steps=2)
        ctx.set_lu_precision(0)
        ctx.set_advance_kernel(0)
        with pytest.raises(RuntimeError, match="Jacobian reuse"):
            ctx.run(perturb=True, max_steps=2)
        ctx.set_advance_kernel(1)
        ctx.set_policy("legacy-local")
        with pytest.raises(RuntimeError, match="legacy-local"):
            ctx.run(perturb=True, max_steps=2)
    with pytest.raises(RuntimeError):
        with wr.Context(8) as ctx:
            ctx.set_jacobian_reuse(0)
