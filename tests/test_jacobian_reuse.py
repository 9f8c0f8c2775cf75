"""Jacobian reuse (a study with no counterpart in the reference): the
oracle's lagged-J integrator.  Age 1 is the reference's integrator (the same
function, pinned bit for bit against the reference build in test_oracle.py);
lagged J can only be checked for what the mathematics fixes -- it leaves the
number of attempts nearly unchanged, removes J evaluations, and the run stays
within the reference's own comparator of the exact-J run -- so the lagged-J
arithmetic beyond that is "parity unpinned" against the reference and pinned
GPU-versus-oracle bit for bit (test_gpu_jacobian_reuse.py)."""
import os

import numpy as np
import pytest


def _run(oracle_wr, wr, n, age, perturb=True):
    oracle_wr._set_jacobian_reuse(age)
    try:
        cells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
        r, done = oracle_wr.run_gridwide(cells, perturb)
    finally:
        oracle_wr._set_jacobian_reuse(1)
    return r, done, cells


def _tot(cells, f):
    return int(cells["stats"][f].sum())


def test_age_one_is_the_reference_integrator(oracle_wr, wr):
    r0, d0, a = _run(oracle_wr, wr, 8, 1)
    cells = oracle_wr.make_cells(8, wr.CELL_DTYPE)
    r1, d1 = oracle_wr.run_gridwide(cells, True)
    assert (r0, d0) == (r1, d1)
    assert a.tobytes() == cells.tobytes()
    # the reference evaluates J once per accepted step (rejections reuse it)
    assert _tot(a, "jacobian_calls") == _tot(a, "accepted_steps")


@pytest.mark.parametrize("age", [2, 4, 8])
def test_lagged_jacobian_within_reference_comparator(oracle_wr, wr, tmp_path, age):
    n, g = 27, 3
    _, _, exact = _run(oracle_wr, wr, n, 1)
    r, done, lag = _run(oracle_wr, wr, n, age)
    assert r == 1 and done == 1000
    # J evaluations: at most one per `age` accepted steps per burn, plus refreshes
    # after rejections; attempts change by controller noise only
    jac, acc = _tot(lag, "jacobian_calls"), _tot(lag, "accepted_steps")
    assert jac < 0.8 * _tot(exact, "jacobian_calls")
    assert jac >= acc / age
    att, att0 = _tot(lag, "internal_steps"), _tot(exact, "internal_steps")
    assert abs(att - att0) < 0.02 * att0
    path = os.fspath(tmp_path / "exact.bin")
    wr.write_final_states(exact, g, path)
    s = wr.compare_final_states(lag, g, path)
    assert s["pass"], s
    assert s["max_thermodynamic_rel_error"] < 1e-4
