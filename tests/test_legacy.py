"""The legacy-local driver policy (the reference CPU build's, reproducer.cpp:
8630-8648 with advance_collapse_step 7646-7685): the oracle's restatement
against the reference's own compiled loop, bit for bit."""
import numpy as np
import pytest


def _equal(a, b):
    mask = np.ones(216, bool)
    mask[156:160] = False
    return np.array_equal(a.view(np.uint8).reshape(len(a), 216)[:, mask],
                          b.view(np.uint8).reshape(len(b), 216)[:, mask])


@pytest.mark.parametrize("n,perturb", [(1, True), (8, True), (8, False)])
def test_legacy_driver_matches_reference(oracle, ref, wr, n, perturb):
    cells = oracle.make_cells(n, wr.CELL_DTYPE)
    rcells = cells.copy()
    r, done, failed = oracle.run_legacy(cells, perturb)
    rr, rdone = ref.run_legacy(rcells, perturb)
    assert r == rr == 1 and done == rdone and failed == -1
    assert _equal(cells, rcells)
    if n == 1:
        assert done == 438  # the grid-1 known answer (legacy-local = gridwide for one cell)
