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


def test_legacy_driver_golden_file(oracle, wr):
    """The committed fixture final_states_grid2_legacy_local.bin was written by
    the reference's writer from its legacy-local loop on 8 cells
    (tests/golden/make_golden.py); the oracle's restatement packs to the same
    bytes."""
    import json
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    cells = oracle.make_cells(8, wr.CELL_DTYPE)
    r, done, failed = oracle.run_legacy(cells, True)
    meta = json.load(open(os.path.join(gold, "MANIFEST.json")))
    assert r == 1 and failed == -1 and done == meta["grid2_legacy_local_steps"]
    data = open(os.path.join(gold, "final_states_grid2_legacy_local.bin"), "rb").read()
    assert data == wr.pack_final_states(cells, 2).tobytes()
