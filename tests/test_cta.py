"""The gridwide-cta32 driver policy of the reference's Mojo front end
(reproducer.mojo:2829-3093, driver 3604-3621): the oracle's restatement.

There is no Mojo toolchain here, so the restatement cannot be run against the
reference.  It is pinned instead by what the policy's definition fixes:
  * with one cell per tile and the density updates in reproducer.cpp's order
    it is the gridwide-independent run (itself pinned against the reference's
    compiled loop), bit for bit in every state and every counter except the
    Jacobian count (one per trial instead of one per accepted step);
  * a tile of identical cells takes the path of any one of them alone (the
    tile error is the largest cell error);
  * cells of one tile share their step counts, end at the same time, conserve
    charge, and stay within the reference comparator's tolerances of the
    independent run (the reference design's acceptance rule for the policy).
What none of these reaches -- the operation order of the two density-driver
updates, (dt * driver) / tff at reproducer.mojo:2858 and 2891 -- follows the
Mojo source by reading."""
import numpy as np
import pytest

D_SPECIES = (4, 5, 7, 9, 10)  # deuterium-bearing, not gated by the comparator


def _bytes(a):
    mask = np.ones(216, bool)
    mask[156:160] = False
    return a.view(np.uint8).reshape(len(a), 216)[:, mask]


def _state_bytes(a):
    """The cell bytes before the counters (state, time, driver, completed)."""
    return a.view(np.uint8).reshape(len(a), 216)[:, :156]


@pytest.mark.parametrize("n,steps", [(27, 120), (8, 438)])
def test_one_cell_tiles_are_the_gridwide_run(oracle_wr, wr, n, steps):
    a = oracle_wr.make_cells(n, wr.CELL_DTYPE)
    b = a.copy()
    ra = oracle_wr.run_gridwide(a, True, steps)
    rb = oracle_wr.run_cta(b, True, steps, tile=1, cpp_order=True)
    assert ra[0] == rb[0] == 1 and ra[1] == rb[1]
    assert np.array_equal(_state_bytes(a), _state_bytes(b))
    for f in a["stats"].dtype.names:
        if f != "jacobian_calls":
            assert np.array_equal(a["stats"][f], b["stats"][f]), f
    # one Jacobian per trial: every completed trial plus the singular ones
    assert np.all(b["stats"]["jacobian_calls"] >= b["stats"]["internal_steps"])
    assert np.all(a["stats"]["jacobian_calls"] == a["stats"]["accepted_steps"])


def test_operation_order_of_the_density_update_matters(oracle_wr, wr):
    """The two orders of the density-driver update give different bits, so
    the restatement's choice is visible in the results."""
    a = oracle_wr.make_cells(8, wr.CELL_DTYPE)
    b = a.copy()
    oracle_wr.run_cta(a, True, 60, tile=1, cpp_order=True)
    oracle_wr.run_cta(b, True, 60, tile=1, cpp_order=False)
    assert not np.array_equal(a["density_driver"], b["density_driver"])
    np.testing.assert_allclose(a["density_driver"], b["density_driver"], rtol=1e-12)


def test_identical_cells_share_the_single_cell_path(oracle_wr, wr):
    """Without perturbations every cell starts from the reference state: a
    32-cell tile then takes exactly the path of one cell alone."""
    n = 40  # a full tile and a partial one
    a = oracle_wr.make_cells(n, wr.CELL_DTYPE)
    b = a.copy()
    ra = oracle_wr.run_cta(a, False, 200, tile=32)
    rb = oracle_wr.run_cta(b, False, 200, tile=1)
    assert ra[:2] == rb[:2] and ra[0] == 1
    assert np.array_equal(_bytes(a), _bytes(b))


def test_tiles_share_step_counts_and_stay_within_comparator(oracle_wr, wr):
    n, steps = 96, 160
    cells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
    ind = cells.copy()
    r, done, lim = oracle_wr.run_cta(cells, True, steps, tile=32)
    ri, donei = oracle_wr.run_gridwide(ind, True, steps)
    assert r == ri == 1 and done == steps and 0 <= lim < n
    st = cells["stats"]
    for t in range(n // 32):
        sl = slice(32 * t, 32 * t + 32)
        for f in ("internal_steps", "accepted_steps", "rejected_steps", "jacobian_calls"):
            assert len(np.unique(st[f][sl])) == 1, (t, f)
    # tiles choose their own substeps
    assert len(np.unique(st["accepted_steps"][::32])) > 1
    # a tile takes at least as many accepted steps as its hardest cell alone
    for t in range(n // 32):
        sl = slice(32 * t, 32 * t + 32)
        assert st["accepted_steps"][32 * t] >= ind["stats"]["accepted_steps"][sl].max() * 0.9
    assert np.all(cells["completed_steps"] == steps)
    assert len(np.unique(cells["time"])) == 1
    # charge balance after the post-step sequence (balance_charge)
    x = cells["current"]["xn"]
    ions = -x[:, 3] - x[:, 7] + x[:, 1] + x[:, 12] + x[:, 6] + x[:, 4] + x[:, 9] + 2.0 * x[:, 11]
    np.testing.assert_allclose(x[:, 0], ions, rtol=1e-12)
    # within the reference comparator's tolerances of the independent run
    # (species rtol of reproducer.cpp:7429-7431, atol 1e-4; T, e 2e-4)
    rtol = np.array([1e-3, 1e-3, 1.1e-3, 1e-3, 1e-4, 10, 1e-3, 1e-4, 3e-4, 1e-4, 10, 1e-4,
                     1e-4, 1e-4])
    xi = ind["current"]["xn"]
    for s in range(14):
        if s in D_SPECIES:
            continue
        bound = 1e-4 + rtol[s] * np.maximum(np.abs(x[:, s]), np.abs(xi[:, s]))
        assert np.all(np.abs(x[:, s] - xi[:, s]) <= bound), s
    T, Ti = cells["current"]["T"], ind["current"]["T"]
    assert np.all(np.abs(T - Ti) <= 2e-4 * np.maximum(T, Ti))
    assert not np.array_equal(_bytes(cells), _bytes(ind))


def test_tile_failure_reports_the_integrator_code(oracle_wr, wr):
    """TOO_MANY_STEPS forced from the host (integrator limit 12).  The tile
    loop tests the limit only while x < dt_grid (reproducer.mojo:2945-2947),
    the per-cell integrator also once after its last accepted step
    (reproducer.cpp:7236-7247), so a burn of exactly 13 steps fails only
    there: with one cell per tile the Mojo loop fails later, never earlier."""
    oracle_wr._set_integrator_max_steps(12)
    try:
        cells = oracle_wr.make_cells(27, wr.CELL_DTYPE)
        r, done, _ = oracle_wr.run_cta(cells, True, 1000, tile=32)
        ind = oracle_wr.make_cells(27, wr.CELL_DTYPE)
        ri, donei = oracle_wr.run_cta(ind, True, 1000, tile=1, cpp_order=True)[:2]
        g = oracle_wr.make_cells(27, wr.CELL_DTYPE)
        rg, doneg = oracle_wr.run_gridwide(g, True, 1000)
    finally:
        oracle_wr._set_integrator_max_steps(0)
    assert r == ri == rg == -4 and 0 < done < 1000
    assert doneg < donei < 1000
