"""The gridwide-cta32 driver policy on the GPU (wr_set_policy(ctx, 2)): the
Mojo front end's policy (reproducer.mojo:2829-3093) -- tiles of 32 cells with
one ROS2S step controller each, a tile being one half of a wave -- checked bit
for bit against the oracle's restatement built with the device libm (the
restatement's own pinning is in tests/test_cta.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _equal(a, b):
    """Byte equality of two cell arrays outside the 4 padding bytes."""
    mask = np.ones(216, bool)
    mask[156:160] = False
    return np.array_equal(a.view(np.uint8).reshape(len(a), 216)[:, mask],
                          b.view(np.uint8).reshape(len(b), 216)[:, mask])


def _run_gpu(wr, cells, perturb, max_steps=1000, graph=True, integ_max_steps=None):
    with wr.Context(len(cells)) as ctx:
        ctx.set_policy("gridwide-cta32")
        ctx.set_graph(graph)
        if integ_max_steps:
            ctx.set_integrator_max_steps(integ_max_steps)
        ctx.upload(cells)
        info = ctx.run(perturb=perturb, max_steps=max_steps)
        ctx.download(cells)
        dts = ctx.dt_history()
    return info, dts


@pytest.mark.parametrize("n,steps,perturb", [(27, 120, True), (100, 80, True), (64, 60, False)])
def test_cta32_matches_oracle(wr, oracle_wr, n, steps, perturb):
    print(f"cta32 reference cells n={n} steps={steps} perturb={perturb}", flush=True)
    cells = wr.make_collapse_states(n)
    ocells = cells.copy()
    info, dts = _run_gpu(wr, cells, perturb, steps)
    r, done, lim = oracle_wr.run_cta(ocells, perturb, steps, tile=32)
    assert info["status"] == r == 1
    assert info["completed_global_steps"] == done == steps
    assert len(dts) == steps
    assert _equal(cells, ocells)
    st = cells["stats"]
    for t in range(n // 32):
        assert len(np.unique(st["accepted_steps"][32 * t:32 * t + 32])) == 1


def test_cta32_lognormal_cells_match_oracle(wr, oracle_wr):
    """Heterogeneous cells (lognormal128 recipe): each tile's controller is
    driven by its hardest cell."""
    n, steps = 96, 40
    print(f"cta32 lognormal n={n} steps={steps}", flush=True)
    cells = wr.make_recipe_states(n, *wr.RECIPES["lognormal128"])
    ocells = cells.copy()
    info, _ = _run_gpu(wr, cells, True, steps)
    r, done, _ = oracle_wr.run_cta(ocells, True, steps, tile=32)
    assert info["status"] == r == 1 and info["completed_global_steps"] == done
    assert _equal(cells, ocells)


def test_cta32_grid16_matches_oracle_and_eager_launches(wr, oracle_wr):
    """16^3 cells, 100 outer steps (perturbations at 20..100): HIP-graph
    replays, eager launches and the oracle agree bit for bit."""
    n, steps = 16 ** 3, 100
    print(f"cta32 grid16 n={n} steps={steps}", flush=True)
    a = wr.make_collapse_states(n)
    b, o = a.copy(), a.copy()
    ia, da = _run_gpu(wr, a, True, steps, graph=True)
    ib, db = _run_gpu(wr, b, True, steps, graph=False)
    r, done, _ = oracle_wr.run_cta(o, True, steps, tile=32)
    assert ia["status"] == ib["status"] == r == 1
    assert ia["completed_global_steps"] == ib["completed_global_steps"] == done == steps
    assert np.array_equal(da, db)
    assert _equal(a, b)
    assert _equal(a, o)


def test_cta32_failure_reports_the_lowest_tile(wr, oracle_wr):
    """TOO_MANY_STEPS forced from the host (integrator limit 12) on 100 cells:
    the run stops in the same step with the same code as the oracle, and the
    failed cell is the first cell of the lowest failing tile."""
    n = 100
    print(f"cta32 failure n={n}", flush=True)
    cells = wr.make_collapse_states(n)
    info, _ = _run_gpu(wr, cells, True, integ_max_steps=12)
    oracle_wr._set_integrator_max_steps(12)
    try:
        ocells = wr.make_collapse_states(n)
        r, done, _ = oracle_wr.run_cta(ocells, True, 1000, tile=32)
    finally:
        oracle_wr._set_integrator_max_steps(0)
    assert info["status"] == r == -4
    assert info["completed_global_steps"] == done
    assert info["stop_reason"] == 4 and info["failed_phase"] == 2
    assert info["failed_cell"] % 32 == 0 and 0 <= info["failed_cell"] < n


def test_cta32_rejects_fp32_factorisation(wr):
    cells = wr.make_collapse_states(8)
    with wr.Context(8) as ctx:
        ctx.set_policy("gridwide-cta32")
        ctx.set_lu_precision(1, 2)
        ctx.upload(cells)
        with pytest.raises(RuntimeError, match="gridwide-cta32"):
            ctx.run(perturb=True)
