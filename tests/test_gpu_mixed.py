"""The FP32-factorisation form of the queue-fed kernel (FP32 LU of the
equilibrated I - gamma h J, FP64 residual and refinement sweeps, FP64
fallback) against the oracle's restatement of the same arithmetic: final
states, the seven counters per cell and the fallback counts identical bit for
bit.  Its distance from the FP64 path is reported, not gated here (the CPU
suite bounds it by the reference comparator's tolerances)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cells_equal(a, b):
    mask = np.ones(216, bool)
    mask[156:160] = False
    return np.array_equal(a.view(np.uint8).reshape(len(a), 216)[:, mask],
                          b.view(np.uint8).reshape(len(b), 216)[:, mask])


@pytest.mark.parametrize("recipe,grid,steps,sweeps", [
    ("reference", 3, 300, 2),
    ("reference", 3, 300, 1),
    ("lognormal128", 4, 60, 2),
])
def test_mixed_runs_bit_exact(wr, oracle_wr, recipe, grid, steps, sweeps):
    n = grid ** 3
    if recipe == "reference":
        cells = wr.make_collapse_states(n)
        ocells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
    else:
        seed, sigma, t0, t1 = wr.RECIPES[recipe]
        cells = wr.make_recipe_states(n, seed, sigma, t0, t1)
        ocells = oracle_wr.make_recipe_cells(n, wr.CELL_DTYPE, seed, sigma, t0, t1)
    fp64 = cells.copy()
    print(f"mixed run: {recipe} {grid}^3 x {steps} steps, {sweeps} sweeps", flush=True)
    with wr.Context(n) as ctx:
        ctx.upload(fp64)
        ctx.run(perturb=True, max_steps=steps)
        ctx.download(fp64)
        ctx.set_lu_precision(1, sweeps)
        ctx.upload(cells)
        info = ctx.run(perturb=True, max_steps=steps)
        ctx.download(cells)
        fb = ctx.mixed_fallbacks()
    try:
        oracle_wr.set_lu_precision(1, sweeps)
        orr, ocompleted = oracle_wr.run_gridwide(ocells, True, max_steps=steps)
        ofb = oracle_wr.mixed_fallbacks()
    finally:
        oracle_wr.set_lu_precision(0, 2)
    assert info["status"] == orr == 1
    assert info["completed_global_steps"] == ocompleted
    assert fb == ofb
    assert _cells_equal(cells, ocells)
    x0, x1 = fp64["current"]["xn"], cells["current"]["xn"]
    m = np.abs(x0) > 1e-20
    rel = np.abs(x1 - x0)[m] / np.abs(x0)[m]
    print(f"  fallbacks {fb}; vs FP64: species rel err max {rel.max():.3e} "
          f"median {np.median(rel):.3e}; accepted diff "
          f"{int(cells['stats']['accepted_steps'].sum()) - int(fp64['stats']['accepted_steps'].sum())}")


def test_mixed_needs_queue_kernel(wr):
    cells = wr.make_collapse_states(8)
    with wr.Context(8) as ctx:
        ctx.set_advance_kernel(0)
        ctx.set_lu_precision(1, 2)
        ctx.upload(cells)
        with pytest.raises(wr.WarpReactorError):
            ctx.run(perturb=True, max_steps=3)
    with pytest.raises(wr.WarpReactorError):
        with wr.Context(8) as ctx:
            ctx.set_lu_precision(1, 0)


def test_mixed_stage_solves_bit_exact(ctx, oracle_wr, wr):
    """FP32 factors and refined stage solves on E = fac I - J of real collapse
    states (several step sizes) and of random network states with a finite
    Jacobian: identical to the
    oracle's bit for bit, convergence flags included."""
    import sys
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    from states import random_states, trajectory_cases
    ys, facs = [], []
    gamma = 0.292893218813452
    for burn, dt in trajectory_cases(oracle_wr, wr.CELL_DTYPE, n_cells=4,
                                     steps=(0, 1, 5, 20, 60, 120, 300)):
        y = oracle_wr.burn_to_y(burn)
        for scale in (1.0, 1e-2, 1e-4, 1e-8):
            ys.append(y)
            facs.append(1.0 / (dt * scale * gamma))
    for y in random_states(64, oracle_wr, seed=41):
        if not np.all(np.isfinite(oracle_wr.jac(y))):
            continue  # extreme synthetic states where the network itself overflows
        ys.append(y)
        facs.append(10.0 ** np.random.default_rng(len(ys)).uniform(-14, 2))
    ys = np.array(ys)
    facs = np.array(facs)
    jac = np.array([oracle_wr.jac(y) for y in ys])
    rng = np.random.default_rng(42)
    b = rng.standard_normal(ys.shape) * (1e-4 + 1e-4 * np.abs(ys))
    for sweeps in (0, 1, 2):
        x, conv, lu = ctx.eval_solve_mixed(jac, facs, ys, b, sweeps)
        for i in range(len(ys)):
            olu, oip, oinfo = oracle_wr.mixed_lu(jac[i], facs[i], ys[i])
            assert np.array_equal(lu[i].view(np.uint32), olu.view(np.uint32)), i
            ox, oc = oracle_wr.solve_mixed(jac[i], facs[i], ys[i], b[i], sweeps)
            assert conv[i] == (-1 if oinfo != 0 else oc), i
            if oinfo == 0:
                assert np.array_equal(x[i].view(np.uint64), ox.view(np.uint64)), i
    print(f"{len(ys)} systems; converged with 2 sweeps: {np.mean(conv == 1):.2f}")
