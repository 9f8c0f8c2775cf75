"""Parity of the HIP path (through the C ABI) with the CPU oracle.

The device evaluates exp/log/cbrt with wr_math.h (correctly rounded basic
operations only).  The oracle built with the same functions
(liboracle_wrmath.so) therefore has to be reproduced bit for bit: RHS, J,
LU, solves, whole ROS2S integrations (all seven counters) and whole collapse
runs.  The oracle built with the host libm is the reference's arithmetic
(bit-identical to the reference CPU build, tests/test_oracle.py); against it
the device differs only through libm, and the comparison uses the reference's
own comparator (reproducer.cpp:8416-8549): species rtol 1e-3 (1.1e-3 H,
3e-4 H2, 1e-4 for the others listed), atol 1e-4, T/e rtol 2e-4, rho 1e-4.
"""
import numpy as np
import pytest

from states import lu_test_matrices, random_matrices, random_states, trajectory_cases

pytestmark = pytest.mark.gpu


def bits(a):
    """Bit patterns with all NaNs mapped to one (IEEE leaves NaN payloads open)."""
    a = np.array(a, dtype=np.float64)
    a[np.isnan(a)] = np.nan
    return a.view(np.uint64)


def assert_bits_equal(got, want, what):
    g, w = bits(got), bits(want)
    bad = np.nonzero(g != w)
    assert len(bad[0]) == 0, f"{what}: {len(bad[0])} values differ, first at {tuple(b[0] for b in bad)}"


# ---- libm ---------------------------------------------------------------------------

def test_device_libm_matches_host_build(ctx, oracle_wr):
    rng = np.random.default_rng(21)
    xe = np.concatenate([rng.uniform(-745, 709, 20000), rng.uniform(-40, 40, 20000),
                         [0.0, -0.0, 1e-300, 709.78, 709.8, -745.1, -746.0, np.inf, -np.inf, np.nan]])
    xl = np.concatenate([10.0 ** rng.uniform(-320, 308, 20000), rng.uniform(0.5, 2.0, 20000),
                         [5e-324, 1e-310, 1.0, 2.0, 0.0, -1.0, np.inf, np.nan]])
    xc = np.concatenate([10.0 ** rng.uniform(-30, 30, 10000), -10.0 ** rng.uniform(-30, 30, 1000),
                         [0.0, -0.0, 1e-310, 8.0, -27.0, np.inf, np.nan]])
    for op, x in (("exp", xe), ("log", xl), ("cbrt", xc)):
        assert_bits_equal(ctx.eval_math(op, x), oracle_wr.libm(op, x), op)
    xs = 10.0 ** rng.uniform(-300, 300, 20000)
    assert_bits_equal(ctx.eval_math("sqrt", xs), np.sqrt(xs), "sqrt (IEEE correctly rounded)")


def test_ocml_versus_host_libm_report(ctx):
    """Not a gate on the product: how far the ROCm device libm is from the
    host libm, in ulps."""
    rng = np.random.default_rng(22)
    x = rng.uniform(-40, 40, 50000)
    d = np.abs(ctx.eval_math("ocml_exp", x).view(np.int64) - np.exp(x).view(np.int64))
    y = 10.0 ** rng.uniform(-30, 30, 50000)
    dl = np.abs(ctx.eval_math("ocml_log", y).view(np.int64) - np.log(y).view(np.int64))
    print(f"ocml exp: identical {np.mean(d == 0):.4f}, max ulp {d.max()}; "
          f"ocml log: identical {np.mean(dl == 0):.4f}, max ulp {dl.max()}")
    assert d.max() <= 4 and dl.max() <= 4


# ---- network, LU, integrator: bit-exact against the oracle ------------------------------

def test_device_rhs_bit_exact(ctx, oracle_wr):
    ys = random_states(2048, oracle_wr, seed=11, t_range=(0.0, 5.5))
    got = ctx.eval_rhs(ys)
    want = np.array([oracle_wr.rhs(y) for y in ys])
    assert_bits_equal(got, want, "rhs")


def test_device_jacobian_bit_exact(ctx, oracle_wr):
    ys = random_states(1024, oracle_wr, seed=12, t_range=(0.0, 5.5))
    got = ctx.eval_jac(ys)
    want = np.array([oracle_wr.jac(y) for y in ys])
    assert_bits_equal(got, want, "jacobian")


def test_device_lu_bit_exact(ctx, oracle):
    mats, rhs = lu_test_matrices()
    mats = np.concatenate([mats, random_matrices(512, seed=3)])
    lu, ip, info = ctx.eval_lu(mats)
    for k in range(len(mats)):
        a, p, i = oracle.lu(mats[k])
        assert i == info[k], k
        assert np.array_equal(p, ip[k]), k
        assert np.array_equal(bits(a), bits(lu[k])), k


def test_device_lu_solve_bit_exact(ctx, oracle):
    mats = random_matrices(256, seed=4) + np.eye(15) * 3.0
    lus, ips, xs = [], [], []
    rng = np.random.default_rng(5)
    for m in mats:
        a, p, i = oracle.lu(m)
        lus.append(a)
        ips.append(p)
        xs.append(rng.standard_normal(15))
    got = ctx.eval_lu_solve(np.array(lus), np.array(ips), np.array(xs))
    for k in range(len(mats)):
        want = oracle.lu_solve(lus[k], ips[k], xs[k])
        assert np.array_equal(bits(want), bits(got[k])), k


def test_lu_reference_cases_info(ctx):
    """The reference's LU cases (tests/gpu/test_warp_lu.mojo:19-47): identity,
    zero last row (info 15), zero first row after the forced swap (info 2),
    pivot tie broken by the first strict maximum, regular dominant cases."""
    mats, _ = lu_test_matrices()
    _, ip, info = ctx.eval_lu(mats)
    assert list(info) == [0, 15, 2, 0, 0, 0, 0, 0]
    assert ip[2][0] == 1 and ip[3][0] == 1
    assert list(ip[0]) == list(range(15))


def test_device_integrate_bit_exact(ctx, oracle_wr, wr):
    cases = trajectory_cases(oracle_wr, wr.CELL_DTYPE, n_cells=6)
    assert len(cases) >= 30
    ys = np.array([oracle_wr.burn_to_y(b) for b, _ in cases])
    dts = np.array([dt for _, dt in cases])
    got = ctx.eval_integrate(ys, dts, dts)
    for k in range(len(cases)):
        want = oracle_wr.integrate(ys[k], dts[k], dts[k])
        assert got["result"][k] == want["result"] == 1
        assert np.array_equal(got["counters"][k], want["counters"]), k
        assert_bits_equal(got["y"][k], want["y"], f"integrate case {k}")
        assert got["t_end"][k] == want["t_end"] and got["dt_end"][k] == want["dt_end"]


def test_device_integrate_fault_codes(ctx, oracle_wr, wr):
    """TOO_MANY_STEPS forced from the host (max_steps = 3) and BAD_INPUTS for
    tout <= t, with the counters the reference leaves."""
    cases = trajectory_cases(oracle_wr, wr.CELL_DTYPE, n_cells=1, steps=(0, 20))
    y = oracle_wr.burn_to_y(cases[0][0])
    dt = cases[0][1]
    got = ctx.eval_integrate(np.array([y, y]), np.array([dt, 0.0]), np.array([dt, dt]), max_steps=3)
    want = oracle_wr.integrate(y, dt, dt, max_steps=3)
    assert got["result"][0] == want["result"] == -4
    assert np.array_equal(got["counters"][0], want["counters"])
    assert got["result"][1] == -1


# ---- whole runs -----------------------------------------------------------------------

def _cells_equal(a, b):
    mask = np.ones(216, bool)
    mask[156:160] = False
    return np.array_equal(a.view(np.uint8).reshape(len(a), 216)[:, mask],
                          b.view(np.uint8).reshape(len(b), 216)[:, mask])


def test_grid1_known_answer_bit_exact(wr, oracle_wr):
    """--grid 1: 438 global steps (reference DESIGN.md:855-858); final state and
    counters identical to the oracle with the same libm."""
    cells = wr.make_collapse_states(1)
    r, completed = wr.run_cells(cells, perturb=True)
    assert r == 1 and completed == 438
    ocells = oracle_wr.make_cells(1, wr.CELL_DTYPE)
    orr, ocompleted = oracle_wr.run_gridwide(ocells, True)
    assert ocompleted == 438
    assert _cells_equal(cells, ocells)


@pytest.mark.parametrize("grid,perturb", [(2, True), (4, True), (3, False)])
def test_small_grid_bit_exact(wr, oracle_wr, grid, perturb):
    n = grid ** 3
    cells = wr.make_collapse_states(n)
    r, completed = wr.run_cells(cells, perturb=perturb)
    ocells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
    orr, ocompleted = oracle_wr.run_gridwide(ocells, perturb)
    assert r == orr == 1
    assert completed == ocompleted
    assert _cells_equal(cells, ocells)


def _run_context(wr, n, kernel, perturb=True, max_steps=1000):
    cells = wr.make_collapse_states(n)
    with wr.Context(n) as ctx:
        ctx.set_advance_kernel(kernel)
        ctx.upload(cells)
        info = ctx.run(perturb=perturb, max_steps=max_steps)
        ctx.download(cells)
        dts = ctx.dt_history()
    return cells, info, dts


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("grid,perturb", [(2, True), (3, False)])
def test_advance_kernels_bit_exact(wr, oracle_wr, kernel, grid, perturb):
    """Both advance kernels (one thread per cell; queue-fed persistent) give
    the oracle's final states and counters bit for bit."""
    n = grid ** 3
    cells, info, dts = _run_context(wr, n, kernel, perturb)
    ocells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
    orr, ocompleted = oracle_wr.run_gridwide(ocells, perturb)
    assert info["status"] == orr == 1
    assert info["completed_global_steps"] == ocompleted
    # the history also holds the dt of a final step that integrated no cell
    extra = 1 if info["stop_reason"] == 3 else 0
    assert len(dts) == ocompleted + extra
    assert _cells_equal(cells, ocells)


@pytest.mark.parametrize("kernel", [0, 1])
def test_advance_failure_path_bit_exact(wr, oracle_wr, kernel):
    """TOO_MANY_STEPS forced from the host (integrator limit 12 instead of
    1e7): on a 3^3 grid one cell fails at outer step 232.  Every cell runs in
    the failing step (27 cells start before any fails), so the grid matches
    the oracle bit for bit, including the failed cell's state: density-scaled
    composition, e from eos_rt, counters added (reproducer.cpp:7760-7784)."""
    n = 27
    cells = wr.make_collapse_states(n)
    with wr.Context(n) as ctx:
        ctx.set_advance_kernel(kernel)
        ctx.set_integrator_max_steps(12)
        ctx.upload(cells)
        info = ctx.run(perturb=True)
        ctx.download(cells)
    oracle_wr._set_integrator_max_steps(12)
    try:
        ocells = oracle_wr.make_cells(n, wr.CELL_DTYPE)
        orr, ocompleted = oracle_wr.run_gridwide(ocells, True)
    finally:
        oracle_wr._set_integrator_max_steps(0)
    assert orr == info["status"] == -4
    assert info["completed_global_steps"] == ocompleted == 232
    assert info["stop_reason"] == 4 and info["failed_phase"] == 2
    failed = np.nonzero(ocells["completed_steps"] == ocompleted)[0]
    assert list(failed) == [info["failed_cell"]]
    assert _cells_equal(cells, ocells)


def test_advance_kernels_agree_16cube(wr):
    """16^3 cells, 300 outer steps: the queue-fed kernel (lanes refilled from
    an atomic queue, cells finishing in any order) and the thread-per-cell
    kernel produce identical grids, dt histories and run info."""
    n = 16 ** 3
    a, ia, da = _run_context(wr, n, 0, max_steps=300)
    b, ib, db = _run_context(wr, n, 1, max_steps=300)
    assert ia == ib
    assert np.array_equal(da, db)
    assert _cells_equal(a, b)


def test_grid_against_reference_comparator(wr, oracle, tmp_path):
    """Against the reference arithmetic (host libm), a 4^3 run passes the
    reference's own final-state comparator."""
    n, g = 64, 4
    ocells = oracle.make_cells(n, wr.CELL_DTYPE)
    oracle.run_gridwide(ocells, True)
    path = str(tmp_path / "final_states_grid4_cpu.bin")
    wr.write_final_states(ocells, g, path)
    cells = wr.make_collapse_states(n)
    r, completed = wr.run_cells(cells, perturb=True)
    assert r == 1
    summary = wr.compare_final_states(cells, g, path)
    print("comparator vs host-libm oracle:", summary)
    assert summary["pass"], summary["first_failure"]


@pytest.mark.parametrize("recipe,grid,steps", [("uniform16", 8, 145), ("lognormal128", 4, 60)])
def test_recipe_runs_bit_exact(wr, oracle_wr, recipe, grid, steps):
    """Benchmark-recipe initial states (log-normal density, fixed or
    log-uniform T) through the default advance kernel: final states and
    counters identical to the oracle."""
    n = grid ** 3
    seed, sigma, t0, t1 = wr.RECIPES[recipe]
    cells = wr.make_recipe_states(n, seed, sigma, t0, t1)
    with wr.Context(n) as ctx:
        ctx.upload(cells)
        info = ctx.run(perturb=True, max_steps=steps)
        ctx.download(cells)
    ocells = oracle_wr.make_recipe_cells(n, wr.CELL_DTYPE, seed, sigma, t0, t1)
    orr, ocompleted = oracle_wr.run_gridwide(ocells, True, max_steps=steps)
    assert info["status"] == orr == 1
    assert info["completed_global_steps"] == ocompleted
    assert _cells_equal(cells, ocells)


def test_graph_replays_match_eager_launches(wr):
    """Outer steps enqueued as a captured HIP graph (8 steps per replay, tens
    of replays per run) against individual launches: identical grids, dt
    histories and run info, for two runs on one context (the graph is
    reused), then for a smaller grid on the same context (captured again),
    with eager work in between."""
    n = 12 ** 3
    init = wr.make_collapse_states(n)
    results = {}
    for graph in (False, True):
        with wr.Context(n) as ctx:
            ctx.set_graph(graph)
            runs = []
            for m, steps in ((n, 200), (n, 200), (n // 2, 77)):
                cells = init[:m].copy()
                ctx.upload(cells)
                info = ctx.run(perturb=True, max_steps=steps)
                ctx.download(cells)
                runs.append((cells, info, ctx.dt_history()))
                ctx.eval_rhs(np.full((4, 15), 1e-3))  # eager work between replays
            results[graph] = runs
    for (a, ia, da), (b, ib, db) in zip(results[False], results[True]):
        assert ia == ib
        assert np.array_equal(da, db)
        assert _cells_equal(a, b)
    # the repeated run on one context reproduces the first
    assert _cells_equal(results[True][0][0], results[True][1][0])
