"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances.  Integer/index work (LU pivots, info codes, step counters,
completed steps) must be identical.  FP64 values: the LU and the triangular
solves involve no libm calls and must be bit-identical.  The network RHS/J
use the device libm for exp/log (ROCm ocml), which may differ from glibc by
an ulp, so single evaluations are held to a relative 1e-11 of the output's
magnitude scale; integrated states to the north-star bar of 1e-9 relative per
species above a 1e-20 abundance floor.
"""
import numpy as np
import pytest

from states import lu_test_matrices, random_matrices, random_states, trajectory_cases

pytestmark = pytest.mark.gpu

REL_EVAL = 1e-11
REL_STATE = 1e-9
FLOOR = 1e-20


def _scaled_err(gpu, cpu, scale):
    return np.abs(gpu - cpu) / np.maximum(scale, 1e-300)


def bits(a):
    """Bit patterns with all NaNs mapped to one (IEEE leaves NaN payloads open)."""
    a = np.array(a, dtype=np.float64)
    a[np.isnan(a)] = np.nan
    return a.view(np.uint64)


def test_device_rhs_matches_oracle(ctx, oracle):
    ys = random_states(2048, oracle, seed=11)
    got = ctx.eval_rhs(ys)
    want = np.array([oracle.rhs(y) for y in ys])
    same = np.mean(np.all(bits(got) == bits(want), axis=1))
    # scale: magnitude of the largest contribution is not visible; use |f| and
    # the row's largest |f_i| * 1e-6 as a cancellation allowance
    scale = np.maximum(np.abs(want), 1e-6 * np.abs(want).max(axis=1, keepdims=True))
    err = _scaled_err(got, want, scale)
    print(f"rhs: bitwise-identical rows {same:.4f}, max scaled err {err.max():.3e}")
    assert np.all(np.isfinite(got) == np.isfinite(want))
    fin = np.isfinite(want)
    assert err[fin].max() <= REL_EVAL


def test_device_jacobian_matches_oracle(ctx, oracle):
    ys = random_states(1024, oracle, seed=12)
    got = ctx.eval_jac(ys)
    want = np.array([oracle.jac(y) for y in ys])
    same = np.mean(np.all(got.reshape(len(ys), -1).view(np.uint64) ==
                          want.reshape(len(ys), -1).view(np.uint64), axis=1))
    scale = np.maximum(np.abs(want), 1e-6 * np.abs(want).max(axis=(1, 2), keepdims=True))
    err = _scaled_err(got, want, scale)
    fin = np.isfinite(want)
    print(f"jac: bitwise-identical matrices {same:.4f}, max scaled err {err[fin].max():.3e}")
    assert np.all(np.isfinite(got) == fin)
    assert err[fin].max() <= REL_EVAL


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
    """Known answers of the reference's LU cases: identity and forced swaps
    are regular, the zero last row is singular at column 15."""
    mats, _ = lu_test_matrices()
    _, ip, info = ctx.eval_lu(mats)
    assert list(info) == [0, 15, 2, 0, 0, 0, 0, 0]
    assert ip[2][0] == 1 and ip[3][0] == 1  # |-3| is the first strict maximum
    assert list(ip[0]) == list(range(15))


def test_device_integrate_matches_oracle(ctx, oracle, wr):
    cases = trajectory_cases(oracle, wr.CELL_DTYPE, n_cells=6)
    ys = np.array([oracle.burn_to_y(b) for b, _ in cases])
    dts = np.array([dt for _, dt in cases])
    got = ctx.eval_integrate(ys, dts, dts)
    n_same = 0
    worst = 0.0
    for k in range(len(cases)):
        want = oracle.integrate(ys[k], dts[k], dts[k])
        assert got["result"][k] == want["result"] == 1
        counters_same = np.array_equal(got["counters"][k], want["counters"])
        n_same += counters_same
        mask = np.abs(want["y"]) > FLOOR
        rel = np.abs(got["y"][k] - want["y"])[mask] / np.abs(want["y"][mask])
        worst = max(worst, rel.max())
    print(f"integrate: {len(cases)} cases, identical counters {n_same}, worst rel {worst:.3e}")
    assert n_same == len(cases)
    assert worst <= REL_STATE


def test_grid1_known_answer(wr, oracle):
    """--grid 1: 438 global steps (DESIGN.md:855-858 of the reference) and the
    oracle's final state and counters."""
    cells = wr.make_collapse_states(1)
    r, completed = wr.run_cells(cells, perturb=True)
    assert r == 1
    assert completed == 438
    ocells = oracle.make_cells(1, wr.CELL_DTYPE)
    orr, ocompleted = oracle.run_gridwide(ocells, True)
    assert ocompleted == 438
    g, o = cells[0], ocells[0]
    assert g["completed_steps"] == o["completed_steps"] == 438
    for f in wr.STATS_FIELDS:
        assert g["stats"][f] == o["stats"][f], f
    mask = o["current"]["xn"] > FLOOR
    rel = np.abs(g["current"]["xn"] - o["current"]["xn"])[mask] / o["current"]["xn"][mask]
    print(f"grid1: T gpu {g['current']['T']!r} oracle {o['current']['T']!r}; max species rel {rel.max():.3e}")
    assert rel.max() <= REL_STATE
    assert abs(g["current"]["T"] - o["current"]["T"]) <= REL_STATE * o["current"]["T"]


@pytest.mark.parametrize("grid,perturb", [(2, True), (4, True), (3, False)])
def test_small_grid_matches_oracle(wr, oracle, grid, perturb):
    n = grid ** 3
    cells = wr.make_collapse_states(n)
    r, completed = wr.run_cells(cells, perturb=perturb)
    ocells = oracle.make_cells(n, wr.CELL_DTYPE)
    orr, ocompleted = oracle.run_gridwide(ocells, perturb)
    assert r == orr == 1
    assert completed == ocompleted
    assert np.array_equal(cells["completed_steps"], ocells["completed_steps"])
    stats_same = np.mean([all(cells[c]["stats"][f] == ocells[c]["stats"][f] for f in wr.STATS_FIELDS)
                          for c in range(n)])
    xg, xo = cells["current"]["xn"], ocells["current"]["xn"]
    mask = xo > FLOOR
    rel = np.abs(xg - xo)[mask] / xo[mask]
    print(f"grid {grid}: identical counters in {stats_same:.3f} of cells, max rel {rel.max():.3e}")
    assert stats_same == 1.0
    assert rel.max() <= REL_STATE
