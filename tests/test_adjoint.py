"""The tangent-linear and adjoint restatement (oracle/adjoint.c, with the
derivative code generated by warp-reactor_amd/gen/ad.py), pinned by what the
mathematics fixes rather than by its own formulas:

  * the generated forward code with every rate multiplier 1 is the network
    the reference-pinned oracle evaluates, bit for bit, and the record run
    equals the plain gridwide run;
  * replaying the recorded steps reproduces the recorded run bit for bit;
  * tangents agree with central finite differences (delta = 1e-6, 8 seeded
    directions over initial states and log-rate multipliers);
  * adjoints agree with tangents through <w, T v> = <T^T w, v> to rounding;
  * the transposed LU solve agrees with numpy.

Workload: the uniform16 recipe (n_H = exp(0.1 z), T = 300 K) with seed 2,
20 outer steps, no perturbations."""
import numpy as np
import pytest

from states import random_states, trajectory_cases

NSTEPS = 20


@pytest.fixture(scope="module")
def adj():
    import oracle_lib
    return oracle_lib.OracleAdjoint()


def _states(adj, wr):
    ys = [adj.burn_to_y(b) for b, _ in trajectory_cases(adj, wr.CELL_DTYPE, n_cells=4,
                                                       steps=(0, 1, 5, 20, 60, 120, 300))]
    ys += [y for y in random_states(120, adj, seed=3)
           if np.isfinite(adj.rhs(y)).all() and np.isfinite(adj.jac(y)).all()][:40]
    return ys


def test_generated_forward_is_the_network(adj, wr):
    M1 = np.ones(adj.R)
    for y in _states(adj, wr):
        f, J = adj.fj(y, M1)
        assert np.array_equal(adj.f(y, M1).view(np.uint64), adj.rhs(y).view(np.uint64))
        assert np.array_equal(f.view(np.uint64), adj.rhs(y).view(np.uint64))
        assert np.array_equal(J.view(np.uint64), adj.jac(y).ravel().view(np.uint64))


def test_rate_multipliers_scale_reaction_fluxes(adj, wr):
    """f is affine in each multiplier and the flux terms conserve what the
    reactions conserve: charge (e- and ion charges) of the species rates is
    zero for any multipliers, as it is for the network."""
    charge = np.array([-1, 1, 0, -1, 1, 0, 1, -1, 0, 1, 0, 2, 1, 0], dtype=float)
    rng = np.random.default_rng(7)
    for y in _states(adj, wr)[:20]:
        M = np.exp(rng.normal(0, 1.0, adj.R))
        f = adj.f(y, M)
        scale = np.max(np.abs(f[:14] * charge)) + 1e-300
        assert abs(np.sum(f[:14] * charge)) <= 1e-10 * scale
        r = rng.integers(adj.R)
        M2 = M.copy()
        M2[r] *= 3.0
        M3 = M.copy()
        M3[r] *= 5.0
        d2, d3 = adj.f(y, M2) - f, adj.f(y, M3) - f
        assert np.allclose(d3, 2.0 * d2, rtol=1e-9, atol=1e-12 * (np.abs(f).max() + 1e-300))


def test_generated_tangents_match_finite_differences(adj, wr):
    rng = np.random.default_rng(8)
    errs = []
    # (the first steps sit at T = 100 K, a branch point of piecewise rate fits,
    # where central differences straddle a jump)
    for y, _ in [(adj.burn_to_y(b), dt) for b, dt in
                 trajectory_cases(adj, wr.CELL_DTYPE, n_cells=4, steps=(5, 20, 60, 120, 300))]:
        M = np.exp(rng.normal(0, 0.1, adj.R))
        dy = rng.standard_normal(15) * np.abs(y)
        dM = rng.standard_normal(adj.R) * M
        eps = 1e-6
        df = adj.f_jvp(y, M, dy, dM)
        fd = (adj.f(y + eps * dy, M + eps * dM) - adj.f(y - eps * dy, M - eps * dM)) / (2 * eps)
        errs.append(np.max(np.abs(df - fd)) / np.max(np.abs(fd)))
        dJ = adj.jac_jvp(y, dy)
        fdJ = (adj.fj(y + eps * dy, M)[1] - adj.fj(y - eps * dy, M)[1]) / (2 * eps)
        errs.append(np.max(np.abs(dJ - fdJ)) / np.max(np.abs(fdJ)))
    assert np.median(errs) < 1e-8 and np.max(errs) < 1e-5, (np.median(errs), np.max(errs))


def test_generated_adjoints_are_transposed_tangents(adj, wr):
    rng = np.random.default_rng(9)
    for y in _states(adj, wr):
        M = np.exp(rng.normal(0, 0.1, adj.R))
        dy = rng.standard_normal(15) * (np.abs(y) + 1e-30)
        dM = rng.standard_normal(adj.R)
        w = rng.standard_normal(15)
        yb, mb = adj.f_vjp(y, M, w)
        lhs = w @ adj.f_jvp(y, M, dy, dM)
        rhs = yb @ dy + mb @ dM
        assert abs(lhs - rhs) <= 1e-11 * (np.abs(w) @ np.abs(adj.f_jvp(y, M, dy, dM)) + abs(lhs))
        W = rng.standard_normal(225)
        dJ = adj.jac_jvp(y, dy)
        lhs = W @ dJ
        rhs = adj.jac_vjp(y, W) @ dy
        assert abs(lhs - rhs) <= 1e-11 * (np.abs(W) @ np.abs(dJ) + abs(lhs))


def test_transposed_lu_solve(adj):
    rng = np.random.default_rng(10)
    for k in range(20):
        a = rng.standard_normal((15, 15)) + np.eye(15) * rng.uniform(0, 3)
        lu, ip, info = adj.lu(a)
        assert info == 0
        b = rng.standard_normal(15)
        x = adj.lu_solve_t(lu, ip, b)
        assert np.allclose(x, np.linalg.solve(a.T, b), rtol=1e-9, atol=1e-12)


@pytest.fixture(scope="module")
def recorded(adj, wr):
    n = 8
    seed, sigma, t0, t1 = wr.RECIPES["uniform16"]
    cells = adj.make_recipe_cells(n, wr.CELL_DTYPE, 2, sigma, t0, t1)
    plain = cells.copy()
    tape = adj.record(cells, NSTEPS, np.ones(adj.R))
    assert tape["status"] == 1
    return cells, plain, tape


def _state_of(cells):
    c = cells
    return np.concatenate([c["current"]["xn"].T, c["current"]["T"][None], c["current"]["e"][None],
                           c["current"]["rho"][None], c["density_driver"][None]])


def test_record_is_the_plain_run(adj, wr, recorded):
    cells, plain, tape = recorded
    r, completed = adj.run_gridwide(plain, False, max_steps=NSTEPS)
    assert r == 1 and completed == NSTEPS
    mask = np.ones(216, bool)
    mask[156:160] = False
    n = len(cells)
    assert np.array_equal(cells.view(np.uint8).reshape(n, 216)[:, mask],
                          plain.view(np.uint8).reshape(n, 216)[:, mask])


def test_replay_reproduces_the_record(adj, recorded):
    cells, _, tape = recorded
    sN = adj.replay(tape, tape["ckpt"][0])
    assert np.array_equal(sN.view(np.uint64), _state_of(cells).view(np.uint64))


def _direction(rng, s0, R):
    v = rng.standard_normal(s0.shape) * np.abs(s0)
    v[15] = 0.0  # e at the first advance is recomputed from T (eos_rt)
    return v, rng.standard_normal(R)


def test_tangent_matches_central_differences(adj, recorded):
    _, _, tape = recorded
    s0 = tape["ckpt"][0]
    M = tape["M"]
    rng = np.random.default_rng(11)
    delta = 1e-6
    for k in range(8):
        v, dth = _direction(rng, s0, adj.R)
        dM = dth * M  # M = exp(theta): dM = M dtheta
        _, tl = adj.replay(tape, s0, ds0=v, dM=dM)
        sp = adj.replay(tape, s0 + delta * v, M=M * np.exp(delta * dth))
        sm = adj.replay(tape, s0 - delta * v, M=M * np.exp(-delta * dth))
        fd = (sp - sm) / (2 * delta)
        for rows in (slice(0, 14), slice(14, 15), slice(15, 16), slice(16, 18)):
            scale = np.max(np.abs(fd[rows])) + 1e-300
            assert np.max(np.abs(fd[rows] - tl[rows])) <= 1e-4 * scale, (k, rows)


def test_adjoint_is_the_transposed_tangent(adj, recorded):
    _, _, tape = recorded
    s0 = tape["ckpt"][0]
    rng = np.random.default_rng(12)
    for k in range(4):
        gN = rng.standard_normal(s0.shape)
        g0, mb = adj.backward(tape, gN)
        v, dM = _direction(rng, s0, adj.R)
        _, tl = adj.replay(tape, s0, ds0=v, dM=dM)
        lhs = np.sum(gN * tl)
        rhs = np.sum(g0 * v) + mb.sum(axis=1) @ dM
        assert abs(lhs - rhs) <= 1e-11 * (np.sum(np.abs(gN * tl)) + abs(lhs)), (lhs, rhs)
        assert np.all(g0[15] == 0.0)
