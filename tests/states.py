"""Seeded test inputs: chemical states y = (n_0..n_13, e) and matrices.

States are drawn so that every temperature branch of the network is visited
(T log-uniform over 3 K .. 1e5 K), abundances span 1e-40 .. 1 relative to
hydrogen with occasional exact zeros (the 1e-100 floor path), and the gas
density spans 1e-3 .. 1e12 cm^-3.  e is set from T through the oracle's
burn_state_from_y so that T(e, X) reproduces the drawn T.
"""
import numpy as np


def random_states(n, oracle, seed=0, t_range=(0.5, 5.0)):
    rng = np.random.default_rng(seed)
    ys = np.zeros((n, 15))
    for k in range(n):
        X = 10.0 ** rng.uniform(-40, 0, 14)
        X[2] = 1.0
        X *= 10.0 ** rng.uniform(-3, 12)
        if k % 7 == 3:
            X[rng.integers(0, 14, 3)] = 0.0
        T = 10.0 ** rng.uniform(*t_range)
        y = np.concatenate([X, [1.0]])
        t1 = oracle.temperature_from_y(y)
        y[14] = T / t1
        ys[k] = y
    return ys


def lu_test_matrices():
    """The reference's LU validation matrices (tests/gpu/test_warp_lu.mojo:19-47):
    identity, exactly singular last row, forced row swap, pivot tie,
    diagonally dominant pseudo-random, near-singular 1e-10 pivot."""
    mats, rhss = [], []
    for case in range(8):
        a = np.zeros((15, 15))
        b = np.zeros(15)
        for r in range(15):
            for c in range(15):
                if case <= 3:
                    v = 1.0 if r == c else 0.0
                else:
                    v = float((r * 17 + c * 13 + case * 7) % 19 - 9) * 0.01
                    if r == c:
                        v += 2.0
                a[r, c] = v
            b[r] = float((r * 5 + case * 3) % 11 - 5)
        if case == 1:
            a[14, :] = 0.0
        elif case == 2:
            a[0, 0] = 0.0
            a[1, 0] = 2.0
        elif case == 3:
            a[0, 0] = 1.0
            a[1, 0] = -3.0
            a[2, 0] = 3.0
        elif case == 7:
            a[14, 14] = 1.0e-10
        mats.append(a)
        rhss.append(b)
    return np.array(mats), np.array(rhss)


def random_matrices(n, seed=0):
    rng = np.random.default_rng(seed)
    out = rng.standard_normal((n, 15, 15))
    for k in range(n):
        if k % 5 == 0:
            out[k, rng.integers(0, 15), :] = 0.0
        if k % 3 == 0:
            out[k, :, rng.integers(0, 15)] *= 1e-8
        if k % 11 == 0:
            out[k] += np.eye(15) * 10.0 ** rng.uniform(-2, 8)
    return out


def trajectory_cases(oracle, cell_dtype, n_cells=8, steps=(0, 1, 5, 20, 60, 120, 200, 300, 400),
                     perturb=True):
    """(y, dt) pairs from a gridwide run of the oracle at the states where the
    collapse actually calls the integrator: the burn input of each outer step."""
    cells = oracle.make_cells(n_cells, cell_dtype)
    cases = []
    completed, grid_time = 0, 0.0
    want = set(steps)
    for step in range(max(steps) + 1):
        cands = []
        for c in range(n_cells):
            dt, f = oracle.prepare_cell(cells[c:c + 1], c, completed, grid_time, step, perturb)
            cands.append(dt)
        dtg = min(cands)
        if dtg == np.finfo(np.float64).max:
            break
        if step in want:
            for c in range(n_cells):
                s = cells[c].copy()
                # the burn input: advance_cell's density update and eos_rt
                local_dt = oracle.collapse_timestep(cells[c:c + 1])
                if local_dt < 10.0:
                    continue
                old = s["density_driver"]
                drv = old + dtg * (old / (local_dt / 0.1))
                if not (drv > 0 and drv <= 2e-6):
                    continue
                ratio = drv / old
                burn = np.concatenate([[s["current"]["rho"] * ratio, s["current"]["T"], 0.0],
                                       s["current"]["xn"] * ratio])
                cases.append((burn, dtg))
        nxt = grid_time + dtg
        for c in range(n_cells):
            oracle.advance_cell(cells[c:c + 1], completed, nxt, dtg)
        grid_time = nxt
        completed += 1
    return cases
