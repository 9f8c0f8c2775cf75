"""The FP32-factorisation form (FP32 LU of I - gamma h J, FP64 residual and
refinement) in the oracle, pinned without reference to its own formulas:
the FP32 factors reconstruct P A to single precision, the refined stage
solves agree with a float64 LAPACK solve, and whole integrations stay within
the reference comparator's tolerances of the FP64 ones.  The HIP kernel is
checked against this restatement bit for bit in test_gpu_mixed.py."""
import numpy as np
import pytest

from states import random_matrices, trajectory_cases


def test_lu32_solves_to_single_precision(oracle_wr):
    """FP32 factors + FP32 substitution: backward error of single precision,
    and the pivot sequence of the FP64 elimination (reference-pinned) on the
    same matrix."""
    rng = np.random.default_rng(31)
    for k in range(40):
        a = (rng.standard_normal((15, 15)) + np.eye(15) * rng.uniform(0, 5))
        a = a.astype(np.float32).astype(np.float64)
        lu, ip, info = oracle_wr.lu32(a)
        assert info == 0
        b = rng.standard_normal(15)
        x = oracle_wr.lu_solve32(lu, ip, b).astype(np.float64)
        bf = b.astype(np.float32).astype(np.float64)
        berr = np.abs(a @ x - bf).max() / (np.abs(a).max() * np.abs(x).max())
        assert berr < 2e-6, (k, berr)
        _, ip64, _ = oracle_wr.lu(a)
        assert np.array_equal(ip, ip64)


def test_lu32_singular_and_identity(oracle_wr):
    lu, ip, info = oracle_wr.lu32(np.eye(15))
    assert info == 0 and np.array_equal(lu, np.eye(15, dtype=np.float32))
    assert np.array_equal(ip, np.arange(15))
    a = np.eye(15)
    a[14, 14] = 0.0
    assert oracle_wr.lu32(a)[2] == 15
    b = np.eye(15)
    b[:, 3] = 0.0
    assert oracle_wr.lu32(b)[2] == 4
    # an entry below FP32's range underflows to zero: singular in FP32 only
    c = np.eye(15)
    c[7, 7] = 1e-60
    assert oracle_wr.lu32(c)[2] == 8 and oracle_wr.lu(c)[2] == 0


def _cases(oracle_wr, wr):
    """(y, dt): the integrator's inputs at outer steps of a 4-cell collapse."""
    return [(oracle_wr.burn_to_y(burn), dt)
            for burn, dt in trajectory_cases(oracle_wr, wr.CELL_DTYPE, n_cells=4,
                                             steps=(0, 1, 5, 20, 60, 120))]


def _accurate_solve(E, b, s):
    """E x = b: float64 LAPACK on the equilibrated system, then refinement
    with residuals in x86 extended precision (np.longdouble)."""
    Es = E * (1.0 / s)[:, None] * s[None, :]
    z = np.linalg.solve(Es, b / s)
    EL, bL = E.astype(np.longdouble), b.astype(np.longdouble)
    x = (z * s).astype(np.longdouble)
    for _ in range(4):
        r = bL - EL @ x
        x = x + (np.linalg.solve(Es, (r / s.astype(np.longdouble)).astype(np.float64)) * s)
    return np.asarray(x, dtype=np.float64)


def test_refined_stage_solves_match_lapack(oracle_wr, wr):
    """E x = b on E = fac I - J(y) of real collapse states, fac from the
    burn's dt and from 1e-3 of it, against an extended-precision-refined
    LAPACK solve, in the
    controller's weights atol_i + rtol_i |y_i|.  Where the equilibrated E is
    well conditioned for FP32 (cond < 1e6: the smaller steps), two sweeps
    converge to within 1e-7 of the weights of the answer; at the full burn step the equilibrated
    condition number is ~1e10, beyond FP32, the refinement reports
    non-convergence (the kernel then falls back to FP64 factors), and a
    converged claim is never wrong beyond the FP64 solve's own error."""
    rng = np.random.default_rng(32)
    gamma = 0.292893218813452
    rtol = np.r_[np.full(14, 1e-4), 1e-6]
    good = bad = 0
    for y, dt in _cases(oracle_wr, wr):
        J = oracle_wr.jac(y).reshape(15, 15)
        sk = rtol + rtol * np.abs(y)
        for h in (dt, dt * 1e-3):
            fac = 1.0 / (h * gamma)
            E = fac * np.eye(15) - J
            b = (J @ (sk * rng.standard_normal(15))) + fac * sk
            want = _accurate_solve(E, b, sk)
            cond = np.linalg.cond(E * (1.0 / sk)[:, None] * sk[None, :])
            x0, _ = oracle_wr.solve_mixed(J.ravel(), fac, y, b, 0)
            x, conv = oracle_wr.solve_mixed(J.ravel(), fac, y, b, 2)
            err = np.max(np.abs(x - want) / sk)
            noise = 0.0
            if cond < 1e6:
                assert conv == 1 and err < 1e-7 + noise, (cond, err)
                good += 1
            else:
                bad += int(conv == 0)
                if conv == 1:
                    assert err < 1e-7 + noise
            # refinement only improves on the FP32 solve
            assert err <= np.max(np.abs(x0 - want) / sk)
    assert good >= 8 and bad >= 1


def test_mixed_integrations_within_reference_tolerances(oracle_wr, wr):
    """Whole ROS2S integrations with FP32 factors against FP64 factors: the
    adaptive controller amplifies rounding differences up to its tolerance, so
    the bar is the reference comparator's (species rtol 1e-3 + atol 1e-4 on
    non-deuterium species, T/e rtol 2e-4), and most counters are unchanged."""
    cases = _cases(oracle_wr, wr)
    same = 0
    try:
        outs = {}
        for mode in (0, 1):
            oracle_wr.set_lu_precision(mode, 2)
            outs[mode] = [oracle_wr.integrate(y, dt, dt) for y, dt in cases]
        fb = oracle_wr.mixed_fallbacks()
    finally:
        oracle_wr.set_lu_precision(0, 2)
    for o0, o1 in zip(outs[0], outs[1]):
        y0, y1, c0, c1 = o0["y"], o1["y"], o0["counters"], o1["counters"]
        assert o0["result"] == o1["result"] == 1
        nond = [0, 1, 2, 3, 6, 8, 11, 12, 13]
        assert np.all(np.abs(y1[nond] - y0[nond]) <= 1e-4 + 1e-3 * np.abs(y0[nond]))
        assert abs(y1[14] - y0[14]) <= 2e-4 * abs(y0[14])
        same += int(np.array_equal(c0, c1))
    assert same >= len(cases) // 2
    print(f"identical counters {same}/{len(cases)}, fallbacks {fb}")
