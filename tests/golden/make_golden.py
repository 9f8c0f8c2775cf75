"""Write the golden fixtures of tests/golden/ from the reference's own CPU code.

Usage (in a checkout where the reference source exists):
    make -C oracle ref            # builds oracle/_ref/libwr_ref.so from reproducer.cpp
    python tests/golden/make_golden.py

Every expected value below is computed by the reference's compiled functions
(PrimordialChem::rhs/jacobian, lu_decomposition/lu_solve, RODAS::integrate,
burn_ros2s, the collapse driver, write_final_states) through
oracle/ref_harness.cpp.  Inputs are seeded.  Files (numpy .npz, no pickles):

  rhs_jac.npz     states y[96,15], f[96,15], J[96,15,15]
  lu.npz          matrices, LU factors, pivots, info, solve rhs and solutions
  integrate.npz   ROS2S integrations from collapse trajectories: y0, tout,
                  dt0, y, counters, t_end, dt_end, result (+ a forced
                  TOO_MANY_STEPS case)
  driver.npz      splitmix64 / perturbation-factor values, prepare/advance
                  semantic cases, final cells of the grid-1 run and of small
                  gridwide runs (216-byte records)
  final_states_grid2.bin   packed file written by the reference writer
  final_states_grid2_legacy_local.bin   the same for the reference CPU build's
                  legacy-local loop (reproducer.cpp:8628-8648) on 8 cells
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle_lib  # noqa: E402
from states import lu_test_matrices, random_matrices, random_states, trajectory_cases  # noqa: E402

CELL_DTYPE = np.dtype([
    ("current", [("rho", "<f8"), ("T", "<f8"), ("e", "<f8"), ("xn", "<f8", (14,))]),
    ("time", "<f8"), ("density_driver", "<f8"), ("completed_steps", "<i4"), ("reserved_", "<i4"),
    ("stats", [(f, "<u8") for f in ("internal_steps", "rhs_calls", "jacobian_calls",
                                    "decompositions", "linear_solves", "accepted_steps",
                                    "rejected_steps")])])


def main():
    ref = oracle_lib.Reference.load()
    if ref is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    meta = {}
    src = os.path.join(os.environ.get("WR_REFERENCE", "/root/reference"), "reproducer.cpp")
    if os.path.exists(src):
        meta["reproducer_cpp_sha256"] = hashlib.sha256(open(src, "rb").read()).hexdigest()

    # -- RHS / Jacobian ---------------------------------------------------------
    ys = random_states(96, ref, seed=101)
    f = np.array([ref.rhs(y) for y in ys])
    J = np.array([ref.jac(y) for y in ys])
    np.savez(os.path.join(HERE, "rhs_jac.npz"), y=ys, f=f, J=J)

    # -- LU -----------------------------------------------------------------------
    mats, rhs = lu_test_matrices()
    extra = random_matrices(20, seed=102)
    # E = fac*I - J of real states at several step sizes (the integrator's matrices)
    e_mats = []
    for k, y in enumerate(ys[:12]):
        for h in (1e2, 1e8, 1e14):
            fac = 1.0 / (h * 0.292893218813452)
            e_mats.append(-J[k] + np.eye(15) * fac)
    allm = np.concatenate([mats, extra, np.array(e_mats)])
    rng = np.random.default_rng(103)
    b = np.concatenate([rhs, rng.standard_normal((len(allm) - len(rhs), 15))])
    lu, ip, info, x = [], [], [], []
    for m, bb in zip(allm, b):
        a, p, i = ref.lu(m)
        lu.append(a)
        ip.append(p)
        info.append(i)
        x.append(ref.lu_solve(a, p, bb) if i == 0 else np.full(15, np.nan))
    np.savez(os.path.join(HERE, "lu.npz"), a=allm, lu=np.array(lu), ipvt=np.array(ip),
             info=np.array(info, dtype=np.int32), b=b, x=np.array(x))

    # -- integrator ---------------------------------------------------------------
    cases = trajectory_cases(ref, CELL_DTYPE, n_cells=4, steps=(0, 3, 20, 80, 160, 240, 320, 400))
    y0 = np.array([ref.burn_to_y(bu) for bu, _ in cases])
    dts = np.array([dt for _, dt in cases])
    outs = [ref.integrate(y, dt, dt) for y, dt in zip(y0, dts)]
    # forced TOO_MANY_STEPS (max_steps = 3) on the first case
    forced = ref.integrate(y0[0], dts[0], dts[0], max_steps=3)
    np.savez(os.path.join(HERE, "integrate.npz"), y0=y0, tout=dts, dt0=dts,
             y=np.array([o["y"] for o in outs]), counters=np.array([o["counters"] for o in outs]),
             t_end=np.array([o["t_end"] for o in outs]), dt_end=np.array([o["dt_end"] for o in outs]),
             result=np.array([o["result"] for o in outs], dtype=np.int32),
             forced_y=forced["y"], forced_counters=forced["counters"],
             forced_result=np.int32(forced["result"]), forced_t_end=forced["t_end"],
             forced_dt_end=forced["dt_end"])

    # -- driver -------------------------------------------------------------------
    seeds = np.array([0, 1, 1234567, (5 << 32) ^ (20 << 16), 2**64 - 1], dtype=np.uint64)
    sm = np.array([ref.splitmix64(int(s)) for s in seeds], dtype=np.uint64)
    pf_cells = np.array([0, 1, 7, 31, 262143], dtype=np.int32)
    pf_steps = np.array([20, 40, 980], dtype=np.int32)
    pf = np.array([[ref.perturbation_factor(int(c), int(s)) for s in pf_steps] for c in pf_cells])

    def grid_run(n, perturb):
        cells = ref.make_cells(n, CELL_DTYPE)
        r, done = ref.run_gridwide(cells, perturb)
        assert r == 1
        return cells, done

    g1, d1 = grid_run(1, True)
    g8, d8 = grid_run(8, True)
    g27, d27 = grid_run(27, False)
    # prepare semantics (tests/integration/test_grid_timestep_reference.mojo:33-71)
    term = ref.make_cells(1, CELL_DTYPE)
    term["current"]["xn"][0] = 0.0
    term["current"]["xn"][0][2] = 1.0e30
    term["density_driver"] = 1.0e-10
    term_dt, term_f = ref.prepare_cell(term, 0, 0, 0.0, 0, False)
    bad = ref.make_cells(1, CELL_DTYPE)
    bad["current"]["xn"][0] = 0.0
    bad_dt, bad_f = ref.prepare_cell(bad, 0, 0, 0.0, 0, False)
    lim = ref.make_cells(1, CELL_DTYPE)
    lim["density_driver"] = 1.99e-6
    dt_lim = ref.collapse_timestep(lim)
    lim_r, lim_f = ref.advance_cell(lim, 0, dt_lim, dt_lim)
    np.savez(os.path.join(HERE, "driver.npz"), splitmix_in=seeds, splitmix_out=sm,
             pf_cells=pf_cells, pf_steps=pf_steps, pf=pf,
             grid1=g1.view(np.uint8), grid1_steps=np.int32(d1),
             grid2=g8.view(np.uint8), grid2_steps=np.int32(d8),
             grid3_noperturb=g27.view(np.uint8), grid3_steps=np.int32(d27),
             terminal_cell=term.view(np.uint8), terminal_dt=term_dt, terminal_failure=np.int32(term_f),
             bad_failure=np.int32(bad_f), limit_cell=lim.view(np.uint8), limit_integrated=np.int32(lim_r),
             limit_failure=np.int32(lim_f))

    # -- packed file written by the reference writer --------------------------------
    path = os.path.join(HERE, "final_states_grid2.bin")
    if os.path.exists(path):
        os.remove(path)
    ref.lib.ref_write_final_states.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_char_p]
    ok = ref.lib.ref_write_final_states(g8.ctypes.data_as(ctypes.c_void_p), 8, 2, path.encode())
    assert ok == 1
    legacy = ref.make_cells(8, CELL_DTYPE)
    r, done = ref.run_legacy(legacy, True)
    assert r == 1
    path = os.path.join(HERE, "final_states_grid2_legacy_local.bin")
    if os.path.exists(path):
        os.remove(path)
    ok = ref.lib.ref_write_final_states(legacy.ctypes.data_as(ctypes.c_void_p), 8, 2, path.encode())
    assert ok == 1
    meta["grid2_legacy_local_steps"] = int(done)
    meta["files"] = {n: hashlib.sha256(open(os.path.join(HERE, n), "rb").read()).hexdigest()
                     for n in sorted(os.listdir(HERE)) if n.endswith((".npz", ".bin"))}
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
