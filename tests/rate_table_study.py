"""Rate-table study on the CPU oracle (not collected by pytest; run it as a
script).  The collapse run (reference initial states, density perturbations)
on a g^3 grid is integrated with the network as generated, with the network
split at its temperature-only frontier (exact values: must be bit-identical),
and with the frontier values from cubic tables in ln T of several sizes.
Printed per table: final-state errors against the exact run (max relative
error over the non-deuterium species and T, entries above 1e-20), the change
in accepted / rejected steps, and the interpolation error of the frontier
values at the temperatures the run visits.

    python tests/rate_table_study.py --grid 16 --steps 1000
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib  # noqa: E402

wr = importlib.import_module("warp-reactor_amd")


def rel_max(a, b, floor=1e-20):
    m = np.abs(b) > floor
    return float((np.abs(a - b)[m] / np.abs(b)[m]).max(initial=0.0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=16)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--tlo", type=float, default=10.0)
    ap.add_argument("--thi", type=float, default=1e5)
    ap.add_argument("--tables", default="2:1024,2:4096,2:16384,3:16384,3:65536,3:262144",
                    help="mode:nodes list (2: cubic in ln|v|, 3: cubic in v)")
    args = ap.parse_args()
    o = oracle_lib.OracleTable()
    n = args.grid ** 3
    init = wr.make_collapse_states(n)

    def run(mode, tlo=1.0, thi=2.0, nodes=4):
        o.rt_set(mode, tlo, thi, nodes)
        c = init.copy()
        t0 = time.perf_counter()
        r, done = o.run_gridwide(c, True, args.steps)
        el = time.perf_counter() - t0
        outside = o.rt_outside()
        o.rt_set(0)
        return c, r, done, el, outside

    ex, r0, d0, e0, _ = run(0)
    sp, r1, d1, e1, _ = run(1)
    same = ex.tobytes() == sp.tobytes()
    print(f"# {n} cells x {d0} steps: generated network {e0:.1f} s, split network {e1:.1f} s, "
          f"bit-identical {same}; T {ex['current']['T'].min():.4g}..{ex['current']['T'].max():.4g} K",
          flush=True)
    nd = [k for k, name in enumerate(wr.SPECIES_NAMES) if "D" not in name]
    rng = np.random.default_rng(3)
    tq = np.exp(rng.uniform(np.log(init["current"]["T"].min()),
                            np.log(ex["current"]["T"].max()), 20000))
    rows = []
    for item in args.tables.split(","):
        mode, nodes = (int(v) for v in item.split(":"))
        c, r, done, el, outside = run(mode, args.tlo, args.thi, nodes)
        o.rt_set(mode, args.tlo, args.thi, nodes)
        te = o.rt_errors(tq)
        o.rt_set(0)
        row = {"mode": "ln|v|" if mode == 2 else "v", "nodes": nodes,
               "table_mb": nodes * o.ntv * 8 / 1e6, "status": r, "steps": done,
               "species_max_rel_err": max(rel_max(c["current"]["xn"][:, k], ex["current"]["xn"][:, k])
                                          for k in nd),
               "T_max_rel_err": rel_max(c["current"]["T"], ex["current"]["T"]),
               "d_accepted": int(c["stats"]["accepted_steps"].sum()) - int(ex["stats"]["accepted_steps"].sum()),
               "d_rejected": int(c["stats"]["rejected_steps"].sum()) - int(ex["stats"]["rejected_steps"].sum()),
               "interp_rel_err_median": float(np.median(te)),
               "interp_rel_err_p99": float(np.quantile(te, 0.99)),
               "outside_range": outside, "seconds": el}
        rows.append(row)
        print(json.dumps(row), flush=True)
    print(json.dumps({"cells": n, "steps": d0, "frontier_values": o.ntv, "split_bit_identical": same,
                      "tables": rows}))


if __name__ == "__main__":
    main()
