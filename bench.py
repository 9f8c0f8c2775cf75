"""Benchmark: cell-advances/s of the 64^3 gridwide-independent collapse run.

One bench step = one complete collapse run of the reference's default
configuration -- 64^3 = 262,144 cells, make_collapse_state initial state,
density perturbations every 20 steps, at most 1000 outer steps
(prepare -> grid-wide minimum dt -> ROS2S advance), the run the reference
times on H200 at 17.23 s (docs/performance/optimization-results.md:48-55).
Cell states are uploaded to HBM before each timed run; the timed region is
the run itself (device-resident outer loop, host polls a done flag).

Prints one JSON line (see the keys below).  `value` = sum over outer steps of
integrated cells / wall time of the K timed runs.  `roofline` is the FP64
roofline of the advance kernel: algorithmic FP64 operations (static counts of
the generated network and of the LU/solve/stage arithmetic, times the ROS2S
counters of the run) per advance launch / its mean duration measured with HIP
events on the launch stream.  `cpu_baseline` times the reference's own
compiled CPU code (oracle/_ref/libwr_ref.so) on a bounded, strided sample of
the same workload -- cells of the run replayed through its dt_grid sequence --
on 1 thread and on all the host threads the box gives the process.
`roofline.traffic` is the HBM bytes per advance dispatch measured in the same
run of bench.py: two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over the
same workload and launch shape, each calibrated by 1 GiB device copies in the
same process (warp-reactor_amd/prof/traffic_run.py).
`plain_fp64` times the same run with the thread-per-cell kernel (no queue,
network left to the compiler's register allocation) and checks that the
tuned default produces bit-identical final states.  `adjoint_default_run`
differentiates the first 100 steps of the same run; it is measured in a child
process started before this one uses the GPU.
"""
import argparse
import importlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP64_PEAK_TFLOPS = 78.6  # MI355X vector FP64 (= FP64 matrix), dense
PROF_DIR = os.path.join(REPO, "warp-reactor_amd", "prof")


def live_traffic(grid, max_outer, shape, kernel_name="k_advance_q", timeout=300):
    """HBM bytes per dispatch of the advance kernel in this build, measured
    now: two runs of prof/traffic_run.py (same workload and launch shape)
    under `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`, each scaled by
    what the counter reads for the run's own 1 GiB device copies.  Returns a
    dict, or None when rocprofv3 is missing or a pass fails."""
    sys.path.insert(0, PROF_DIR)
    import pmc_summary
    sys.path.pop(0)
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    out = {}
    tmp = tempfile.mkdtemp(prefix="wr_pmc_")
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "t", "--",
                   sys.executable, os.path.join(PROF_DIR, "traffic_run.py"), "--grid", str(grid),
                   "--max-outer", str(max_outer), "--threads", str(shape[0]), "--per-cu",
                   str(shape[1]), "--solo", str(shape[2])]
            if len(shape) > 3:
                cmd += ["--coop", str(shape[3])]
            env = dict(os.environ, TMPDIR="/tmp")
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd="/tmp",
                               env=env)
            csvs = [os.path.join(dp, f) for dp, _, fs in os.walk(d) for f in fs
                    if f.endswith("counter_collection.csv")]
            if r.returncode != 0 and not csvs:
                print(f"# traffic: rocprofv3 {counter} failed ({r.returncode}): "
                      f"{r.stderr[-300:]}", file=sys.stderr, flush=True)
                return None
            summ = {}
            for c in csvs:
                summ.update(pmc_summary.summarize(c))
            copy = [v[counter] for k, v in summ.items() if "copyBuffer" in k and counter in v]
            kern = [v[counter] for k, v in summ.items() if kernel_name in k and counter in v]
            if not copy or not kern:
                print(f"# traffic: no {counter} rows for the copy or {kernel_name}",
                      file=sys.stderr, flush=True)
                return None
            # the runtime splits each 1 GiB copy into several dispatches: the
            # scale is the bytes the copies read (or wrote), 3 GiB, over what
            # the counter summed over all their dispatches
            copy_kib = sum(nd * mean for nd, mean in copy)
            scale = 3.0 * (1 << 20) / copy_kib
            out[counter] = {"dispatches": kern[0][0], "kib_per_dispatch_raw": kern[0][1],
                            "copy_dispatches": sum(nd for nd, _ in copy),
                            "copy_kib_total_raw": copy_kib, "scale": scale,
                            "bytes_per_dispatch": kern[0][1] * scale * 1024.0}
    except subprocess.TimeoutExpired:
        print("# traffic: rocprofv3 pass timed out", file=sys.stderr, flush=True)
        return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    total = out["FETCH_SIZE"]["bytes_per_dispatch"] + out["WRITE_SIZE"]["bytes_per_dispatch"]
    return {"bytes_per_dispatch": total, "passes": out}


def network_op_counts():
    path = os.path.join(REPO, "warp-reactor_amd", "csrc", "_gen", "chem_counts.json")
    with open(path) as f:
        c = json.load(f)
    flops = ("add", "sub", "mul", "div", "sqrt", "exp", "log")
    return {k: sum(c[k].get(op, 0) for op in flops) for k in ("chem_fused", "chem_rhs")}


def work_model():
    """FP64 operations per unit of ROS2S work (transcendentals count 1)."""
    n = 15
    eos = 2 * 14 + 14 + 5 * 14 + 4 + 2  # density sum, abarinv/gammasinv sums, T = e/(...)
    lu = sum(1 + (n - 1 - k) + 2 * (n - 1 - k) ** 2 for k in range(n - 1))
    solve = 2 * sum(n - 1 - k for k in range(n - 1)) + n + 2 * sum(range(n))
    stage = 476  # stage vectors, error norm, controller (per attempt)
    c = network_op_counts()
    return {"jac_eval": c["chem_fused"] + eos, "rhs_eval": c["chem_rhs"] + eos,
            "decomp": n + lu, "solve": solve, "attempt": stage}


def run_flops(stats_before, stats_after):
    """Algorithmic FP64 operations of a run from the ROS2S counters:
    per J evaluation the fused J + f(y); per attempt two more RHS, the
    stage vectors and three solves; per decomposition E and its LU."""
    w = work_model()
    d = {f: int(stats_after[f].sum()) - int(stats_before[f].sum()) for f in stats_after.dtype.names}
    attempts = d["internal_steps"]
    return (d["jacobian_calls"] * w["jac_eval"] + 2 * attempts * w["rhs_eval"]
            + d["decompositions"] * w["decomp"] + d["linear_solves"] * w["solve"]
            + attempts * w["attempt"]), d


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(init, dts, budget_s=12.0):
    """The reference's own compiled CPU code (oracle/_ref/libwr_ref.so, built
    from reproducer.cpp) on a bounded, strided sample of the same workload:
    cells of the 64^3 run replayed through the GPU run's dt_grid of every
    outer step (the reference's prepare + advance kernel bodies in its host
    loop's order), about budget_s seconds on 1 thread and again on all the
    host threads the box gives this process (OpenMP over cells)."""
    import ctypes
    so = os.path.join(REPO, "oracle", "_ref", "libwr_ref.so")
    if not os.path.exists(so):
        return None
    lib = ctypes.CDLL(so)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    fn = lib.ref_replay_cells
    fn.argtypes = [vp, vp, i32, vp, i32, i32, vp, i32]
    fn.restype = ctypes.c_longlong
    n = len(init)
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    order = (np.arange(n, dtype=np.int64) * 104729) % n  # prime stride over the grid
    nthreads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16),
                          os.cpu_count() or 1, 16))

    def measure(threads, chunk):
        k, adv, att, t0 = 0, 0, 0, time.perf_counter()
        while time.perf_counter() - t0 < budget_s and k < n:
            idx = np.ascontiguousarray(order[k:k + chunk], dtype=np.int32)
            cells = init[idx].copy()
            status = np.zeros(len(idx), dtype=np.int32)
            att += fn(cells.ctypes.data, idx.ctypes.data, len(idx), dts.ctypes.data, len(dts), 1,
                      status.ctypes.data, threads)
            if not np.all(status == 1):
                return None
            adv += int(cells["completed_steps"].sum())
            k += len(idx)
        el = time.perf_counter() - t0
        return {"cells": k, "cell_advances": adv, "attempts": att, "seconds": el,
                "value": adv / el, "attempts_per_s": att / el}

    one = measure(1, 8)
    allc = measure(nthreads, 8 * nthreads)
    if one is None or allc is None:
        return None
    return {"value": allc["value"], "unit": "cell-advances/s", "cores": nthreads,
            "kind": "reference",
            "attempts_per_s": allc["attempts_per_s"],
            "per_core": {"value": one["value"], "attempts_per_s": one["attempts_per_s"],
                         "cores": 1, "cells": one["cells"], "seconds": one["seconds"]},
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{allc['cells']} cells of the {round(n ** (1 / 3))}^3 run (prime-stride "
                      f"sample) each replayed through all {len(dts)} outer steps of the GPU run's "
                      f"dt_grid with the reference's compiled prepare/advance/ROS2S code: "
                      f"{allc['cell_advances']} cell-advances, {allc['attempts']} ROS2S attempts in "
                      f"{allc['seconds']:.1f} s on {nthreads} threads; 1 thread: {one['cells']} "
                      f"cells in {one['seconds']:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--max-outer", type=int, default=1000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of each of the two CPU-baseline measurements (1 / all threads)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure HBM traffic")
    ap.add_argument("--advance-kernel", type=int, default=1,
                    help="1: queue-fed persistent kernel (default); 0: one thread per cell")
    ap.add_argument("--recipe", default="reference",
                    choices=["reference", "uniform16", "lognormal128"],
                    help="initial states: the reference's make_collapse_state (default) or a "
                         "benchmark recipe (log-normal density, fixed / log-uniform T)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the thread-per-cell comparison run")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the FP32-factorisation run")
    ap.add_argument("--no-jac-reuse", action="store_true",
                    help="skip the Jacobian-reuse study runs")
    ap.add_argument("--jac-ages", default="2,4",
                    help="accepted steps per Jacobian of the Jacobian-reuse study runs")
    ap.add_argument("--no-lognormal", action="store_true",
                    help="skip the 128^3 log-normal recipe measurement")
    ap.add_argument("--no-legacy", action="store_true",
                    help="skip the legacy-local policy run")
    ap.add_argument("--no-cta32", action="store_true",
                    help="skip the gridwide-cta32 policy run")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint measurements (32^3, 20 outer steps; and the "
                         "default run's first 100 steps in a child process)")
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue outer steps as individual launches instead of HIP graphs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape of the queue-fed kernel")
    ap.add_argument("--tune-steps", type=int, default=100,
                    help="outer steps of each autotuner candidate run")
    ap.add_argument("--adjoint-default-run", action="store_true",
                    help="measure only the adjoint of the default run's first 100 outer steps "
                         "(64^3, perturbations), in a process that has run nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    wr = importlib.import_module("warp-reactor_amd")
    if args.adjoint_default_run:
        print(json.dumps({"adjoint_default_run": adjoint_default_run_measure(wr)}), flush=True)
        return
    adjoint_default_run = None
    if not args.no_adjoint and args.grid == 64 and args.recipe == "reference":
        adjoint_default_run = adjoint_default_run_child()
    n = args.grid ** 3
    if args.recipe == "reference":
        init = wr.make_collapse_states(n)
        data = ("synthetic: reference initial state (make_collapse_state) with splitmix64 "
                "density perturbations")
    else:
        seed, sigma, t0, t1 = wr.RECIPES[args.recipe]
        init = wr.make_recipe_states(n, seed, sigma, t0, t1)
        data = (f"synthetic: recipe {args.recipe} (n_H = exp({sigma} z) cm^-3, seed {seed}, "
                f"T {t0:g}-{t1:g} K) with splitmix64 density perturbations")
    ctx = wr.Context(n)
    ctx.set_advance_kernel(args.advance_kernel)
    ctx.set_graph(not args.no_graph)
    # prime: load the kernels with one outer step (not timed, not counted)
    ctx.upload(init)
    ctx.run(perturb=True, max_steps=1)
    tuned = None
    if args.advance_kernel == 1 and not args.no_autotune:
        autotune = importlib.import_module("warp-reactor_amd.autotune")
        chosen, table = autotune.autotune(ctx, init, steps=min(args.tune_steps, args.max_outer))
        tuned = {"chosen": {"threads": chosen[0], "blocks_per_cu": chosen[1],
                            "solo_lanes": chosen[2], "coop_cells": chosen[3]},
                 "steps_per_candidate": min(args.tune_steps, args.max_outer), "candidates": table}
        print(f"# autotune: {chosen} of {len(table)} shapes", file=sys.stderr, flush=True)

    ctx_shape = (ctx.queue_launch()[:3] + (ctx.coop()[0],)) if args.advance_kernel == 1 else None

    def one_run():
        ctx.upload(init)  # inputs resident in HBM before the clock starts
        t0 = time.perf_counter()
        info = ctx.run(perturb=True, max_steps=args.max_outer)
        return info, time.perf_counter() - t0

    for _ in range(args.warmup):
        one_run()
    cells = np.zeros(n, dtype=wr.CELL_DTYPE)
    # timed runs: the production path (outer steps replayed as HIP graphs)
    total_adv, total_s, infos = 0, 0.0, []
    for _ in range(args.steps):
        info, el = one_run()
        total_adv += info["cell_advances"]
        total_s += el
        infos.append(info)
        print(f"# run: {el:.3f} s, outer steps {info['completed_global_steps']}, "
              f"cell-advances {info['cell_advances']}, status {info['status']}",
              file=sys.stderr, flush=True)
    ctx.download(cells)
    # one more run with HIP events around every advance launch (individual
    # launches) for the kernel's mean duration
    ctx.set_timing(True)
    info_t, el_t = one_run()
    adv_ms, adv_launches = ctx.advance_timing()
    ctx.set_timing(False)
    timed_cells = np.zeros(n, dtype=wr.CELL_DTYPE)
    ctx.download(timed_cells)
    assert timed_cells.tobytes() == cells.tobytes(), "instrumented run differs"
    print(f"# instrumented run: {el_t:.3f} s, advance kernel {adv_ms:.1f} ms over "
          f"{adv_launches} launches", file=sys.stderr, flush=True)
    flops, counters = run_flops(init["stats"], cells["stats"])
    mean_launch_s = adv_ms / 1e3 / max(adv_launches, 1)
    per_launch_flops = flops / max(adv_launches, 1)
    achieved = per_launch_flops / mean_launch_s / 1e12 if adv_launches else None
    value = total_adv / total_s
    line = {
        "metric": f"cell-advances/s, {args.grid}^3 gridwide-independent ROS2S collapse (FP64; "
                  f"tuned queue-fed kernel, final states bit-identical to the plain "
                  f"thread-per-cell kernel)",
        "value": value,
        "unit": "cell-advances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": data,
        "config": {"workload": f"primordial collapse {args.grid}^3 = {n} cells, "
                               f"{args.recipe} initial states, <= "
                               f"{args.max_outer} outer steps, perturb every 20, plain FP64",
                   "cells": n, "outer_steps": infos[-1]["completed_global_steps"],
                   "cell_advances_per_run": infos[-1]["cell_advances"],
                   "run_seconds": total_s / args.steps,
                   "h200_reference_run_seconds": 17.2347},
        "roofline": {"bound": "valu_fp64", "compute_pipe": "VALU FP64 (vector FP64 instructions "
                     "only: no MFMA; MI355X FP64 peak 78.6 TFLOP/s, vector = matrix)", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None,
                     "traffic": None,
                     "kernel": "k_advance_q" if args.advance_kernel == 1 else "k_advance",
                     "mean_launch_ms": mean_launch_s * 1e3,
                     "flops_per_launch": per_launch_flops},
        "counters": counters,
    }
    if tuned:
        line["autotune"] = tuned
    dts = ctx.dt_history()
    # the reference-equivalent region (reproducer.cpp:8626-8650): device
    # allocation, host-to-device copy, the run, device-to-host copy, free
    host_cells = init.copy()
    t0 = time.perf_counter()
    r_eq, _ = wr.run_cells(host_cells, perturb=True)
    line["config"]["reference_equivalent_region_seconds"] = time.perf_counter() - t0
    line["config"]["reference_equivalent_region_note"] = (
        "wr_run_cells on host cells: context create (hipMalloc), upload, run, download, destroy; "
        "status %d" % r_eq)
    if not args.no_plain and args.advance_kernel == 1:
        # the same run with the thread-per-cell kernel: time and bit identity
        ctx.set_advance_kernel(0)
        info0, el0 = one_run()
        plain = np.zeros(n, dtype=wr.CELL_DTYPE)
        ctx.download(plain)
        mask = np.ones(216, bool)
        mask[156:160] = False
        same = bool(np.array_equal(cells.view(np.uint8).reshape(n, 216)[:, mask],
                                   plain.view(np.uint8).reshape(n, 216)[:, mask]))
        a, b = cells["current"]["xn"], plain["current"]["xn"]
        floor = np.abs(b) > 1e-20
        rel = np.abs(a - b)[floor] / np.abs(b)[floor]
        plain_value = info0["cell_advances"] / el0
        line["plain_fp64"] = {"kernel": "k_advance (one thread per cell)", "run_seconds": el0,
                              "value": plain_value, "tuned_speedup": value / plain_value,
                              "tuned_final_states_bit_identical": same,
                              "tuned_max_species_rel_err": float(rel.max()) if rel.size else 0.0}
        print(f"# plain run: {el0:.3f} s, bit-identical {same}", file=sys.stderr, flush=True)
        ctx.set_advance_kernel(args.advance_kernel)
    if not args.no_mixed and args.advance_kernel == 1:
        line["mixed_fp32_lu"] = mixed_measure(ctx, one_run, cells, value, n, wr)
    if not args.no_jac_reuse and args.advance_kernel == 1 and args.recipe == "reference":
        ages = [int(a) for a in args.jac_ages.split(",") if a]
        line["jacobian_reuse"] = jac_reuse_measure(ctx, one_run, cells, value, n, args.grid, wr,
                                                   ages)
    if not args.no_legacy and args.advance_kernel == 1:
        line["legacy_local"] = legacy_measure(ctx, init, args.max_outer)
    if not args.no_cta32 and args.advance_kernel == 1:
        line["gridwide_cta32"] = cta32_measure(ctx, init, args.max_outer, cells, n, wr)
    if not args.no_lognormal and args.advance_kernel == 1 and args.grid == 64:
        line["lognormal128"] = lognormal128_measure(wr, ctx_shape)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_measure(wr)
        if adjoint_default_run is not None:
            line["adjoint_default_run"] = adjoint_default_run
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(init, dts, args.cpu_budget)
    kname = line["roofline"]["kernel"]
    if not args.no_traffic and args.advance_kernel == 1 and args.recipe == "reference":
        ctx.close()
        shape = ctx_shape if ctx_shape else (256, 0, 64, 0)
        tr = live_traffic(args.grid, args.max_outer, shape, kname)
        if tr:
            line["roofline"]["traffic"] = tr["bytes_per_dispatch"]
            line["roofline"]["traffic_passes"] = tr["passes"]
            if adv_launches:
                line["roofline"]["traffic_gbs"] = tr["bytes_per_dispatch"] / mean_launch_s / 1e9
            line["roofline"]["algorithmic_state_bytes_per_dispatch"] = n * 424
        print(f"# traffic: {tr['bytes_per_dispatch'] if tr else None} B per dispatch",
              file=sys.stderr, flush=True)
    print(json.dumps(line))


def rel_errors(a, b, floor=1e-20):
    """max / median of |a - b| / |b| over entries with |b| > floor."""
    m = np.abs(b) > floor
    r = np.abs(a - b)[m] / np.abs(b)[m]
    return (float(r.max()), float(np.median(r))) if r.size else (0.0, 0.0)


def mixed_measure(ctx, one_run, fp64_cells, fp64_value, n, wr):
    """The same run with FP32 factorisation of I - gamma h J (FP64 residual, two
    refinement sweeps, FP64 fallback): speed and error against the FP64 run."""
    ctx.set_advance_kernel(1)
    ctx.set_lu_precision(1, 2)
    info, el = one_run()
    fb = ctx.mixed_fallbacks()
    ctx.set_lu_precision(0)
    got = np.zeros(n, dtype=wr.CELL_DTYPE)
    ctx.download(got)
    out = {"run_seconds": el, "value": info["cell_advances"] / el,
           "speed_vs_fp64": (info["cell_advances"] / el) / fp64_value,
           "fallbacks": {"singular_fp32_factors": fb[0], "unconverged_refinements": fb[1]}}
    per = {}
    for k, name in enumerate(wr.SPECIES_NAMES):
        mx, md = rel_errors(got["current"]["xn"][:, k], fp64_cells["current"]["xn"][:, k])
        per[name] = [mx, md]
    out["species_rel_err_max_median"] = per
    out["T_rel_err_max_median"] = list(rel_errors(got["current"]["T"], fp64_cells["current"]["T"]))
    for f in ("accepted_steps", "rejected_steps"):
        out[f"delta_{f}"] = int(got["stats"][f].sum()) - int(fp64_cells["stats"][f].sum())
    worst = max([v[0] for k, v in per.items() if "D" not in k] + [out["T_rel_err_max_median"][0]])
    out["budget"] = "<= 1e-6 relative error in every non-deuterium species and T, >= 2x FP64 speed"
    out["meets_budget"] = bool(worst <= 1.0e-6 and out["speed_vs_fp64"] >= 2.0)
    print(f"# mixed run: {el:.3f} s, fallbacks {fb}", file=sys.stderr, flush=True)
    return out


def jac_reuse_measure(ctx, one_run, fp64_cells, fp64_value, n, grid, wr, ages):
    """The same run with the Jacobian lagged up to `age` accepted steps (J
    refreshed on rejection): speed, J evaluations, attempts, and the final
    states against the exact-J run through the reference's comparator."""
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "exact.bin")
    wr.write_final_states(fp64_cells, grid, path)
    st0 = fp64_cells["stats"]
    runs = []
    try:
        for age in ages:
            ctx.set_jacobian_reuse(age)
            info, el = one_run()
            got = np.zeros(n, dtype=wr.CELL_DTYPE)
            ctx.download(got)
            cmp = wr.compare_final_states(got, grid, path)
            st = got["stats"]
            nd = [k for k, name in enumerate(wr.SPECIES_NAMES) if "D" not in name]
            sp = max(rel_errors(got["current"]["xn"][:, k], fp64_cells["current"]["xn"][:, k])[0]
                     for k in nd)
            runs.append({
                "age_max": age, "run_seconds": el, "value": info["cell_advances"] / el,
                "speed_vs_exact_j": (info["cell_advances"] / el) / fp64_value,
                "jacobian_calls": int(st["jacobian_calls"].sum()),
                "jacobian_calls_exact_j": int(st0["jacobian_calls"].sum()),
                "attempts": int(st["internal_steps"].sum()),
                "attempts_exact_j": int(st0["internal_steps"].sum()),
                "delta_accepted_steps": int(st["accepted_steps"].sum()) - int(st0["accepted_steps"].sum()),
                "delta_rejected_steps": int(st["rejected_steps"].sum()) - int(st0["rejected_steps"].sum()),
                "outer_steps": info["completed_global_steps"],
                "reference_comparator_pass": cmp["pass"],
                "max_non_deuterium_species_rel_err": sp,
                "T_rel_err_max_median": list(rel_errors(got["current"]["T"],
                                                        fp64_cells["current"]["T"]))})
            print(f"# jacobian reuse {age}: {el:.3f} s, comparator {cmp['pass']}",
                  file=sys.stderr, flush=True)
    finally:
        ctx.set_jacobian_reuse(1)
        shutil.rmtree(tmp, ignore_errors=True)
    return {"rule": "J at a burn's first attempt and after age_max accepted steps; "
                    "refreshed when an attempt with a lagged J is rejected", "runs": runs}


def lognormal128_measure(wr, shape, steps=145, compare_steps=20):
    """The 128^3 log-normal recipe (seed 1, n_H = exp(z) cm^-3, T log-uniform in
    [100, 8000] K: substep counts per cell differ 10-100x): the queue-fed
    kernel for `steps` outer steps (about the 1e6 density growth at ~10 % per
    step), and both kernels on identical inputs for `compare_steps` steps
    (time and bit identity of the final states)."""
    n = 128 ** 3
    seed, sigma, t0, t1 = wr.RECIPES["lognormal128"]
    init = wr.make_recipe_states(n, seed, sigma, t0, t1)
    out = {"workload": f"128^3 = {n} cells, lognormal128 recipe (seed {seed}, sigma {sigma}, "
                       f"T {t0:g}-{t1:g} K), perturbations on"}
    with wr.Context(n) as ctx:
        if shape:
            ctx.set_queue_launch(*shape[:3])
            if len(shape) > 3:
                ctx.set_coop(shape[3], shape[3])
        ctx.upload(init)
        ctx.run(perturb=True, max_steps=1)  # prime
        ctx.upload(init)
        ta = time.perf_counter()
        info = ctx.run(perturb=True, max_steps=steps)
        el = time.perf_counter() - ta
        cells = np.zeros(n, dtype=wr.CELL_DTYPE)
        ctx.download(cells)
        growth = float(cells["density_driver"].max() / init["density_driver"].max())
        att = cells["stats"]["internal_steps"]
        out["queue_fed"] = {"outer_steps": info["completed_global_steps"], "run_seconds": el,
                            "cell_advances": info["cell_advances"],
                            "value": info["cell_advances"] / el,
                            "densest_cell_growth": growth,
                            "attempts_per_cell_min_median_max": [int(att.min()), int(np.median(att)),
                                                                 int(att.max())]}
        res = {}
        for kern in (1, 0):
            ctx.set_advance_kernel(kern)
            ctx.upload(init)
            ta = time.perf_counter()
            inf = ctx.run(perturb=True, max_steps=compare_steps)
            e = time.perf_counter() - ta
            c = np.zeros(n, dtype=wr.CELL_DTYPE)
            ctx.download(c)
            res[kern] = (inf, e, c)
        ctx.set_advance_kernel(1)
    mask = np.ones(216, bool)
    mask[156:160] = False
    same = bool(np.array_equal(res[1][2].view(np.uint8).reshape(n, 216)[:, mask],
                               res[0][2].view(np.uint8).reshape(n, 216)[:, mask]))
    out["identical_inputs"] = {"outer_steps": compare_steps,
                               "queue_fed_seconds": res[1][1], "plain_seconds": res[0][1],
                               "speedup": res[0][1] / res[1][1], "bit_identical": same}
    print(f"# lognormal128: {steps} steps {el:.3f} s (growth {growth:.3g}); {compare_steps} steps "
          f"queue {res[1][1]:.3f} s vs plain {res[0][1]:.3f} s, identical {same}",
          file=sys.stderr, flush=True)
    return out


def legacy_measure(ctx, init, max_outer):
    """The reference CPU build's driver policy (legacy-local: each cell steps
    with its own free-fall dt) on the same cells: one run, timed."""
    ctx.set_policy("legacy-local")
    try:
        ctx.upload(init)
        t0 = time.perf_counter()
        info = ctx.run(perturb=True, max_steps=max_outer)
        el = time.perf_counter() - t0
    finally:
        ctx.set_policy("gridwide")
    print(f"# legacy-local: {el:.3f} s, {info['completed_global_steps']} steps",
          file=sys.stderr, flush=True)
    return {"run_seconds": el, "value": info["cell_advances"] / el, "unit": "cell-advances/s",
            "outer_steps": info["completed_global_steps"], "cell_advances": info["cell_advances"],
            "status": info["status"], "stop_reason": info["stop_reason"]}


def cta32_measure(ctx, init, max_outer, ref_cells, n, wr):
    """The Mojo front end's gridwide-cta32 policy (one step controller per
    32-cell tile) on the same cells: one run, timed, and its final states
    against the gridwide-independent run's (the reference design's acceptance
    rule for the policy: the comparator's tolerances)."""
    ctx.set_policy("gridwide-cta32")
    try:
        ctx.upload(init)
        t0 = time.perf_counter()
        info = ctx.run(perturb=True, max_steps=max_outer)
        el = time.perf_counter() - t0
        cells = np.zeros(n, dtype=wr.CELL_DTYPE)
        ctx.download(cells)
    finally:
        ctx.set_policy("gridwide")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "independent.bin")
        grid = round(n ** (1.0 / 3.0))
        wr.write_final_states(ref_cells, grid, path)
        cmp = wr.compare_final_states(cells, grid, path)
    st = cells["stats"]
    print(f"# gridwide-cta32: {el:.3f} s, {info['completed_global_steps']} steps, comparator "
          f"{'PASS' if cmp['pass'] else 'FAIL'}", file=sys.stderr, flush=True)
    return {"run_seconds": el, "value": info["cell_advances"] / el, "unit": "cell-advances/s",
            "outer_steps": info["completed_global_steps"], "cell_advances": info["cell_advances"],
            "status": info["status"], "limiting_cell": info["limiting_cell"],
            "attempts": int(st["internal_steps"].sum()),
            "attempts_independent": int(ref_cells["stats"]["internal_steps"].sum()),
            "accepted_steps": int(st["accepted_steps"].sum()),
            "vs_independent_comparator": {k: cmp[k] for k in (
                "pass", "failures", "max_non_deuterium_species_rel_error",
                "max_thermodynamic_rel_error", "max_deuterium_species_rel_error")}}


def adjoint_default_run_child(timeout=600):
    """adjoint_default_run_measure in a process of its own (this script with
    --adjoint-default-run), started before this process uses the GPU; its
    result, or the error it ended with."""
    cmd = [sys.executable, os.path.abspath(__file__), "--adjoint-default-run"]
    try:
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout} s"}
    sys.stderr.write("".join(l + "\n" for l in p.stderr.splitlines() if l.startswith("#")))
    sys.stderr.flush()
    if p.returncode != 0:
        return {"error": f"exit status {p.returncode}", "stderr_tail": p.stderr[-600:]}
    return json.loads(p.stdout.strip().splitlines()[-1])["adjoint_default_run"]


def adjoint_default_run_measure(wr, grid=64, nsteps=100, hcap=4096, directions=2, delta=1e-6):
    """The discrete adjoint of the first `nsteps` outer steps of the default
    run (reference initial state, density perturbations at steps 20, 40, ...):
    forward, record and backward times, and the gradient of sum(log T) against
    central differences of the frozen-map replay along seeded directions.
    Runs only under --adjoint-default-run, in a process that has run nothing
    else (the default bench starts it as a child first): run after the other
    measurements in one process (torch's autograd among them), the first
    64^3 reverse dispatch was refused its scratch memory
    (HSA_STATUS_ERROR_OUT_OF_RESOURCES)."""
    n = grid ** 3
    init = wr.make_collapse_states(n)
    R = wr.adjoint_ntheta()
    with wr.Context(n) as ctx:
        ctx.upload(init)
        ctx.run(perturb=True, max_steps=1)  # loads the kernels
        ctx.upload(init)
        t0 = time.perf_counter()
        ctx.run(perturb=True, max_steps=nsteps)  # the production forward of the same steps
        fwd = time.perf_counter() - t0
        ctx.upload(init)
        t0 = time.perf_counter()
        ctx.adjoint_record(nsteps, np.ones(R), hcap=hcap, perturb=True)
        rec = time.perf_counter() - t0
        final = np.zeros(n, dtype=wr.CELL_DTYPE)
        ctx.download(final)
        attempts = int(final["stats"]["internal_steps"].sum())
        gN = np.zeros((18, n))
        gN[14] = 1.0 / final["current"]["T"]  # d sum(log T) / dT
        ctx.adjoint_backward(gN)  # allocates the window buffers
        t0 = time.perf_counter()
        g0, mbar = ctx.adjoint_backward(gN)
        bwd = time.perf_counter() - t0
        s0 = np.concatenate([init["current"]["xn"].T, init["current"]["T"][None],
                             init["current"]["e"][None], init["current"]["rho"][None],
                             init["density_driver"][None]])
        rng = np.random.default_rng(5)
        errs = []
        for _ in range(directions):
            v = rng.standard_normal(s0.shape) * np.abs(s0)
            v[15] = 0.0
            dth = rng.standard_normal(R)
            lp = np.sum(np.log(ctx.adjoint_replay(s0 + delta * v, np.exp(delta * dth))[14]))
            lm = np.sum(np.log(ctx.adjoint_replay(s0 - delta * v, np.exp(-delta * dth))[14]))
            fd = (lp - lm) / (2 * delta)
            an = float(np.sum(g0 * v) + mbar.sum(axis=1) @ dth)
            errs.append(abs(fd - an) / abs(an))
    print(f"# adjoint of the default run: {nsteps} steps, forward {fwd:.3f} s, record {rec:.3f} s, "
          f"backward {bwd:.3f} s, FD max rel err {max(errs):.2e}", file=sys.stderr, flush=True)
    return {"workload": f"{grid}^3 = {n} cells, reference initial state, the first {nsteps} outer "
                        "steps of the default run with its density perturbations, loss "
                        "sum(log T), 45 log-rate multipliers + 18 state fields per cell",
            "forward_seconds": fwd, "record_seconds": rec, "backward_seconds": bwd,
            "backward_over_forward": bwd / fwd, "record_plus_backward_over_forward": (rec + bwd) / fwd,
            "attempts": attempts,
            "fd_check": {"directions": directions, "delta": delta, "max_rel_err": max(errs)}}


def adjoint_measure(wr, grid=32, nsteps=20, seed=2, directions=8, delta=1e-6):
    """Discrete adjoint on the uniform16 recipe with the given seed: the
    production forward run, the recording run and the backward pass (HIP-event
    and wall times), and the gradient of sum(H2 mass fraction + log T) against
    central differences of the frozen-map replay along seeded directions."""
    import torch
    ag = importlib.import_module("warp-reactor_amd.autograd")
    n = grid ** 3
    _, sigma, t0, t1 = wr.RECIPES["uniform16"]
    init = wr.make_recipe_states(n, seed, sigma, t0, t1)
    ctx = wr.Context(n)
    ctx.upload(init)
    ctx.run(perturb=False, max_steps=1)
    # the forward pass the ratios compare with: the production path at the
    # launch shape (and cooperative setting) that is fastest on this workload
    autotune = importlib.import_module("warp-reactor_amd.autotune")
    chosen, _ = autotune.autotune(ctx, init, steps=nsteps, perturb=False,
                                  space={"coop_cells": (-1, 0)})
    fwd = float("inf")
    for _ in range(3):
        ctx.upload(init)
        tf = time.perf_counter()
        info = ctx.run(perturb=False, max_steps=nsteps)
        fwd = min(fwd, time.perf_counter() - tf)
    R = wr.adjoint_ntheta()
    ctx.upload(init)
    tr = time.perf_counter()
    ctx.adjoint_record(nsteps, np.ones(R), hcap=1024)
    rec = time.perf_counter() - tr
    rec_ms = ctx.adjoint_last_ms()
    final = np.zeros(n, dtype=wr.CELL_DTYPE)
    ctx.download(final)
    sN = torch.tensor(ag.state_of(final), dtype=torch.float64, requires_grad=True)
    loss = ag.h2_log_t_loss(sN)
    loss.backward()
    ctx.adjoint_backward(sN.grad.numpy())  # allocates the window buffers
    tb = time.perf_counter()
    g0, mbar = ctx.adjoint_backward(sN.grad.numpy())
    bwd = time.perf_counter() - tb
    bwd_ms = ctx.adjoint_last_ms()
    gth = mbar.sum(axis=1)  # theta = 0: M = 1
    s0 = ag.state_of(init)
    rng = np.random.default_rng(seed)
    errs = []
    for _ in range(directions):
        v = rng.standard_normal(s0.shape) * np.abs(s0)
        v[15] = 0.0
        dth = rng.standard_normal(R)
        lp = ag.h2_log_t_loss(torch.from_numpy(ctx.adjoint_replay(s0 + delta * v, np.exp(delta * dth))))
        lm = ag.h2_log_t_loss(torch.from_numpy(ctx.adjoint_replay(s0 - delta * v, np.exp(-delta * dth))))
        fd = (lp.item() - lm.item()) / (2 * delta)
        an = float(np.sum(g0 * v) + gth @ dth)
        errs.append(abs(fd - an) / abs(an))
    ctx.close()
    print(f"# adjoint: forward {fwd:.3f} s, record {rec:.3f} s, backward {bwd:.3f} s, "
          f"FD max rel err {max(errs):.2e}", file=sys.stderr, flush=True)
    return {"workload": f"{grid}^3 = {n} cells, uniform16 recipe seed {seed}, {nsteps} outer steps, "
                        f"loss sum(H2 mass fraction + log T), {R} log-rate multipliers + 18 state "
                        "fields per cell",
            "forward_seconds": fwd, "forward_launch_shape": {
                "threads": chosen[0], "blocks_per_cu": chosen[1], "solo_lanes": chosen[2],
                "coop_cells": chosen[3]},
            "record_seconds": rec, "backward_seconds": bwd,
            "record_kernel_ms": rec_ms, "backward_kernel_ms": bwd_ms,
            "backward_over_forward": bwd / fwd, "record_plus_backward_over_forward": (rec + bwd) / fwd,
            # the differentiable function's own forward pass is the recording run
            "backward_over_record": bwd / rec,
            "cell_advances": info["cell_advances"],
            "fd_check": {"directions": directions, "delta": delta, "max_rel_err": max(errs)}}


if __name__ == "__main__":
    main()
