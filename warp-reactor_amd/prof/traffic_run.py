"""The program bench.py runs under `rocprofv3 --pmc FETCH_SIZE` and
`--pmc WRITE_SIZE` to measure the advance kernel's HBM traffic in the same
build and launch shape as the line it reports:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o t -- \
        python3 warp-reactor_amd/prof/traffic_run.py --grid 64 --max-outer 1000 \
            --threads 256 --per-cu 0 --solo 0

First three device-to-device copies of 1 GiB (hipMemcpy: the runtime's copy
kernel, 3 GiB read and 3 GiB written over its dispatches) calibrate what the
counters read for a known number of bytes; then one collapse run of the workload on
the library (make_collapse_state cells, perturbations, the given outer steps,
the queue-fed advance kernel with the given launch shape, HIP graphs).  With
--timed (not under the profiler) one outer step loads the kernels first and
the run's wall time is printed.  Needs no torch."""
import argparse
import ctypes
import importlib
import os
import sys
import time

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

COPY_BYTES = 1 << 30


def calibration_copies(n=3):
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), COPY_BYTES) == 0
    assert hip.hipMalloc(ctypes.byref(b), COPY_BYTES) == 0
    assert hip.hipMemset(a, 1, COPY_BYTES) == 0
    for _ in range(n):
        assert hip.hipMemcpy(b, a, COPY_BYTES, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"calibration: {n} copies of {COPY_BYTES} bytes", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--max-outer", type=int, default=1000)
    ap.add_argument("--threads", type=int, default=256)
    ap.add_argument("--per-cu", type=int, default=0)
    ap.add_argument("--solo", type=int, default=0)
    ap.add_argument("--coop", type=int, default=-1)
    ap.add_argument("--timed", action="store_true")
    args = ap.parse_args()
    if not args.timed:
        calibration_copies()
    wr = importlib.import_module("warp-reactor_amd")
    n = args.grid ** 3
    init = wr.make_collapse_states(n)
    with wr.Context(n) as ctx:
        ctx.set_queue_launch(args.threads, args.per_cu, args.solo)
        if args.coop >= 0:
            ctx.set_coop(args.coop, args.coop)
        if args.timed:
            ctx.upload(init)
            ctx.run(perturb=True, max_steps=1)  # loads the kernels
        ctx.upload(init)
        t0 = time.perf_counter()
        info = ctx.run(perturb=True, max_steps=args.max_outer)
        el = time.perf_counter() - t0
        lanes = ctx.queue_launch()[3] * args.threads
    print(f"run: {info['completed_global_steps']} outer steps, {info['cell_advances']} "
          f"cell-advances, status {info['status']}, {el:.3f} s, {lanes} resident lanes", flush=True)


if __name__ == "__main__":
    main()
