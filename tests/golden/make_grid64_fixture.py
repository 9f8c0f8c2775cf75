"""Write the full-size reference-arithmetic fixture of the default run.

Runs the reference's own compiled CPU code (oracle/_ref/libwr_ref.so, built by
oracle/Makefile from reproducer.cpp where it lies) through its gridwide host
loop restated in oracle/ref_harness.cpp (run_cells_cuda's loop,
reproducer.cpp:7859-7959): 64^3 cells from make_collapse_state (7533-7546),
perturbations on, at most 1000 outer steps.  The whole final grid is packed
into the 172-byte records of write_final_states (8289-8378); the fixture keeps
the SHA-256 of that 45,088,768-byte file, the run's step count, and a strided
sample of its records (every 512th cell plus the first and last 64), which the
GPU tests compare against with the reference comparator's tolerances.

Usage: OMP_NUM_THREADS=6 python tests/golden/make_grid64_fixture.py
(about 15-25 minutes on 6 cores; test infrastructure, never run by the product).
"""
import ctypes
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)


def sample_index(n):
    return np.unique(np.concatenate([np.arange(0, n, 512), np.arange(64), np.arange(n - 64, n)]))


def main():
    import importlib
    wr = importlib.import_module("warp-reactor_amd")
    lib = ctypes.CDLL(os.path.join(REPO, "oracle", "_ref", "libwr_ref.so"))
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    lib.ref_make_collapse_state.argtypes = [vp]
    lib.ref_run_gridwide.argtypes = [vp, i32, i32, i32, ctypes.POINTER(i32)]
    lib.ref_run_gridwide.restype = i32
    lib.ref_write_final_states.argtypes = [vp, i32, i32, ctypes.c_char_p]
    lib.ref_write_final_states.restype = i32
    grid = 64
    n = grid ** 3
    cells = np.zeros(n, dtype=wr.CELL_DTYPE)
    base, item = cells.ctypes.data, cells.itemsize
    for c in range(n):
        lib.ref_make_collapse_state(vp(base + c * item))
    cells["reserved_"] = 0
    done = i32(0)
    t0 = time.time()
    status = lib.ref_run_gridwide(vp(base), n, 1, 1000, ctypes.byref(done))
    secs = time.time() - t0
    print(f"reference gridwide run: status {status}, {done.value} steps, {secs:.0f} s", flush=True)
    path = os.path.join(HERE, "_grid64_reference.bin")
    if os.path.exists(path):
        os.remove(path)
    assert lib.ref_write_final_states(vp(base), n, grid, path.encode()) == 1
    raw = open(path, "rb").read()
    os.remove(path)
    assert len(raw) == n * 172
    idx = sample_index(n)
    rec = np.frombuffer(raw, dtype=np.uint8).reshape(n, 172)[idx]
    np.savez_compressed(os.path.join(HERE, "grid64_reference.npz"), index=idx.astype(np.int32),
                        records=rec, sha256=np.array(hashlib.sha256(raw).hexdigest()),
                        status=np.int32(status), completed_global_steps=np.int32(done.value),
                        threads=np.int32(int(os.environ.get("OMP_NUM_THREADS", "0"))),
                        seconds=np.float64(secs))
    print("sha256", hashlib.sha256(raw).hexdigest(), "sample", len(idx))


if __name__ == "__main__":
    main()
