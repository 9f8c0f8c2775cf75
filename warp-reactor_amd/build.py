"""Build the gfx950 library, the CLI and (for tests) the CPU oracle.

    python warp-reactor_amd/build.py          # library + CLI
    python warp-reactor_amd/build.py --all    # + oracle (and the reference CPU
                                              #   build when its source exists)

Outputs (in-tree, git-ignored):
    warp-reactor_amd/lib/libwarp_reactor.so   C ABI of include/warp_reactor.h
    warp-reactor_amd/bin/wr_reproducer        reproducer-compatible CLI
    warp-reactor_amd/csrc/_gen/chem_network.h generated network (from the IR)
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(PKG, "gen")
LIB = os.path.join(PKG, "lib", "libwarp_reactor.so")
CLI = os.path.join(PKG, "bin", "wr_reproducer")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: every product and sum is rounded separately, in the
# reference's order.  Packed-FP32 instructions are disabled for all HIP code.
# -disable-machine-licm: the advance kernel's loop body holds the straight-line
# network code; hoisting its constants out of the loop only spills them.
HIP_FLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-disable-machine-licm", "-Wno-unused-command-line-argument",
]


def run(cmd, cwd=None):
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def run_parallel(cmds):
    procs = []
    for cmd in cmds:
        print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    codes = [p.wait() for p in procs]
    for cmd, code in zip(cmds, codes):
        if code != 0:
            raise subprocess.CalledProcessError(code, cmd)


def generate():
    out = os.path.join(CSRC, "_gen", "chem_network.h")
    counts = os.path.join(CSRC, "_gen", "chem_counts.json")
    run([sys.executable, os.path.join(GEN, "emit.py"), out, counts])
    # wave-cooperative programs of the same network (wr_coop.hpp)
    run([sys.executable, os.path.join(GEN, "coop.py"), os.path.join(CSRC, "_gen", "coop_network.h"),
         os.path.join(CSRC, "_gen", "coop_stats.json")])
    # derivative code: the device subset (the CPU header is the oracle's)
    run([sys.executable, os.path.join(GEN, "ad.py"), os.path.join(CSRC, "_gen", "chem_adjoint_cpu.h"),
         os.path.join(CSRC, "_gen", "chem_adjoint.json"),
         os.path.join(CSRC, "_gen", "chem_adjoint_dev.h")])
    return out


def build_library(extra=()):
    generate()
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    io_obj = os.path.join(objdir, "wr_io.o")
    k_obj = os.path.join(objdir, "wr_kernels.o")
    a_obj = os.path.join(objdir, "wr_adjoint.o")
    n_obj = os.path.join(objdir, "wr_noalloc.o")
    run(["g++", "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas",
         "-c", "-o", io_obj, os.path.join(CSRC, "wr_io.cpp")])
    # the two HIP translation units compile concurrently
    run_parallel([[HIPCC, *HIP_FLAGS, *extra, "-c", "-o", k_obj, os.path.join(CSRC, "wr_kernels.hip")],
                  [HIPCC, *HIP_FLAGS, *extra, "-c", "-o", a_obj, os.path.join(CSRC, "wr_adjoint.hip")],
                  [HIPCC, *HIP_FLAGS, *extra, "-c", "-o", n_obj, os.path.join(CSRC, "wr_noalloc.hip")]])
    tmp = LIB + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, k_obj, a_obj, n_obj, io_obj])
    os.replace(tmp, LIB)
    return LIB


def network_sha256():
    """The sha256 of the network source the committed IR was imported from."""
    with open(os.path.join(GEN, "primordial_network.ir")) as f:
        for line in f:
            if line.startswith("# source-sha256 "):
                return line.split()[2]
    return "unknown"


def build_cli():
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    tmp = CLI + ".tmp"
    run(["g++", "-std=c++17", "-O2", "-Wall", f'-DWR_NETWORK_SHA256="{network_sha256()}"',
         "-o", tmp, os.path.join(CSRC, "wr_cli.cpp"),
         f"-I{os.path.join(REPO, 'include')}", f"-L{os.path.dirname(LIB)}", "-lwarp_reactor",
         "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle(with_reference=True):
    run(["make", "-C", os.path.join(REPO, "oracle"), "-j4"])
    ref_src = os.path.join(os.environ.get("WR_REFERENCE", "/root/reference"), "reproducer.cpp")
    if with_reference and os.path.exists(ref_src):
        run(["make", "-C", os.path.join(REPO, "oracle"), "-j4", "ref",
             f"REF={os.path.dirname(ref_src)}"])


def main(argv):
    build_library()
    build_cli()
    if "--all" in argv:
        build_oracle()


if __name__ == "__main__":
    main(sys.argv[1:])
