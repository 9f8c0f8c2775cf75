"""The wave-cooperative network programs (warp-reactor_amd/gen/coop.py)
against the plain generated network, on the CPU.

The generator spreads the fused J + f DAG (and the f DAG) over the 64 lanes of
a wave: vector steps with per-lane LDS addresses from a table, uniform
register code for thin chains, exact rebalancing of boolean and select
chains.  Its emitted header is compiled here as C (each vector statement a
loop over the lanes, a byte array as the LDS region) together with the plain
code of gen/emit.py, both with this build's libm, and the two must agree bit
for bit on states along collapse trajectories and at random states over the
temperature branches of the rate fits.  The scheduler's own model (simulate)
is checked against the DAG evaluation the same way.
"""
import ctypes
import math
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(REPO, "warp-reactor_amd", "gen")
sys.path.insert(0, GEN)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = tmp_path_factory.mktemp("coop")
    subprocess.run([sys.executable, os.path.join(GEN, "emit.py"), str(out / "chem_network.h")],
                   check=True, capture_output=True)
    subprocess.run([sys.executable, os.path.join(GEN, "coop.py"), str(out / "coop_network.h")],
                   check=True, capture_output=True)
    so = out / "libcoop.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    "-fno-strict-aliasing", "-mfma", "-std=c11", f"-I{out}",
                    f"-I{os.path.join(REPO, 'warp-reactor_amd', 'csrc')}",
                    "-o", str(so), os.path.join(REPO, "tests", "coop_harness.c"), "-lm"],
                   check=True)
    lib = ctypes.CDLL(str(so))
    vp, d = ctypes.c_void_p, ctypes.c_double
    for name in ("coop_eval_fused", "plain_eval_fused"):
        getattr(lib, name).argtypes = [d, vp, vp, vp]
    for name in ("coop_eval_rhs", "plain_eval_rhs"):
        getattr(lib, name).argtypes = [d, vp, vp]
    return lib


def _states():
    """(T, X) along collapse trajectories of the reference recipe and of the
    log-normal recipe, plus random states over 20 K - 1e5 K."""
    import importlib
    import oracle_lib
    wr = importlib.import_module("warp-reactor_amd")
    orc = oracle_lib.OracleWrMath()
    out = []
    cells = orc.make_cells(4, wr.CELL_DTYPE)
    for _ in range(8):
        for c in cells:
            out.append((float(c["current"]["T"]), np.maximum(c["current"]["xn"], 1e-100)))
        orc.run_gridwide(cells, True, 60)
    rec = orc.make_recipe_cells(24, wr.CELL_DTYPE, 1, 1.0, 100.0, 8000.0)
    for c in rec:
        out.append((float(c["current"]["T"]), np.maximum(c["current"]["xn"], 1e-100)))
    rng = np.random.default_rng(11)
    for _ in range(64):
        out.append((float(10 ** rng.uniform(1.3, 5.0)), 10 ** rng.uniform(-30, 1, 14)))
    return out


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _same(a, b):
    """Bit for bit, except that any two NaNs match: the cooperative code
    computes a - b as a + (-b), which the gfx950 plain code does too (there is
    no FP64 subtract; fsub is an add with a negated operand), while the C
    build of the plain code uses x86 subsd, which can give a NaN of the other
    sign.  NaNs appear only at far-off random states."""
    ua, ub = a.view(np.uint64), b.view(np.uint64)
    return np.all((ua == ub) | (np.isnan(a) & np.isnan(b)))


def test_coop_fused_is_bit_identical(harness):
    n_bad, n_nan, n_trajectory_nan = 0, 0, 0
    states = _states()
    for k, (T, X) in enumerate(states):
        X = np.ascontiguousarray(X, dtype=np.float64)
        f1, J1, f2, J2 = np.zeros(15), np.zeros(225), np.zeros(15), np.zeros(225)
        harness.coop_eval_fused(T, _p(X), _p(f1), _p(J1))
        harness.plain_eval_fused(T, _p(X), _p(f2), _p(J2))
        n_bad += int(not _same(f1, f2)) + int(not _same(J1, J2))
        nan = bool(np.isnan(f2).any() or np.isnan(J2).any())
        n_nan += nan
        n_trajectory_nan += nan and k < len(states) - 64
    print(f"{len(states)} states, {n_nan} with NaN outputs ({n_trajectory_nan} on trajectories)")
    assert n_bad == 0
    assert n_trajectory_nan == 0


def test_coop_rhs_is_bit_identical(harness):
    n_bad = 0
    for T, X in _states():
        X = np.ascontiguousarray(X, dtype=np.float64)
        f1, f2 = np.zeros(15), np.zeros(15)
        harness.coop_eval_rhs(T, _p(X), _p(f1))
        harness.plain_eval_rhs(T, _p(X), _p(f2))
        n_bad += int(not _same(f1, f2))
    assert n_bad == 0


def test_schedule_model_matches_dag():
    """The scheduler's item list, slots and tables, interpreted in Python,
    against the DAG evaluated in source order (same exp/log on both sides)."""
    import coop
    import dag as dagmod
    import emit
    d0 = emit.specialise(dagmod.read_ir(os.path.join(GEN, "primordial_network.ir")))
    progs, lay, oaddr = coop.build_programs(d0)
    exp = lambda x: math.exp(min(x, 700.0))
    log = lambda x: math.log(x) if x > 0 else (-math.inf if x == 0 else math.nan)
    f_roots = [d0.outputs[("ydot", i)] for i in range(14)] + [d0.outputs[("edot",)]]
    j_roots = [d0.outputs[("jac", r, c)] for r in range(15) for c in range(15)]
    rng = np.random.default_rng(5)
    for _ in range(3):
        T = float(10 ** rng.uniform(1.5, 4.5))
        X = list(10 ** rng.uniform(-12, 0, 14))
        vals = dagmod.evaluate(d0, T, X, exp=exp, log=log)
        for prog in progs:
            mem = coop.simulate(prog, lay, oaddr, T, X, exp, log, math.sqrt)
            keys = [("f0" if prog.name == "fused" else "dy", i) for i in range(15)]
            if prog.name == "fused":
                keys += [("J", k) for k in range(225)]
            for key, r in zip(keys, f_roots + j_roots):
                got, want = mem[oaddr(key)][1], vals[r]
                assert (got != got and want != want) or \
                    struct.pack("<d", got) == struct.pack("<d", want), (prog.name, key)


def test_schedule_respects_the_pipelining_rule():
    """A vector step never reads a value produced by the item right before it
    (its loads are issued before that item's store)."""
    import coop
    import dag as dagmod
    import emit
    d0 = emit.specialise(dagmod.read_ir(os.path.join(GEN, "primordial_network.ir")))
    progs, _, _ = coop.build_programs(d0)
    for prog in progs:
        for k, it in enumerate(prog.items):
            if it[0] != "v":
                continue
            for n in it[2]:
                for a in prog.d.nodes[n].args:
                    if a in prog.where:
                        assert prog.where[a][1] <= k - 2, (prog.name, k, n, a)


def test_programs_stay_inside_the_lds_region():
    """Every LDS address the programs use (table words, uniform loads and
    stores) is 8-byte aligned and inside the region the kernel declares."""
    import coop
    import dag as dagmod
    import emit
    d0 = emit.specialise(dagmod.read_ir(os.path.join(GEN, "primordial_network.ir")))
    progs, lay, _ = coop.build_programs(d0)
    for prog in progs:
        for row in prog.table:
            for w in row:
                a = w & coop.ADDR_MASK
                assert a % 8 == 0 and a + 8 <= lay.total
        for n in prog.in_lds:
            a = prog.a_of(n)
            assert a % 8 == 0 and a + 8 <= lay.total
    assert lay.total <= 64 * 1024  # one cooperative wave per block, 160 KiB per CU


def test_lds_slots_hold_one_type(tmp_path):
    """No LDS slot is read or written both as a double and as a 32-bit
    boolean word by the emitted programs.  Where a slot's last load of one
    type was followed by a store of the other type (a slot reused across
    types), the GPU compiler's type-based alias analysis took the two accesses
    as independent and was free to move the store above the load; the slot
    pools are separate so that no such pair exists."""
    subprocess.run([sys.executable, os.path.join(GEN, "emit.py"), str(tmp_path / "chem_network.h")],
                   check=True, capture_output=True)
    subprocess.run([sys.executable, os.path.join(GEN, "coop.py"), str(tmp_path / "coop_network.h")],
                   check=True, capture_output=True)
    so = tmp_path / "libcoop_typed.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    "-fno-strict-aliasing", "-mfma", "-std=c11", "-DCOOP_TYPE_CHECK",
                    f"-I{tmp_path}", f"-I{os.path.join(REPO, 'warp-reactor_amd', 'csrc')}",
                    "-o", str(so), os.path.join(REPO, "tests", "coop_harness.c"), "-lm"],
                   check=True)
    lib = ctypes.CDLL(str(so))
    vp, d = ctypes.c_void_p, ctypes.c_double
    lib.coop_eval_fused.argtypes = [d, vp, vp, vp]
    lib.coop_eval_rhs.argtypes = [d, vp, vp]
    lib.coop_type_conflicts.restype = ctypes.c_long
    rng = np.random.default_rng(5)
    f, J = np.zeros(15), np.zeros(225)
    for _ in range(4):
        X = 10 ** rng.uniform(-12, 0, 14)
        T = float(10 ** rng.uniform(1.5, 4.5))
        lib.coop_eval_fused(T, _p(X), _p(f), _p(J))
        lib.coop_eval_rhs(T, _p(X), _p(f))
    assert lib.coop_type_conflicts() == 0


@pytest.fixture(scope="module")
def emulation(tmp_path_factory):
    out = tmp_path_factory.mktemp("coop_emul")
    subprocess.run([sys.executable, os.path.join(GEN, "coop.py"), str(out / "coop_network.h")],
                   check=True, capture_output=True)
    so = out / "libemul.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    "-fno-strict-aliasing", "-mfma", "-std=c++17", f"-I{out}",
                    f"-I{os.path.join(REPO, 'warp-reactor_amd', 'csrc')}",
                    "-o", str(so), os.path.join(REPO, "tests", "coop_emul.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    lib.coop_emul_advance.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int]
    lib.coop_emul_advance.restype = ctypes.c_int
    return lib


def test_cooperative_advance_emulated_matches_oracle(emulation):
    """The hand-written cooperative advance (wr_coop.hpp: wave-level LU with
    row labels, lane-per-row solves, the controller), emulated lane by lane on
    the CPU with the generated programs, advances cells exactly as the oracle
    does: 27 cells of the reference recipe through 400 outer steps of their
    own gridwide run (perturbations included) and 16 log-normal recipe cells
    (T 100-8000 K) through 30 steps, compared after every advance, all seven
    ROS2S counters included.  (The same check over the 64 cells that take the
    most attempts in the 64^3 run, through all its 1000 steps, was run while
    building this path and found no difference.)"""
    import importlib
    import oracle_lib
    wr = importlib.import_module("warp-reactor_amd")
    orc = oracle_lib.OracleWrMath()
    mask = np.ones(216, bool)
    mask[156:160] = False
    for cells, steps in ((orc.make_cells(27, wr.CELL_DTYPE), 400),
                         (orc.make_recipe_cells(16, wr.CELL_DTYPE, 1, 1.0, 100.0, 8000.0), 30)):
        t, advances = 0.0, 0
        for s in range(steps):
            dts = []
            for c in range(len(cells)):
                dt, f = orc.prepare_cell(cells[c:c + 1], c, s, t, s, True)
                assert f == 1
                dts.append(dt)
            dt = min(dts)
            for c in range(len(cells)):
                one = cells[c:c + 1]
                cp = one.copy()
                r1 = emulation.coop_emul_advance(cp.ctypes.data, s, t + dt, dt, 0)
                r2, f = orc.advance_cell(one, s, t + dt, dt)
                assert f == 1 and r1 == r2
                advances += r2
                assert np.array_equal(cp.view(np.uint8).reshape(1, 216)[:, mask],
                                      one.view(np.uint8).reshape(1, 216)[:, mask]), (s, c)
            t = t + dt
        assert advances > 0
