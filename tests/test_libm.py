"""Accuracy of this build's exp / log / cbrt (warp-reactor_amd/csrc/wr_math.h).

The GPU path evaluates the network with wr_exp / wr_log and the controller
with wr_cbrt; the oracle it is compared with bit for bit is built from the same
source (oracle/_build/liboracle_wrmath.so, whose or_libm_* entry points run
exactly those functions).  Nothing in that bit-identity says how far the
functions are from the true values, so they are pinned here against
correctly rounded results (mpmath at 200 bits, rounded to nearest double):

* the arguments are the network's own: every exp and log argument the fused
  RHS + Jacobian computes (untaken select branches included, as the
  branch-free device code evaluates them) at states taken along collapse
  trajectories of the reference recipe and of the log-normal 128^3 recipe
  (T from 100 to 8000 K), >= 1e5 of each;
* exp and log must be within one ulp everywhere and correctly rounded on at
  least 90 % of the arguments; cbrt (the controller's cbrt(err) and
  cbrt(err^2/erracc), err from 1e-12 to 1e6) within one ulp and correctly
  rounded on at least 99.9 %;
* special cases (NaN, infinities, zeros, negative arguments, subnormals, the
  overflow and underflow cut-offs of exp) are checked one by one.

The report against the host libm (glibc, which the reference's CPU build
calls) follows the method of the reference's tools/transcendental_probe.py:84-201
(ulp distance on ordered bit patterns) and is printed by
test_ulp_report_against_glibc.
"""
import ctypes
import ctypes.util
import math
import os
import struct
import sys

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "warp-reactor_amd", "gen"))


def _ordered(x):
    b = np.asarray(x, dtype=np.float64).view(np.uint64).astype(object)
    out = []
    for v in b:
        v = int(v)
        out.append((~v) & 0xFFFFFFFFFFFFFFFF if v >> 63 else v | (1 << 63))
    return np.array(out, dtype=object)


def ulp_distance(a, b):
    """Distance in units in the last place between doubles (ordered bits)."""
    da, db = _ordered(a), _ordered(b)
    return np.array([abs(int(x) - int(y)) for x, y in zip(da, db)], dtype=np.int64)


def correctly_rounded(fn, xs):
    mp = mpmath.mp
    old = mp.prec
    mp.prec = 200
    try:
        return np.array([float(fn(mpmath.mpf(float(x)))) for x in xs])
    finally:
        mp.prec = old


@pytest.fixture(scope="module")
def network_args():
    """exp and log arguments of the fused network along collapse trajectories."""
    import importlib
    import dag as dagmod
    import emit
    import oracle_lib

    wr = importlib.import_module("warp-reactor_amd")
    d = emit.specialise(dagmod.read_ir(os.path.join(REPO, "warp-reactor_amd", "gen",
                                                    "primordial_network.ir")))
    orc = oracle_lib.OracleWrMath()
    states = []
    # reference recipe: 8 cells stepped through the gridwide loop, 25 steps at a time
    cells = orc.make_cells(8, wr.CELL_DTYPE)
    for _ in range(16):
        for c in cells:
            states.append((float(c["current"]["T"]), np.array(c["current"]["xn"], dtype=float)))
        orc.run_gridwide(cells, True, 25)
    # log-normal 128^3 recipe (n_H spread e^+-3, T log-uniform 100-8000 K): initial
    # states and states after a few steps
    rec = orc.make_recipe_cells(160, wr.CELL_DTYPE, 1, 1.0, 100.0, 8000.0)
    for k in range(2):
        for c in rec:
            states.append((float(c["current"]["T"]), np.array(c["current"]["xn"], dtype=float)))
        orc.run_gridwide(rec, False, 3)
    args = {"exp": [], "log": []}

    def rec_exp(x):
        args["exp"].append(x)
        return math.exp(min(x, 700.0))

    def rec_log(x):
        args["log"].append(x)
        return math.log(x) if x > 0 else (-math.inf if x == 0 else math.nan)

    for T, xn in states:
        X = np.maximum(xn, 1.0e-100)
        dagmod.evaluate(d, T, list(X), exp=rec_exp, log=rec_log)
    return {k: np.array(v) for k, v in args.items()}, len(states)


def _wr(orc, op, xs):
    return orc.libm(op, xs)


def test_network_argument_sample_is_large(network_args):
    args, nstates = network_args
    print(f"states {nstates}: exp args {len(args['exp'])}, log args {len(args['log'])}")
    assert len(args["exp"]) >= 100_000
    assert len(args["log"]) >= 25_000


@pytest.mark.parametrize("op", ["exp", "log"])
def test_within_one_ulp_of_correctly_rounded(network_args, oracle_wr, op):
    args, _ = network_args
    xs = args[op]
    if op == "log":
        # log of the network is also taken of its many distinct rate arguments;
        # pad the sample to 1e5 with the neighbourhoods of the recorded ones
        extra = np.nextafter(np.repeat(xs, 3), np.inf)
        xs = np.concatenate([xs, extra])[:max(100_000, len(xs))]
    xs = xs[np.isfinite(xs)]
    if op == "log":
        xs = xs[xs > 0]
    else:
        xs = xs[(xs > -745.0) & (xs < 709.0)]
    got = _wr(oracle_wr, op, xs)
    want = correctly_rounded(mpmath.exp if op == "exp" else mpmath.log, xs)
    dist = ulp_distance(got, want)
    cr = float(np.mean(dist == 0))
    print(f"wr_{op}: {len(xs)} network arguments, max {dist.max()} ulp, correctly rounded {cr:.4f}")
    assert dist.max() <= 1
    assert cr >= 0.90


def test_cbrt_within_one_ulp(oracle_wr):
    rng = np.random.default_rng(7)
    err = 10.0 ** rng.uniform(-12, 6, 60_000)
    erracc = np.maximum(1.0e-2, 10.0 ** rng.uniform(-3, 0, 60_000))
    xs = np.concatenate([err, err * err / erracc])
    got = _wr(oracle_wr, "cbrt", xs)
    want = correctly_rounded(mpmath.cbrt, xs)
    dist = ulp_distance(got, want)
    cr = float(np.mean(dist == 0))
    print(f"wr_cbrt: {len(xs)} controller arguments, max {dist.max()} ulp, correctly rounded {cr:.5f}")
    assert dist.max() <= 1
    assert cr >= 0.999


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def test_special_cases(oracle_wr):
    e = lambda x: oracle_wr.libm("exp", [x])[0]
    lg = lambda x: oracle_wr.libm("log", [x])[0]
    cb = lambda x: oracle_wr.libm("cbrt", [x])[0]
    inf, nan = math.inf, math.nan
    assert math.isnan(e(nan)) and e(inf) == inf and e(-inf) == 0.0
    assert e(0.0) == 1.0 and e(-0.0) == 1.0
    # overflow / underflow cut-offs (ln DBL_MAX = 709.78..., ln of half the least subnormal)
    assert e(709.78) < inf and e(709.79) == inf and e(1000.0) == inf
    assert e(-745.13) > 0.0 and e(-745.14) == 0.0 and e(-1000.0) == 0.0
    assert e(-740.0) == correctly_rounded(mpmath.exp, [-740.0])[0]  # subnormal result
    assert math.isnan(lg(nan)) and lg(inf) == inf and lg(0.0) == -inf and lg(-0.0) == -inf
    assert math.isnan(lg(-1.0)) and math.isnan(lg(-inf))
    assert lg(1.0) == 0.0 and _bits(lg(1.0)) == 0
    for x in (5e-324, 1e-310, 2.2250738585072014e-308):  # subnormals and the least normal
        assert ulp_distance([lg(x)], correctly_rounded(mpmath.log, [x]))[0] <= 1
    assert cb(0.0) == 0.0 and _bits(cb(-0.0)) == _bits(-0.0)
    assert cb(inf) == inf and cb(-inf) == -inf and math.isnan(cb(nan))
    assert cb(-8.0) == -2.0 and cb(27.0) == 3.0 and cb(1e-300) == correctly_rounded(mpmath.cbrt, [1e-300])[0]
    assert ulp_distance([cb(5e-324)], correctly_rounded(mpmath.cbrt, [5e-324]))[0] <= 1


def test_ulp_report_against_glibc(network_args, oracle_wr):
    """ULP distance of wr_exp / wr_log / wr_cbrt from the host libm the
    reference's CPU build calls (glibc), on the same arguments."""
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.cbrt.restype = ctypes.c_double
    libm.cbrt.argtypes = [ctypes.c_double]
    args, _ = network_args
    report = {}
    for op, host in (("exp", math.exp), ("log", math.log)):
        xs = args[op][np.isfinite(args[op])]
        xs = xs[(xs > 0)] if op == "log" else xs[(xs > -745.0) & (xs < 709.0)]
        got = _wr(oracle_wr, op, xs)
        ref = np.array([host(float(x)) for x in xs])
        dist = ulp_distance(got, ref)
        report[op] = (len(xs), float(np.mean(dist == 0)), int(dist.max()))
    xs = 10.0 ** np.random.default_rng(3).uniform(-12, 6, 50_000)
    got = _wr(oracle_wr, "cbrt", xs)
    ref = np.array([libm.cbrt(float(x)) for x in xs])
    dist = ulp_distance(got, ref)
    report["cbrt"] = (len(xs), float(np.mean(dist == 0)), int(dist.max()))
    for op, (n, same, mx) in report.items():
        print(f"wr_{op} vs glibc {op}: {n} arguments, identical {same:.4f}, max {mx} ulp")
    # both sides are within about one ulp of the true value
    assert report["exp"][2] <= 2 and report["log"][2] <= 2 and report["cbrt"][2] <= 4
