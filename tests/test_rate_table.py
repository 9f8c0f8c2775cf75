"""Rate-table study on the CPU oracle (oracle/rate_table.c, gen/table.py).

The network split at its temperature-only frontier must reproduce the
generated network bit for bit; the cubic table must be exact at its nodes and
converge as h^4 between them (which a wrong interpolation weight breaks); a
table run must take the table path and stay close to the exact run.  The
full-size numbers (the ≤1e-6 budget gate) come from tests/rate_table_study.py.
"""
import importlib
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle_lib  # noqa: E402

wr = importlib.import_module("warp-reactor_amd")


@pytest.fixture(scope="module")
def orc():
    o = oracle_lib.OracleTable()
    yield o
    o.rt_set(0)


def _run(o, init, mode, steps, **kw):
    o.rt_set(mode, **kw)
    c = init.copy()
    try:
        r, done = o.run_gridwide(c, True, steps)
    finally:
        o.rt_set(0)
    return c, r, done


def test_split_network_bit_identical(orc):
    init = wr.make_collapse_states(64)
    ex, r0, d0 = _run(orc, init, 0, 45)  # crosses the perturbations at 20 and 40
    sp, r1, d1 = _run(orc, init, 1, 45)
    assert (r0, d0) == (r1, d1) == (1, 45)
    assert ex.tobytes() == sp.tobytes()
    # and J, f on their own, on states along the run and scaled abundances
    rng = np.random.default_rng(4)
    for cell in (ex[0], ex[17], init[5]):
        y = np.concatenate([cell["current"]["xn"] * rng.uniform(0.5, 2.0, 14), [cell["current"]["e"]]])
        f0, j0 = orc.rhs(y), orc.jac(y)
        orc.rt_set(1)
        try:
            f1, j1 = orc.rhs(y), orc.jac(y)
        finally:
            orc.rt_set(0)
        assert f0.tobytes() == f1.tobytes() and j0.tobytes() == j1.tobytes()


def test_table_exact_at_nodes(orc):
    tlo, thi, n = 10.0, 1e5, 512
    orc.rt_set(2, tlo, thi, n)
    try:
        h = (np.log(thi) - np.log(tlo)) / (n - 1)
        nodes = np.exp(np.log(tlo) + h * np.arange(2, n - 3))
        nodes = nodes[(nodes > 100.0) & (nodes < 1e4)]  # the collapse run's temperatures
        err = orc.rt_errors(nodes)
    finally:
        orc.rt_set(0)
    # only the rounding of ln T onto the grid and of exp(ln v) separate a
    # node's value from its table row; the few values that change sign in the
    # table's range are interpolated in v, where that rounding is multiplied
    # by their neighbours' magnitude
    assert np.median(err) < 1e-13
    assert np.quantile(err, 0.99) < 1e-10


def test_table_converges_as_h4(orc):
    tq = np.exp(np.random.default_rng(8).uniform(np.log(300.0), np.log(3000.0), 4000))
    errs = []
    for n in (2048, 4096):
        orc.rt_set(3, 10.0, 1e5, n)
        try:
            errs.append(orc.rt_errors(tq))
        finally:
            orc.rt_set(0)
    coarse, fine = errs
    band = (coarse > 1e-11) & (coarse < 1e-5) & (fine > 0)
    assert band.sum() > 50
    ratio = np.median(coarse[band] / fine[band])
    assert 12.0 < ratio < 20.0, ratio  # 2^4 for a cubic; 2 or 4 if a weight were wrong


def test_table_run_uses_table(orc):
    init = wr.make_collapse_states(64)
    ex, _, _ = _run(orc, init, 0, 100)
    tb, r, done = _run(orc, init, 2, 100, tlo=10.0, thi=1e5, n=4096)
    assert (r, done) == (1, 100)
    nd = [k for k, name in enumerate(wr.SPECIES_NAMES) if "D" not in name]
    a, b = tb["current"]["xn"][:, nd], ex["current"]["xn"][:, nd]
    m = np.abs(b) > 1e-20
    rel = np.abs(a - b)[m] / np.abs(b)[m]
    assert rel.max() > 0.0  # the table path ran
    assert rel.max() < 1e-4


def test_rt_set_rejects_bad_arguments(orc):
    for args in ((4, 10.0, 1e5, 64), (2, 0.0, 1e5, 64), (2, 1e5, 10.0, 64), (3, 10.0, 1e5, 3)):
        with pytest.raises(ValueError):
            orc.rt_set(*args)
    orc.rt_set(0)
