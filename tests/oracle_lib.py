"""Test-side loaders of the CPU oracle (oracle/_build/liboracle.so) and of the
reference's own CPU code (oracle/_ref/libwr_ref.so).  Both are checkers; the
product never loads them."""
import ctypes
import json
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "liboracle.so")
ORACLE_WR_SO = os.path.join(REPO, "oracle", "_build", "liboracle_wrmath.so")
ORACLE_TABLE_SO = os.path.join(REPO, "oracle", "_build", "liboracle_table.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libwr_ref.so")
REF_CLI = os.path.join(REPO, "oracle", "_ref", "reproducer_cpu")

_VP, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def _p(a):
    return a.ctypes.data_as(_VP)


class _Cpu:
    prefix = ""

    def __init__(self, path):
        self.lib = ctypes.CDLL(path)
        p = self.prefix
        sig = {
            "rhs": (None, [_VP, _VP]), "jac": (None, [_VP, _VP]),
            "temperature_from_y": (_D, [_VP]),
            "lu": (_I, [_VP, _VP]), "lu_solve": (None, [_VP, _VP, _VP]),
            "integrate": (_I, [_VP, _D, _D, _I, _VP, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
            "burn": (_I, [_VP, _D, _VP]),
            "make_collapse_state": (None, [_VP]),
            "splitmix64": (ctypes.c_uint64, [ctypes.c_uint64]),
            "perturbation_factor": (_D, [_I, _I]),
            "apply_perturbation": (None, [_VP, _I, _I, _I]),
            "collapse_timestep": (_D, [_VP]),
            "post_step": (None, [_VP]),
            "eos_rt": (None, [_VP]),
            "prepare_cell": (_D, [_VP, _I, _I, _D, _I, _I, ctypes.POINTER(_I)]),
            "advance_cell": (_I, [_VP, _I, _D, _D, ctypes.POINTER(_I)]),
            "run_gridwide": (_I, [_VP, _I, _I, _I, ctypes.POINTER(_I)]),
        }
        if p == "or_":
            sig.update({"set_integrator_max_steps": (None, [_I]),
                        "set_lu_precision": (None, [_I, _I]),
                        "set_jacobian_reuse": (None, [_I]),
                        "run_legacy": (_I, [_VP, _I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
                        "run_cta": (_I, [_VP, _I, _I, _I, _I, _I, ctypes.POINTER(_I),
                                         ctypes.POINTER(_I)]),
                        "mixed_fallbacks": (None, [_VP, _VP]),
                        "reset_mixed_fallbacks": (None, []),
                        "lu32": (_I, [_VP, _VP]), "lu_solve32": (None, [_VP, _VP, _VP]),
                        "solve_mixed": (_I, [_VP, _D, _VP, _I, _VP]),
                        "mixed_lu": (_I, [_VP, _D, _VP, _VP, _VP]),
                        "make_recipe_state": (None, [_VP, _I, ctypes.c_uint64, _D, _D, _D]),
                        "libm_exp": (_D, [_D]), "libm_log": (_D, [_D]), "libm_cbrt": (_D, [_D]),
                        "prepare_grid": (None, [_VP, _I, _I, _D, _I, _I, ctypes.POINTER(_D),
                                                ctypes.POINTER(_I), ctypes.POINTER(_I)]),
                        "replay_cells": (None, [_VP, _VP, _I, _VP, _I, _I, _I, _D, _VP])})
        for name, (res, args) in sig.items():
            fn = getattr(self.lib, p + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, "_" + name, fn)

    def rhs(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(15)
        self._rhs(_p(y), _p(out))
        return out

    def jac(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros((15, 15))
        self._jac(_p(y), _p(out))
        return out

    def temperature_from_y(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        return self._temperature_from_y(_p(y))

    def lu(self, a):
        a = np.array(a, dtype=np.float64).reshape(15, 15).copy()
        ip = np.zeros(15, dtype=np.int32)
        info = self._lu(_p(a), _p(ip))
        return a, ip, info

    def lu_solve(self, lu, ip, x):
        lu = np.ascontiguousarray(lu, dtype=np.float64)
        ip = np.ascontiguousarray(ip, dtype=np.int32)
        x = np.array(x, dtype=np.float64).copy()
        self._lu_solve(_p(lu), _p(ip), _p(x))
        return x

    def integrate(self, y, tout, dt0, max_steps=0):
        y = np.array(y, dtype=np.float64).copy()
        cnt = np.zeros(7, dtype=np.int32)
        te, de = _D(), _D()
        r = self._integrate(_p(y), tout, dt0, max_steps, _p(cnt), ctypes.byref(te), ctypes.byref(de))
        return {"y": y, "counters": cnt, "t_end": te.value, "dt_end": de.value, "result": r}

    def burn(self, burn, dt, stats):
        burn = np.array(burn, dtype=np.float64).copy()
        stats = np.array(stats, dtype=np.uint64).copy()
        r = self._burn(_p(burn), dt, _p(stats))
        return r, burn, stats

    def make_cells(self, n, dtype):
        cells = np.zeros(n, dtype=dtype)
        for i in range(n):
            self._make_collapse_state(ctypes.c_void_p(cells.ctypes.data + i * dtype.itemsize))
        cells["reserved_"] = 0
        return cells

    def replay_cells(self, cells, index, dt_hist, perturb=True, completed0=0, time0=0.0):
        """Replay cells (their initial states in `cells`, grid indices `index`)
        through the given dt_grid sequence; returns per-cell status codes."""
        index = np.ascontiguousarray(index, dtype=np.int32)
        dt_hist = np.ascontiguousarray(dt_hist, dtype=np.float64)
        status = np.zeros(len(cells), dtype=np.int32)
        self._replay_cells(_p(cells), _p(index), len(cells), _p(dt_hist), len(dt_hist),
                           int(perturb), completed0, time0, _p(status))
        cells["reserved_"] = 0
        return status

    def burn_to_y(self, burn):
        """y = (xn, e) with e from eos_rt, as burn_ros2s builds it (7595-7606)."""
        b = np.array(burn, dtype=np.float64).copy()
        self._eos_rt(_p(b))
        return np.concatenate([b[3:], [b[2]]])

    def splitmix64(self, v):
        return self._splitmix64(v)

    def perturbation_factor(self, cell, step):
        return self._perturbation_factor(cell, step)

    def collapse_timestep(self, cell):
        return self._collapse_timestep(_p(cell))

    def prepare_cell(self, cell, index, completed, grid_time, step, perturb):
        f = _I(1)
        dt = self._prepare_cell(_p(cell), index, completed, grid_time, step, int(perturb),
                                ctypes.byref(f))
        return dt, f.value

    def advance_cell(self, cell, completed, next_time, dt_grid):
        f = _I(1)
        r = self._advance_cell(_p(cell), completed, next_time, dt_grid, ctypes.byref(f))
        return r, f.value

    def make_recipe_cells(self, n, dtype, seed, sigma, t_min, t_max):
        cells = np.zeros(n, dtype=dtype)
        item = cells.itemsize
        base = cells.ctypes.data
        for i in range(n):
            self._make_recipe_state(_VP(base + i * item), i, seed, sigma, t_min, t_max)
        return cells

    def run_legacy(self, cells, perturb=True, max_steps=0):
        """The legacy-local driver (reference CPU policy); (status, completed, failed cell)."""
        done, failed = _I(0), _I(0)
        r = self._run_legacy(_p(cells), len(cells), int(perturb), max_steps, ctypes.byref(done),
                             ctypes.byref(failed))
        cells["reserved_"] = 0
        return r, done.value, failed.value

    def set_lu_precision(self, mode, sweeps=2):
        """0: FP64 factors; 1: FP32 factors + FP64 refinement (oracle global)."""
        self._set_lu_precision(int(mode), int(sweeps))
        self._reset_mixed_fallbacks()

    def lu32(self, a):
        a = np.array(a, dtype=np.float32).reshape(15, 15).copy()
        ip = np.zeros(15, dtype=np.int32)
        info = self._lu32(_p(a), _p(ip))
        return a, ip, info

    def lu_solve32(self, lu, ip, x):
        x = np.array(x, dtype=np.float32).copy()
        self._lu_solve32(_p(np.ascontiguousarray(lu, dtype=np.float32)),
                         _p(np.ascontiguousarray(ip, dtype=np.int32)), _p(x))
        return x

    def solve_mixed(self, jac, fac, y, b, sweeps=2):
        x = np.array(b, dtype=np.float64).copy()
        r = self._solve_mixed(_p(np.ascontiguousarray(jac, dtype=np.float64)), float(fac),
                              _p(np.ascontiguousarray(y, dtype=np.float64)), int(sweeps), _p(x))
        return x, r

    def mixed_lu(self, jac, fac, y):
        lu = np.zeros(225, dtype=np.float32)
        ip = np.zeros(15, dtype=np.int32)
        info = self._mixed_lu(_p(np.ascontiguousarray(jac, dtype=np.float64)), float(fac),
                              _p(np.ascontiguousarray(y, dtype=np.float64)), _p(lu), _p(ip))
        return lu, ip, info

    def mixed_fallbacks(self):
        out = np.zeros(2, dtype=np.uint64)
        self._mixed_fallbacks(_p(out[:1]), _p(out[1:]))
        return int(out[0]), int(out[1])

    def prepare_grid(self, cells, completed, grid_time, step, perturb=True):
        """prepare of one outer step on every cell (in place) and the host
        loop's minimum: (dt_grid, lowest cell holding it or -1, failure code)."""
        dt, cell, f = _D(), _I(), _I()
        self._prepare_grid(_p(cells), len(cells), completed, grid_time, step, int(perturb),
                           ctypes.byref(dt), ctypes.byref(cell), ctypes.byref(f))
        cells["reserved_"] = 0
        return dt.value, cell.value, f.value

    def run_gridwide(self, cells, perturb=True, max_steps=0, completed0=0):
        """run_cells_cuda's host loop; the loop index starts at 0 and the
        completed count at completed0, as in the reference (7791-7859)."""
        done = _I(completed0)
        r = self._run_gridwide(_p(cells), len(cells), int(perturb), max_steps, ctypes.byref(done))
        cells["reserved_"] = 0
        return r, done.value

    def run_cta(self, cells, perturb=True, max_steps=0, tile=32, completed0=0, cpp_order=False):
        """The Mojo front end's gridwide-cta32 policy (reproducer.mojo:2829-3093,
        3604-3621) with `tile` cells per step controller: (status, completed
        steps, limiting cell of the last step or -1).  cpp_order: the density
        updates in reproducer.cpp's operation order (for pinning)."""
        done, lim = _I(completed0), _I(-1)
        r = self._run_cta(_p(cells), len(cells), int(perturb), max_steps, tile, int(cpp_order),
                          ctypes.byref(done), ctypes.byref(lim))
        cells["reserved_"] = 0
        return r, done.value, lim.value


class Oracle(_Cpu):
    prefix = "or_"

    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (make -C oracle)")
        super().__init__(ORACLE_SO)


class OracleWrMath(_Cpu):
    """The oracle evaluated with the GPU path's exp/log/cbrt (wr_math.h)."""
    prefix = "or_"

    def __init__(self):
        if not os.path.exists(ORACLE_WR_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_WR_SO} (make -C oracle)")
        super().__init__(ORACLE_WR_SO)

    def libm(self, op, x):
        fn = {"exp": self._libm_exp, "log": self._libm_log, "cbrt": self._libm_cbrt}[op]
        return np.array([fn(float(v)) for v in np.ravel(x)])


class OracleTable(_Cpu):
    """The oracle with the rate-table study compiled in (oracle/rate_table.c):
    mode 0 the network as generated, 1 the network split at its T-only
    frontier with exact values, 2 the frontier values from a cubic table in
    ln T.  The mode is process-wide; reset it to 0 after use."""
    prefix = "or_"

    def __init__(self):
        if not os.path.exists(ORACLE_TABLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_TABLE_SO} (make -C oracle)")
        super().__init__(ORACLE_TABLE_SO)
        self.lib.or_rt_set.restype = _I
        self.lib.or_rt_set.argtypes = [_I, _D, _D, _I]
        self.lib.or_rt_ntv.restype = _I
        self.lib.or_rt_outside.restype = ctypes.c_longlong
        self.lib.or_rt_errors.argtypes = [_VP, _I, _VP]
        self.ntv = self.lib.or_rt_ntv()

    def rt_set(self, mode, tlo=1.0, thi=2.0, n=4):
        if self.lib.or_rt_set(mode, tlo, thi, n) != 0:
            raise ValueError(f"or_rt_set({mode}, {tlo}, {thi}, {n}) refused")

    def rt_outside(self):
        """Frontier evaluations outside the table's range since rt_set."""
        return self.lib.or_rt_outside()

    def rt_errors(self, T):
        """Per frontier value, the max relative interpolation error at T."""
        T = np.ascontiguousarray(T, dtype=np.float64)
        err = np.zeros(self.ntv)
        self.lib.or_rt_errors(_p(T), len(T), _p(err))
        return err


class Reference(_Cpu):
    prefix = "ref_"

    @classmethod
    def load(cls):
        if not os.path.exists(REF_SO):
            return None
        return cls(REF_SO)

    def run_legacy(self, cells, perturb=True):
        """The reference's own CPU driver loop (legacy-local policy)."""
        fn = self.lib.ref_run_legacy
        fn.restype = _I
        fn.argtypes = [_VP, _I, _I, ctypes.POINTER(_I)]
        done = _I(0)
        r = fn(_p(cells), len(cells), int(perturb), ctypes.byref(done))
        cells["reserved_"] = 0
        return r, done.value


class OracleAdjoint(OracleWrMath):
    """The oracle with the tangent-linear and adjoint restatement (adjoint.c),
    evaluated with the GPU path's libm."""

    def __init__(self):
        path = os.path.join(REPO, "oracle", "_build", "liboracle_adj.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (make -C oracle)")
        _Cpu.__init__(self, path)
        L = self.lib
        sig = {
            "or_adj_ntheta": (_I, []),
            "or_lu_solve_t": (None, [_VP, _VP, _VP]),
            "or_adj_npre": (_I, [_I]),
            "or_adj_record": (_I, [_VP, _I, _I, _VP, _VP, _VP, _VP, _VP, _I, _I, _VP]),
            "or_adj_replay": (_I, [_VP, _VP, _I, _I, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I]),
            "or_adj_backward": (_I, [_VP, _I, _I, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I,
                                     _VP]),
            "or_adj_f": (None, [_VP, _VP, _VP]),
            "or_adj_fj": (None, [_VP, _VP, _VP, _VP]),
            "or_adj_f_jvp": (None, [_VP, _VP, _VP, _VP, _VP]),
            "or_adj_f_vjp": (None, [_VP, _VP, _VP, _VP, _VP]),
            "or_adj_jac_jvp": (None, [_VP, _VP, _VP, _VP, _VP]),
            "or_adj_jac_vjp": (None, [_VP, _VP, _VP, _VP, _VP]),
            "or_adj_jacd_vjp": (None, [_VP, _VP, _VP, _VP, _VP]),
            "or_adj_jacd_mbar": (None, [_VP, _VP, _VP]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        self.R = L.or_adj_ntheta()
        table = os.path.join(REPO, "oracle", "_build", "chem_adjoint.json")
        with open(table) as fh:
            self._reactions = json.load(fh)["reactions"]
        if len(self._reactions) != self.R:
            raise RuntimeError(f"{table} does not describe {path} (make -C oracle)")

    def reactions(self):
        """The reaction of each multiplier (name, stoichiometry nu over the 14
        species, rate terms, energy-rate monomials), as the generator that
        wrote the derivative code recovered them."""
        return self._reactions

    def lu_solve_t(self, lu, ip, x):
        x = np.array(x, dtype=np.float64).copy()
        self.lib.or_lu_solve_t(_p(np.ascontiguousarray(lu, dtype=np.float64)),
                               _p(np.ascontiguousarray(ip, dtype=np.int32)), _p(x))
        return x

    def f(self, y, M):
        out = np.zeros(15)
        self.lib.or_adj_f(_p(np.ascontiguousarray(y, dtype=np.float64)),
                          _p(np.ascontiguousarray(M, dtype=np.float64)), _p(out))
        return out

    def fj(self, y, M):
        f, J = np.zeros(15), np.zeros(225)
        self.lib.or_adj_fj(_p(np.ascontiguousarray(y, dtype=np.float64)),
                           _p(np.ascontiguousarray(M, dtype=np.float64)), _p(f), _p(J))
        return f, J

    def f_jvp(self, y, M, dy, dM):
        out = np.zeros(15)
        self.lib.or_adj_f_jvp(*[_p(np.ascontiguousarray(a, dtype=np.float64)) for a in (y, M, dy, dM)],
                              _p(out))
        return out

    def f_vjp(self, y, M, w):
        yb, mb = np.zeros(15), np.zeros(self.R)
        self.lib.or_adj_f_vjp(*[_p(np.ascontiguousarray(a, dtype=np.float64)) for a in (y, M, w)],
                              _p(yb), _p(mb))
        return yb, mb

    def jac_jvp(self, y, M, dy, dM):
        out = np.zeros(225)
        self.lib.or_adj_jac_jvp(*[_p(np.ascontiguousarray(a, dtype=np.float64))
                                  for a in (y, M, dy, dM)], _p(out))
        return out

    def jac_vjp(self, y, M, W):
        yb, mb = np.zeros(15), np.zeros(self.R)
        self.lib.or_adj_jac_vjp(*[_p(np.ascontiguousarray(a, dtype=np.float64)) for a in (y, M, W)],
                                _p(yb), _p(mb))
        return yb, mb

    def jacd_vjp(self, y, M, W):
        """Adjoint of D(y, M) (J_theta = J - D) with seed W: (ybar, Mbar)."""
        yb, mb = np.zeros(15), np.zeros(self.R)
        self.lib.or_adj_jacd_vjp(*[_p(np.ascontiguousarray(a, dtype=np.float64)) for a in (y, M, W)],
                                 _p(yb), _p(mb))
        return yb, mb

    def jacd_mbar(self, y, W):
        """The M-adjoint of D alone (it does not depend on M)."""
        mb = np.zeros(self.R)
        self.lib.or_adj_jacd_mbar(*[_p(np.ascontiguousarray(a, dtype=np.float64)) for a in (y, W)],
                                  _p(mb))
        return mb

    def record(self, cells, nsteps, M, hcap=512, perturb=False):
        """Forward run with multipliers M (and the perturbations when
        `perturb`) recording the tape; returns a dict."""
        n = len(cells)
        M = np.ascontiguousarray(M, dtype=np.float64)
        npre = self.lib.or_adj_npre(nsteps)
        t = {"n": n, "nsteps": nsteps, "hcap": hcap, "M": M, "perturb": bool(perturb),
             "dts": np.zeros(nsteps), "ckpt": np.zeros((nsteps, 18, n)),
             "pre": np.zeros((max(npre, 1), 18, n)),
             "hoff": np.zeros((nsteps + 1, n), dtype=np.int32), "htape": np.zeros((n, hcap))}
        r = self.lib.or_adj_record(_p(cells), n, nsteps, _p(M), _p(t["dts"]), _p(t["ckpt"]),
                                   _p(t["hoff"]), _p(t["htape"]), hcap, int(perturb),
                                   _p(t["pre"]))
        cells["reserved_"] = 0
        t["status"] = r
        return t

    def replay(self, t, s0, M=None, ds0=None, dM=None):
        """Frozen map from state s0 [18, n]; with tangents when ds0 is given."""
        n, ns = t["n"], t["nsteps"]
        M = t["M"] if M is None else np.ascontiguousarray(M, dtype=np.float64)
        s0 = np.ascontiguousarray(s0, dtype=np.float64)
        sN = np.zeros((18, n))
        dsN = np.zeros((18, n)) if ds0 is not None else None
        if ds0 is not None:
            ds0 = np.ascontiguousarray(ds0, dtype=np.float64)
            dM = np.ascontiguousarray(dM if dM is not None else np.zeros(self.R), dtype=np.float64)
        r = self.lib.or_adj_replay(_p(s0), _p(ds0) if ds0 is not None else None, n, ns,
                                   _p(t["dts"]), _p(t["hoff"]), _p(t["htape"]), t["hcap"], _p(M),
                                   _p(dM) if ds0 is not None else None, _p(sN),
                                   _p(dsN) if ds0 is not None else None,
                                   int(t.get("perturb", False)))
        assert r == 1, r
        return (sN, dsN) if ds0 is not None else sN

    def backward(self, t, gN):
        n, ns = t["n"], t["nsteps"]
        g0 = np.zeros((18, n))
        mb = np.zeros((self.R, n))
        r = self.lib.or_adj_backward(_p(t["ckpt"]), n, ns, _p(t["dts"]), _p(t["hoff"]),
                                     _p(t["htape"]), t["hcap"], _p(t["M"]),
                                     _p(np.ascontiguousarray(gN, dtype=np.float64)), _p(g0), _p(mb),
                                     int(t.get("perturb", False)), _p(t["pre"]))
        assert r == 1, r
        return g0, mb
