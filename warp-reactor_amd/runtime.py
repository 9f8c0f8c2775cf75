"""ctypes binding of libwarp_reactor.so (include/warp_reactor.h).

Cell arrays are numpy structured arrays with the 216-byte CollapseState
layout (reproducer.cpp:7455-7461); packed final states use the 172-byte
PackedFinalState layout (reproducer.cpp:7488-7505).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))

NUM_SPEC = 14
NEQS = 15
SPECIES_NAMES = ("E", "Hp", "H", "Hm", "Dp", "D", "H2p", "Dm", "H2", "HDp", "HD",
                 "HEpp", "HEp", "HE")
STATS_FIELDS = ("internal_steps", "rhs_calls", "jacobian_calls", "decompositions",
                "linear_solves", "accepted_steps", "rejected_steps")

BURN_DTYPE = np.dtype([("rho", "<f8"), ("T", "<f8"), ("e", "<f8"), ("xn", "<f8", (NUM_SPEC,))])
STATS_DTYPE = np.dtype([(f, "<u8") for f in STATS_FIELDS])
CELL_DTYPE = np.dtype([("current", BURN_DTYPE), ("time", "<f8"), ("density_driver", "<f8"),
                       ("completed_steps", "<i4"), ("reserved_", "<i4"), ("stats", STATS_DTYPE)])
PACKED_DTYPE = np.dtype([("cell", "<i4"), ("i", "<i4"), ("j", "<i4"), ("k", "<i4"),
                         ("completed_steps", "<i4"), ("time", "<f8"), ("density_driver", "<f8"),
                         ("rho", "<f8"), ("T", "<f8"), ("e", "<f8"), ("xn", "<f8", (NUM_SPEC,))])
assert CELL_DTYPE.itemsize == 216 and PACKED_DTYPE.itemsize == 172

SUCCESS, BAD_INPUTS, DT_UNDERFLOW, TOO_MANY_STEPS = 1, -1, -2, -4
TOO_MUCH_ACCURACY_REQUESTED, LU_DECOMPOSITION_ERROR = -5, -7
ERR_HIP, ERR_ARGUMENT, ERR_CAPACITY, ERR_IO = -100, -101, -102, -103


class WarpReactorError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{message} (status {code})")
        self.code = code


class RunInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("stop_reason", ctypes.c_int32),
                ("completed_global_steps", ctypes.c_int32), ("failed_step", ctypes.c_int32),
                ("failed_phase", ctypes.c_int32), ("failed_cell", ctypes.c_int32),
                ("grid_time", ctypes.c_double), ("cell_advances", ctypes.c_uint64),
                ("limiting_cell", ctypes.c_int32), ("reserved_", ctypes.c_int32)]


class ComparisonSummary(ctypes.Structure):
    _fields_ = [("pass_", ctypes.c_int32), ("failures", ctypes.c_int32),
                ("max_deuterium_species_rel_error", ctypes.c_double),
                ("max_non_deuterium_species_rel_error", ctypes.c_double),
                ("max_thermodynamic_rel_error", ctypes.c_double),
                ("max_rho_rel_error", ctypes.c_double),
                ("first_failure", ctypes.c_char * 512)]


def library_path():
    return os.environ.get("WR_LIBRARY", os.path.join(PKG, "lib", "libwarp_reactor.so"))


_LIB = None

_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SIGS = {
    "wr_abi_version": (_I, []),
    "wr_last_error": (ctypes.c_char_p, []),
    "wr_run_cells": (_I, [_VP, _I, _I, ctypes.POINTER(_I)]),
    "wr_context_create": (_I, [_I, _I, ctypes.POINTER(_VP)]),
    "wr_context_destroy": (None, [_VP]),
    "wr_upload_cells": (_I, [_VP, _VP, _I]),
    "wr_download_cells": (_I, [_VP, _VP, _I]),
    "wr_run": (_I, [_VP, _I, _I, _I, _D]),
    "wr_run_from": (_I, [_VP, _I, _I, _I, _D, _I]),
    "wr_set_coop": (_I, [_VP, _I, _I]),
    "wr_get_coop": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "wr_run_info_get": (_I, [_VP, ctypes.POINTER(RunInfo)]),
    "wr_prepare": (_I, [_VP, _I, _D, _I, _I, ctypes.POINTER(_D), ctypes.POINTER(_I)]),
    "wr_advance": (_I, [_VP, _I, _D, _D, ctypes.POINTER(_I)]),
    "wr_prepare_cells_noalloc": (_I, [_VP, _I, _I, _D, _I, _I, _VP, _VP, _VP, _VP]),
    "wr_dt_history": (_I, [_VP, _VP, _I, ctypes.POINTER(_I)]),
    "wr_set_advance_kernel": (_I, [_VP, _I]),
    "wr_set_policy": (_I, [_VP, _I]),
    "wr_set_jacobian_reuse": (_I, [_VP, _I]),
    "wr_set_timing": (_I, [_VP, _I]),
    "wr_set_integrator_max_steps": (_I, [_VP, _I]),
    "wr_set_graph": (_I, [_VP, _I]),
    "wr_set_lu_precision": (_I, [_VP, _I, _I]),
    "wr_set_queue_launch": (_I, [_VP, _I, _I, _I]),
    "wr_get_queue_launch": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                                 ctypes.POINTER(_I)]),
    "wr_mixed_fallbacks": (_I, [_VP, ctypes.POINTER(ctypes.c_ulonglong),
                                ctypes.POINTER(ctypes.c_ulonglong)]),
    "wr_last_advance_timing": (_I, [_VP, ctypes.POINTER(_D), ctypes.POINTER(_I)]),
    "wr_eval_rhs": (_I, [_VP, _VP, _I, _VP]),
    "wr_eval_math": (_I, [_VP, _I, _VP, _I, _VP]),
    "wr_eval_jac": (_I, [_VP, _VP, _I, _VP]),
    "wr_eval_lu": (_I, [_VP, _VP, _I, _VP, _VP]),
    "wr_eval_lu_solve": (_I, [_VP, _VP, _VP, _I, _VP]),
    "wr_eval_integrate": (_I, [_VP, _VP, _VP, _VP, _I, _I, _VP, _VP, _VP, _VP]),
    "wr_eval_solve_mixed": (_I, [_VP, _VP, _VP, _VP, _I, _I, _VP, _VP, _VP]),
    "wr_adjoint_ntheta": (_I, []),
    "wr_adjoint_record": (_I, [_VP, _I, _VP, _I, ctypes.POINTER(_I)]),
    "wr_adjoint_replay": (_I, [_VP, _VP, _VP, _VP]),
    "wr_adjoint_backward": (_I, [_VP, _VP, _VP, _VP]),
    "wr_adjoint_generation": (_I, [_VP, ctypes.POINTER(ctypes.c_ulonglong)]),
    "wr_adjoint_backward_gen": (_I, [_VP, ctypes.c_ulonglong, _VP, _VP, _VP]),
    "wr_adjoint_set_perturb": (_I, [_VP, _I]),
    "wr_adjoint_tape": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "wr_adjoint_last_ms": (_I, [_VP, ctypes.POINTER(_D)]),
    "wr_adjoint_set_window": (_I, [_VP, _I]),
    "wr_adjoint_set_lanes": (_I, [_VP, _I]),
    "wr_make_collapse_states": (_I, [_VP, _I]),
    "wr_make_recipe_states": (_I, [_VP, _I, ctypes.c_uint64, _D, _D, _D]),
    "wr_pack_final_states": (_I, [_VP, _I, _I, _VP]),
    "wr_write_final_states": (_I, [_VP, _I, _I, ctypes.c_char_p]),
    "wr_read_final_states": (_I, [ctypes.c_char_p, _I, _VP]),
    "wr_compare_final_states": (_I, [_VP, _I, _I, ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.POINTER(ComparisonSummary)]),
}
EXPORTED = tuple(_SIGS)


def library():
    """Load libwarp_reactor.so once; raises if it is missing (no fallback)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise WarpReactorError(ERR_ARGUMENT, f"HIP library not built: {path} "
                                   "(run python warp-reactor_amd/build.py)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(_VP)


def _check(code, what):
    if code != SUCCESS:
        msg = library().wr_last_error().decode(errors="replace")
        raise WarpReactorError(code, f"{what}: {msg}")


def adjoint_ntheta():
    """Number of rate multipliers (reactions of the network) of the adjoint."""
    return library().wr_adjoint_ntheta()


def reaction_table():
    """The reaction behind each rate multiplier of the adjoint, as the code
    generator recovered it from the network's rate code: a list of dicts with
    "reaction" (e.g. "H + H- -> e- + H2"), "nu" (stoichiometry over the 14
    species), "rate_terms" and "thermal_monomials" (terms of the energy rate
    that carry its heat; 0 when the network books none)."""
    import json
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "_gen",
                        "chem_adjoint.json")
    with open(path) as fh:
        table = json.load(fh)["reactions"]
    if len(table) != adjoint_ntheta():
        raise WarpReactorError(ERR_ARGUMENT, f"{path} does not describe the built library")
    return table


def make_collapse_states(n):
    cells = np.zeros(n, dtype=CELL_DTYPE)
    _check(library().wr_make_collapse_states(_ptr(cells), n), "wr_make_collapse_states")
    return cells


RECIPES = {
    # name: (seed, sigma of ln n_H, T_min, T_max)
    "uniform16": (0, 0.1, 300.0, 300.0),
    "lognormal128": (1, 1.0, 100.0, 8000.0),
}


def make_recipe_states(n, seed=0, sigma_ln_nh=0.1, t_min=300.0, t_max=None):
    """Benchmark-recipe initial states (wr_make_recipe_states)."""
    cells = np.zeros(n, dtype=CELL_DTYPE)
    _check(library().wr_make_recipe_states(_ptr(cells), n, seed, sigma_ln_nh, t_min,
                                           t_min if t_max is None else t_max),
           "wr_make_recipe_states")
    return cells


def run_cells(cells, perturb=True, completed_global_steps=0):
    """run_cells_cuda (reproducer.cpp:7791): advances `cells` in place.

    Returns (IntegratorResult code, completed_global_steps)."""
    assert cells.dtype == CELL_DTYPE and cells.flags.c_contiguous
    completed = ctypes.c_int(int(completed_global_steps))
    r = library().wr_run_cells(_ptr(cells), len(cells), int(bool(perturb)), ctypes.byref(completed))
    return r, completed.value


def pack_final_states(cells, grid_dim):
    out = np.zeros(len(cells), dtype=PACKED_DTYPE)
    _check(library().wr_pack_final_states(_ptr(cells), len(cells), grid_dim, _ptr(out)),
           "wr_pack_final_states")
    return out


def write_final_states(cells, grid_dim, path):
    _check(library().wr_write_final_states(_ptr(cells), len(cells), grid_dim,
                                           os.fsencode(path)), "wr_write_final_states")


def read_final_states(path, expected):
    out = np.zeros(expected, dtype=PACKED_DTYPE)
    _check(library().wr_read_final_states(os.fsencode(path), expected, _ptr(out)),
           "wr_read_final_states")
    return out


def compare_final_states(cells, grid_dim, path, csv_path=None):
    s = ComparisonSummary()
    _check(library().wr_compare_final_states(
        _ptr(cells), len(cells), grid_dim, os.fsencode(path),
        os.fsencode(csv_path) if csv_path else None, ctypes.byref(s)), "wr_compare_final_states")
    return {"pass": bool(s.pass_), "failures": s.failures,
            "max_deuterium_species_rel_error": s.max_deuterium_species_rel_error,
            "max_non_deuterium_species_rel_error": s.max_non_deuterium_species_rel_error,
            "max_thermodynamic_rel_error": s.max_thermodynamic_rel_error,
            "max_rho_rel_error": s.max_rho_rel_error,
            "first_failure": s.first_failure.decode()}


class Context:
    """A device context holding up to `capacity` cells (SoA on the GPU)."""

    def __init__(self, capacity, device=0):
        self._lib = library()
        h = _VP()
        _check(self._lib.wr_context_create(device, capacity, ctypes.byref(h)), "wr_context_create")
        self._h = h
        self.capacity = capacity

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wr_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, cells):
        assert cells.dtype == CELL_DTYPE and cells.flags.c_contiguous
        _check(self._lib.wr_upload_cells(self._h, _ptr(cells), len(cells)), "wr_upload_cells")
        self._n = len(cells)

    def download(self, cells):
        assert cells.dtype == CELL_DTYPE and cells.flags.c_contiguous
        _check(self._lib.wr_download_cells(self._h, _ptr(cells), len(cells)), "wr_download_cells")

    def upload_ptr(self, ptr, n):
        """Upload n cells from a raw address (host or device memory of the
        context's device, e.g. a torch tensor's data_ptr(); its data must be
        complete before the call)."""
        _check(self._lib.wr_upload_cells(self._h, ctypes.c_void_p(int(ptr)), int(n)),
               "wr_upload_cells")
        self._n = int(n)

    def download_ptr(self, ptr, n):
        _check(self._lib.wr_download_cells(self._h, ctypes.c_void_p(int(ptr)), int(n)),
               "wr_download_cells")

    def run(self, perturb=True, max_steps=1000, completed=0, grid_time=0.0, first_step=None):
        """Outer loop on resident cells; returns the run info as a dict.
        first_step: loop index of the first step (None: the completed count,
        wr_run's resume semantics; 0: run_cells_cuda's loop)."""
        if first_step is None:
            r = self._lib.wr_run(self._h, int(bool(perturb)), max_steps, completed, grid_time)
        else:
            r = self._lib.wr_run_from(self._h, int(bool(perturb)), max_steps, completed,
                                      grid_time, int(first_step))
        if r <= ERR_HIP:
            _check(r, "wr_run")
        info = RunInfo()
        _check(self._lib.wr_run_info_get(self._h, ctypes.byref(info)), "wr_run_info_get")
        return {f: getattr(info, f) for f, _ in RunInfo._fields_ if f != "reserved_"}

    def dt_history(self):
        out = np.zeros(4096)
        n = ctypes.c_int()
        _check(self._lib.wr_dt_history(self._h, _ptr(out), len(out), ctypes.byref(n)),
               "wr_dt_history")
        return out[:n.value].copy()

    def set_policy(self, policy):
        """Driver policy of subsequent runs: "gridwide" (default, run_cells_cuda),
        "legacy-local" (the reference's CPU driver: each cell steps with its
        own free-fall dt) or "gridwide-cta32" (the Mojo front end's policy:
        one step controller per 32-cell tile)."""
        code = {"gridwide": 0, "legacy-local": 1, "gridwide-cta32": 2}.get(policy, policy)
        _check(self._lib.wr_set_policy(self._h, int(code)), "wr_set_policy")

    def set_jacobian_reuse(self, age_max):
        """Accepted steps per Jacobian in the queue-fed kernel (1: the reference;
        k > 1: J lagged up to k accepted steps, refreshed on rejection)."""
        _check(self._lib.wr_set_jacobian_reuse(self._h, int(age_max)), "wr_set_jacobian_reuse")

    def set_advance_kernel(self, kernel):
        """1: queue-fed persistent kernel (default); 0: one thread per cell."""
        _check(self._lib.wr_set_advance_kernel(self._h, int(kernel)), "wr_set_advance_kernel")

    def set_integrator_max_steps(self, max_steps):
        """Integrator step limit of subsequent runs (<= 0: 1e7, the reference's)."""
        _check(self._lib.wr_set_integrator_max_steps(self._h, int(max_steps)),
               "wr_set_integrator_max_steps")

    def set_graph(self, enable=True):
        """Enqueue outer steps as captured HIP graphs (default) or launches."""
        _check(self._lib.wr_set_graph(self._h, int(bool(enable))), "wr_set_graph")

    def set_queue_launch(self, threads_per_block=256, blocks_per_cu=0, solo_lanes=64):
        """Launch shape of the queue-fed kernel (results do not depend on it)."""
        _check(self._lib.wr_set_queue_launch(self._h, int(threads_per_block), int(blocks_per_cu),
                                             int(solo_lanes)), "wr_set_queue_launch")

    def set_coop(self, cells, waves=None):
        """Advance the `cells` most expensive cells of each step with whole
        waves (wave 0 of the first `waves` blocks; default waves = cells);
        0 turns it off, -1 (the context's default) takes one cell per 2,048
        cells, at most 128.  Results are bit-identical."""
        if waves is None:
            waves = cells if cells > 0 else 0
        _check(self._lib.wr_set_coop(self._h, int(cells), int(waves)), "wr_set_coop")

    def coop(self):
        """(cells, waves, LDS bytes per cooperative wave) of the current
        setting for this context's grid."""
        v = [ctypes.c_int() for _ in range(3)]
        _check(self._lib.wr_get_coop(self._h, *[ctypes.byref(x) for x in v]), "wr_get_coop")
        return tuple(x.value for x in v)

    def queue_launch(self):
        """(threads per block, blocks per CU, solo lanes, grid blocks last used)."""
        v = [ctypes.c_int() for _ in range(4)]
        _check(self._lib.wr_get_queue_launch(self._h, *[ctypes.byref(x) for x in v]),
               "wr_get_queue_launch")
        return tuple(x.value for x in v)

    def set_lu_precision(self, mode, sweeps=2):
        """Factorisation in the queue-fed kernel: 0 FP64 (default); 1 FP32
        factors with FP64 residuals and `sweeps` refinement sweeps per stage
        solve (FP64 fallback when singular or not converged)."""
        _check(self._lib.wr_set_lu_precision(self._h, int(mode), int(sweeps)),
               "wr_set_lu_precision")

    def mixed_fallbacks(self):
        """(attempts with singular FP32 factors, stage solves whose refinement
        did not converge) of the last run."""
        a, b = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        _check(self._lib.wr_mixed_fallbacks(self._h, ctypes.byref(a), ctypes.byref(b)),
               "wr_mixed_fallbacks")
        return a.value, b.value

    def set_timing(self, enable=True):
        _check(self._lib.wr_set_timing(self._h, int(bool(enable))), "wr_set_timing")

    def advance_timing(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(self._lib.wr_last_advance_timing(self._h, ctypes.byref(ms), ctypes.byref(n)),
               "wr_last_advance_timing")
        return ms.value, n.value

    def prepare(self, completed, grid_time, step, perturb=True):
        dt, cell = ctypes.c_double(), ctypes.c_int()
        r = self._lib.wr_prepare(self._h, completed, grid_time, step, int(bool(perturb)),
                                 ctypes.byref(dt), ctypes.byref(cell))
        if r <= ERR_HIP:
            _check(r, "wr_prepare")
        return r, dt.value, cell.value

    def advance(self, completed, next_grid_time, dt_grid):
        n = ctypes.c_int()
        r = self._lib.wr_advance(self._h, completed, next_grid_time, dt_grid, ctypes.byref(n))
        if r <= ERR_HIP:
            _check(r, "wr_advance")
        return r, n.value

    # -- device building blocks (verification) ---------------------------------
    def eval_rhs(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, NEQS)
        out = np.zeros_like(y)
        _check(self._lib.wr_eval_rhs(self._h, _ptr(y), len(y), _ptr(out)), "wr_eval_rhs")
        return out

    MATH_OPS = {"exp": 0, "log": 1, "cbrt": 2, "sqrt": 3, "ocml_exp": 4, "ocml_log": 5,
                "ocml_cbrt": 6}

    def eval_math(self, op, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        out = np.zeros_like(x)
        _check(self._lib.wr_eval_math(self._h, self.MATH_OPS[op], _ptr(x), len(x), _ptr(out)),
               "wr_eval_math")
        return out

    def eval_jac(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1, NEQS)
        out = np.zeros((len(y), NEQS, NEQS))
        _check(self._lib.wr_eval_jac(self._h, _ptr(y), len(y), _ptr(out)), "wr_eval_jac")
        return out

    def eval_lu(self, a):
        a = np.array(a, dtype=np.float64).reshape(-1, NEQS, NEQS).copy()
        ipvt = np.zeros((len(a), NEQS), dtype=np.int32)
        info = np.zeros(len(a), dtype=np.int32)
        _check(self._lib.wr_eval_lu(self._h, _ptr(a), len(a), _ptr(ipvt), _ptr(info)),
               "wr_eval_lu")
        return a, ipvt, info

    def eval_lu_solve(self, lu, ipvt, x):
        lu = np.ascontiguousarray(lu, dtype=np.float64).reshape(-1, NEQS, NEQS)
        ipvt = np.ascontiguousarray(ipvt, dtype=np.int32).reshape(-1, NEQS)
        x = np.array(x, dtype=np.float64).reshape(-1, NEQS).copy()
        _check(self._lib.wr_eval_lu_solve(self._h, _ptr(lu), _ptr(ipvt), len(x), _ptr(x)),
               "wr_eval_lu_solve")
        return x

    # ---- discrete adjoint (frozen controller decisions of a recorded run) ----
    def adjoint_record(self, nsteps, multipliers, hcap=512, perturb=False):
        """Run nsteps outer steps of the uploaded cells with rate multipliers
        (and the density perturbations at steps 20, 40, ... when `perturb`),
        recording the tape; returns the completed step count."""
        _check(self._lib.wr_adjoint_set_perturb(self._h, int(bool(perturb))),
               "wr_adjoint_set_perturb")
        M = np.ascontiguousarray(multipliers, dtype=np.float64)
        done = ctypes.c_int()
        r = self._lib.wr_adjoint_record(self._h, int(nsteps), _ptr(M), int(hcap), ctypes.byref(done))
        if r != SUCCESS:
            _check(r, "wr_adjoint_record")
        self._adj = (int(nsteps), int(hcap))
        return done.value

    def adjoint_replay(self, s0=None, multipliers=None):
        n = self._n
        sN = np.zeros((18, n))
        a = None if s0 is None else np.ascontiguousarray(s0, dtype=np.float64)
        m = None if multipliers is None else np.ascontiguousarray(multipliers, dtype=np.float64)
        r = self._lib.wr_adjoint_replay(self._h, _ptr(a) if a is not None else None,
                                        _ptr(m) if m is not None else None, _ptr(sN))
        if r != SUCCESS:
            _check(r, "wr_adjoint_replay")
        return sN

    def adjoint_generation(self):
        """Number of successful recordings on this context (0: none)."""
        g = ctypes.c_ulonglong()
        _check(self._lib.wr_adjoint_generation(self._h, ctypes.byref(g)), "wr_adjoint_generation")
        return g.value

    def adjoint_backward_ptr(self, gN, g0, mbar, generation):
        """wr_adjoint_backward_gen on raw addresses (host or device memory of
        the context's device): gN, g0 [18][n], mbar [R][n] float64."""
        r = self._lib.wr_adjoint_backward_gen(self._h, int(generation), ctypes.c_void_p(int(gN)),
                                              ctypes.c_void_p(int(g0)), ctypes.c_void_p(int(mbar)))
        if r != SUCCESS:
            _check(r, "wr_adjoint_backward")

    def adjoint_backward(self, gN, generation=None):
        """dL/d(first state), dL/dM per cell; `generation` (adjoint_generation()
        right after the recording) makes the call fail if the context's tape
        has since been replaced by another recording."""
        n = self._n
        g0 = np.zeros((18, n))
        mb = np.zeros((adjoint_ntheta(), n))
        g = np.ascontiguousarray(gN, dtype=np.float64)
        if generation is None:
            r = self._lib.wr_adjoint_backward(self._h, _ptr(g), _ptr(g0), _ptr(mb))
        else:
            r = self._lib.wr_adjoint_backward_gen(self._h, int(generation), _ptr(g), _ptr(g0),
                                                  _ptr(mb))
        if r != SUCCESS:
            _check(r, "wr_adjoint_backward")
        return g0, mb

    def adjoint_tape(self):
        nsteps, hcap = self._adj
        n = self._n
        t = {"hoff": np.zeros((nsteps + 1, n), dtype=np.int32), "htape": np.zeros((hcap, n)),
             "dts": np.zeros(nsteps), "ckpt": np.zeros((nsteps, 18, n))}
        _check(self._lib.wr_adjoint_tape(self._h, _ptr(t["hoff"]), _ptr(t["htape"]), _ptr(t["dts"]),
                                         _ptr(t["ckpt"])), "wr_adjoint_tape")
        return t

    def adjoint_set_window(self, max_steps=0):
        """Cap the steps per backward window (0: by the memory budget)."""
        _check(self._lib.wr_adjoint_set_window(self._h, int(max_steps)), "wr_adjoint_set_window")

    def adjoint_set_lanes(self, cells_per_wave=0):
        """Cells per wave of the reverse sweep, 1..64 (0: by grid size)."""
        _check(self._lib.wr_adjoint_set_lanes(self._h, int(cells_per_wave)), "wr_adjoint_set_lanes")

    def adjoint_last_ms(self):
        ms = ctypes.c_double()
        _check(self._lib.wr_adjoint_last_ms(self._h, ctypes.byref(ms)), "wr_adjoint_last_ms")
        return ms.value

    def eval_solve_mixed(self, jac, fac, y, b, sweeps=2):
        """FP32-factorisation stage solves; returns (x, conv, lu32)."""
        jac = np.ascontiguousarray(jac, dtype=np.float64).reshape(-1, 225)
        n = len(jac)
        fac = np.ascontiguousarray(np.broadcast_to(fac, (n,)), dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(n, 15)
        x = np.array(b, dtype=np.float64).reshape(n, 15).copy()
        conv = np.zeros(n, dtype=np.int32)
        lu = np.zeros((n, 225), dtype=np.float32)
        _check(self._lib.wr_eval_solve_mixed(self._h, _ptr(jac), _ptr(fac), _ptr(y), n,
                                             int(sweeps), _ptr(x), _ptr(conv), _ptr(lu)),
               "wr_eval_solve_mixed")
        return x, conv, lu

    def eval_integrate(self, y, tout, dt0, max_steps=0):
        y = np.array(y, dtype=np.float64).reshape(-1, NEQS).copy()
        n = len(y)
        tout = np.ascontiguousarray(np.broadcast_to(tout, (n,)), dtype=np.float64)
        dt0 = np.ascontiguousarray(np.broadcast_to(dt0, (n,)), dtype=np.float64)
        counters = np.zeros((n, 7), dtype=np.int32)
        t_end = np.zeros(n)
        dt_end = np.zeros(n)
        result = np.zeros(n, dtype=np.int32)
        _check(self._lib.wr_eval_integrate(self._h, _ptr(y), _ptr(tout), _ptr(dt0), n, max_steps,
                                           _ptr(counters), _ptr(t_end), _ptr(dt_end),
                                           _ptr(result)), "wr_eval_integrate")
        return {"y": y, "counters": counters, "t_end": t_end, "dt_end": dt_end, "result": result}
