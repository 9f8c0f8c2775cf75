"""Emit straight-line FP64 code for the primordial network from the DAG IR.

Two functions are emitted into one header that compiles both as C (the CPU
oracle) and as HIP device code (gfx950):

``chem_fused(T, X, f)``
    right-hand side f = (ydot[0..13], edot) *and* the 15x15 Jacobian at the
    same state.  Each Jacobian entry is handed to the macro
    ``CHEM_JAC_STORE(r, c, value)`` as soon as it is computed; the macro may
    use the extra parameters the includer declares in ``CHEM_SINK_PARAMS``
    (e.g. a pointer and a stride), so the caller decides where it goes (a row-major array on the CPU; the E = I/(h*gamma)-J
    workspace on the GPU).  This is the reference's J evaluation
    (``reproducer.cpp:7154``, ``7052-7060``) fused with the first stage RHS at
    the same y (``7275``), which share every rate coefficient.
``chem_rhs(T, X, f)``
    right-hand side only (``reproducer.cpp:7040-7050``, ``1515-1535``).

Every operation is emitted in the DAG's order of operands; compile with
``-ffp-contract=off`` so no multiply-add is fused.  The redshift ``z`` is the
constant 30 (``reproducer.cpp:201-226``: device builds hard-code it).

Static op counts of each emitted function are written to a JSON sidecar; they
are the per-call work figures used for the FP64 roofline.
"""
from __future__ import annotations

import json
import os
import struct
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dag as dagmod  # noqa: E402

REDSHIFT = 30.0


def specialise(d, z=REDSHIFT):
    """Rebuild the DAG with z replaced by a constant (re-folding what that enables)."""
    out = dagmod.Dag(meta=dict(d.meta))
    remap = {}
    for nid, n in enumerate(d.nodes):
        if n.op == "z":
            remap[nid] = out.const(z)
        elif n.op == "const":
            remap[nid] = out.const(d.cval(nid))
        elif n.op in ("bconst", "T"):
            remap[nid] = out.add(n.op, (), n.val)
        elif n.op == "X":
            remap[nid] = out.add("X", (), n.val)
        else:
            remap[nid] = out.add(n.op, tuple(remap[a] for a in n.args))
    out.outputs = {k: remap[v] for k, v in d.outputs.items()}
    return out


def _lit(v):
    if v != v:
        return "(0.0/0.0)"
    if v in (float("inf"), float("-inf")):
        return "(1.0/0.0)" if v > 0 else "(-1.0/0.0)"
    return v.hex()


class _Emitter:
    def __init__(self, d):
        self.d = d
        self.name = {}

    def ref(self, nid):
        n = self.d.nodes[nid]
        if n.op == "const":
            v = self.d.cval(nid)
            s = _lit(v)
            return f"({s})" if s.startswith("-") else s
        if n.op == "bconst":
            return "1" if n.val else "0"
        if n.op == "T":
            return "T"
        if n.op == "X":
            return f"X[{n.val}]"
        return self.name[nid]

    def expr(self, nid):
        n = self.d.nodes[nid]
        a = [self.ref(x) for x in n.args]
        op = n.op
        if op == "neg":
            return f"-{a[0]}"
        if op == "abs":
            return f"fabs({a[0]})"
        if op in ("exp", "log"):
            return f"CHEM_{op.upper()}({a[0]})"
        if op == "sqrt":
            return f"sqrt({a[0]})"
        binop = {"add": "+", "sub": "-", "mul": "*", "div": "/", "lt": "<", "le": "<=",
                 "gt": ">", "ge": ">=", "eq": "==", "ne": "!=", "and": "&&", "or": "||"}
        if op in binop:
            return f"{a[0]} {binop[op]} {a[1]}"
        if op == "not":
            return f"!{a[0]}"
        if op == "sel":
            return f"{a[0]} ? {a[1]} : {a[2]}"
        raise ValueError(op)

    def function(self, fname, roots_f, roots_j):
        """roots_f: list of 15 node ids; roots_j: dict (r,c)->node id or None."""
        d = self.d
        roots = list(roots_f) + (list(roots_j.values()) if roots_j else [])
        live = d.live_nodes(roots)
        order = sorted(live)
        jac_at = {}
        if roots_j:
            for rc, nid in roots_j.items():
                jac_at.setdefault(nid, []).append(rc)
        lines = []
        sig = "double T, const double* X, double* f"
        if roots_j:
            sig += " CHEM_SINK_PARAMS"
        lines.append(f"CHEM_FN void {fname}({sig})")
        lines.append("{")
        # outputs that are constants / inputs are stored first
        for nid in order:
            n = d.nodes[nid]
            if n.op in ("const", "bconst", "T", "X") and nid in jac_at:
                for (r, c) in jac_at[nid]:
                    lines.append(f"    CHEM_JAC_STORE({r}, {c}, {self.ref(nid)});")
        counts = {}
        for nid in order:
            n = d.nodes[nid]
            if n.op in ("const", "bconst", "T", "X"):
                continue
            counts[n.op] = counts.get(n.op, 0) + 1
            is_b = d.is_bool(nid)
            nm = ("b" if is_b else "v") + str(nid)
            self.name[nid] = nm
            ty = "int" if is_b else "double"
            lines.append(f"    const {ty} {nm} = {self.expr(nid)};")
            for (r, c) in jac_at.get(nid, []):
                lines.append(f"    CHEM_JAC_STORE({r}, {c}, {nm});")
        for i, nid in enumerate(roots_f):
            lines.append(f"    f[{i}] = {self.ref(nid)};")
        lines.append("}")
        return "\n".join(lines), counts


def emit_staged_pair(d, f_roots, j_roots, r_fused, r_rhs, prefetch=500, prefetch_extra=32):
    """Register-budgeted versions (stage.py): outputs leave through
    CHEM_OUT(J, r, c, v) / CHEM_OUT(F, i, v) as soon as computed; values
    that do not fit R doubles go through STAGE_STORE/STAGE_LOAD slots, and
    each reload is issued up to `prefetch` operations before its use."""
    import stage
    outs = {}
    for i, nid in enumerate(f_roots):
        outs.setdefault(nid, []).append(f"F, {i}")
    rhs_outs = {k: list(v) for k, v in outs.items()}
    for (r, c), nid in j_roots.items():
        outs.setdefault(nid, []).append(f"J, {r}, {c}")
    fused, pf = stage.emit_staged(d, "chem_fused_staged", outs, r_fused, " CHEM_STAGE_PARAMS", "CHEM_OUT",
                                  prefetch=prefetch, prefetch_extra=prefetch_extra)
    rhs, pr = stage.emit_staged(d, "chem_rhs_staged", rhs_outs, r_rhs, " CHEM_STAGE_PARAMS", "CHEM_OUT",
                                prefetch=prefetch, prefetch_extra=prefetch_extra)
    info = {"fused": {"R": r_fused, "slots": pf.nslots, "stores": pf.stores, "loads": pf.loads},
            "rhs": {"R": r_rhs, "slots": pr.nslots, "stores": pr.stores, "loads": pr.loads},
            "prefetch": prefetch, "prefetch_extra": prefetch_extra}
    text = "\n".join([
        "#ifdef CHEM_STAGE_PARAMS",
        f"#define CHEM_FUSED_NSLOT {max(pf.nslots, 1)}",
        f"#define CHEM_RHS_NSLOT {max(pr.nslots, 1)}",
        fused, "", rhs,
        "#endif /* CHEM_STAGE_PARAMS */", ""])
    return text, info


def emit_header(ir_path, out_path, counts_path=None, r_fused=128, r_rhs=96):
    d = specialise(dagmod.read_ir(ir_path))
    f_roots = [d.outputs[("ydot", i)] for i in range(dagmod.NUM_SPEC)] + [d.outputs[("edot",)]]
    j_roots = {(r, c): d.outputs[("jac", r, c)] for r in range(dagmod.NEQS) for c in range(dagmod.NEQS)}
    em = _Emitter(d)
    fused, c_fused = em.function("chem_fused", f_roots, j_roots)
    em = _Emitter(d)
    rhs, c_rhs = em.function("chem_rhs", f_roots, None)
    staged, staged_info = emit_staged_pair(d, f_roots, j_roots, r_fused, r_rhs)
    sha = d.meta.get("source_sha256", "?")
    hdr = [
        "/* Generated by warp-reactor_amd/gen/emit.py from primordial_network.ir",
        f" * (network source sha256 {sha}). Do not edit.",
        " * Straight-line FP64 right-hand side and Jacobian of the 14-species +",
        " * energy primordial network (reproducer.cpp:341-6994). Compile with",
        " * -ffp-contract=off. The includer defines CHEM_FN, CHEM_JAC_STORE,",
        " * CHEM_SINK_PARAMS and the exp/log it wants as CHEM_EXP / CHEM_LOG. */",
        "#ifndef CHEM_EXP",
        "#error define CHEM_EXP(x) and CHEM_LOG(x) before including the generated network",
        "#endif",
        "#ifndef CHEM_FN",
        "#error define CHEM_FN before including the generated network",
        "#endif",
        "#ifndef CHEM_JAC_STORE",
        "#error define CHEM_JAC_STORE(r, c, v) before including the generated network",
        "#endif",
        "#ifndef CHEM_SINK_PARAMS",
        "#error define CHEM_SINK_PARAMS (extra parameters of chem_fused, leading comma)",
        "#endif",
        "",
        fused,
        "",
        rhs,
        "",
        staged,
    ]
    text = "\n".join(hdr)
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    tmp = out_path + ".tmp"
    with open(tmp, "w") as f:
        f.write(text)
    os.replace(tmp, out_path)
    if counts_path:
        with open(counts_path, "w") as f:
            json.dump({"chem_fused": c_fused, "chem_rhs": c_rhs, "staged": staged_info,
                       "source_sha256": sha}, f, indent=1, sort_keys=True)
    return c_fused, c_rhs


def main(argv):
    here = os.path.dirname(os.path.abspath(__file__))
    ir = os.path.join(here, "primordial_network.ir")
    out = argv[1]
    counts = argv[2] if len(argv) > 2 else None
    cf, cr = emit_header(ir, out, counts)
    print("chem_fused:", cf)
    print("chem_rhs:", cr)


if __name__ == "__main__":
    main(sys.argv)
