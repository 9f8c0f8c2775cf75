"""Split the network at its temperature-only values, for rate-table studies.

Every value of the fused J + f evaluation that depends on T alone (rate
fits, their exp/log, T powers, the piecewise selections between fits) is
what a rate table would replace.  This emitter finds the *frontier* of that
subgraph: the smooth T-only values (no piecewise selection or range test
below them) that anything else consumes -- an abundance-dependent value or a
selection between fits -- or that are outputs themselves (a Jacobian entry
that is a bare rate).  It writes one header with

``chem_tvals(T, tv)``
    the CHEM_NTV frontier values at T, computed exactly as the network does;
``chem_fused_tv(tv, T, X, f, ...)`` / ``chem_rhs_tv(tv, T, X, f)``
    the network with each frontier value read from ``tv[k]``.

With exact tv the pair reproduces ``chem_fused`` / ``chem_rhs`` bit for bit
(same operations, same operands, same order); a table study hands the second
pair interpolated tv instead.  The selections between fits and the range
tests stay in the second pair, computed from T and the frontier values, so a
table never interpolates across a branch point.  Compile with
-ffp-contract=off like the network header.
"""
from __future__ import annotations

import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dag as dagmod  # noqa: E402
import emit  # noqa: E402

LEAF_OPS = ("const", "bconst", "T", "X")


def smooth(d):
    """smooth[nid]: a T-only value built from T and constants without a
    selection or a test (no piecewise branch below it), so a cubic in ln T
    can follow it."""
    out = []
    for n in d.nodes:
        if n.op in ("const", "T"):
            out.append(True)
        elif n.op in ("X", "bconst", "sel") or n.op in dagmod.BOOL_OPS:
            out.append(False)
        else:
            out.append(all(out[a] for a in n.args))
    return out


def frontier(d, roots):
    """Ordered frontier node ids of the subgraph live from `roots`: smooth
    T-only values consumed by a value that is not smooth (abundance-dependent,
    or a piecewise selection, which the split network computes from T and the
    frontier) or that are outputs."""
    live = d.live_nodes(roots)
    sm = smooth(d)
    rootset = set(roots)
    needed = set()
    for nid in live:
        if sm[nid]:
            continue
        for a in d.nodes[nid].args:
            needed.add(a)
    fr = [nid for nid in sorted(live)
          if sm[nid] and d.nodes[nid].op not in LEAF_OPS and (nid in needed or nid in rootset)]
    return fr


def live_with_leaves(d, roots, leaves):
    """Nodes reachable from roots without passing through `leaves`."""
    seen = set()
    stack = list(roots)
    while stack:
        n = stack.pop()
        if n in seen:
            continue
        seen.add(n)
        if n in leaves:
            continue
        stack.extend(d.nodes[n].args)
    return seen


class _TvEmitter(emit._Emitter):
    def __init__(self, d, slot):
        super().__init__(d)
        self.slot = slot

    def ref(self, nid):
        if nid in self.slot:
            return f"tv[{self.slot[nid]}]"
        return super().ref(nid)

    def tv_function(self, fname, roots_f, roots_j):
        d = self.d
        roots = list(roots_f) + (list(roots_j.values()) if roots_j else [])
        live = live_with_leaves(d, roots, self.slot)
        jac_at = {}
        if roots_j:
            for rc, nid in roots_j.items():
                jac_at.setdefault(nid, []).append(rc)
        sig = "const double* tv, double T, const double* X, double* f"
        if roots_j:
            sig += " CHEM_SINK_PARAMS"
        lines = [f"CHEM_FN void {fname}({sig})", "{"]
        ops = 0
        for nid in sorted(live):
            n = d.nodes[nid]
            if nid in self.slot or n.op in LEAF_OPS:
                for (r, c) in jac_at.get(nid, []):
                    lines.append(f"    CHEM_JAC_STORE({r}, {c}, {self.ref(nid)});")
                continue
            ops += 1
            nm = ("b" if d.is_bool(nid) else "v") + str(nid)
            self.name[nid] = nm
            ty = "int" if d.is_bool(nid) else "double"
            lines.append(f"    const {ty} {nm} = {self.expr(nid)};")
            for (r, c) in jac_at.get(nid, []):
                lines.append(f"    CHEM_JAC_STORE({r}, {c}, {nm});")
        for i, nid in enumerate(roots_f):
            lines.append(f"    f[{i}] = {self.ref(nid)};")
        lines.append("}")
        return "\n".join(lines), ops


def tvals_function(d, fr):
    em = emit._Emitter(d)
    live = d.live_nodes(fr)
    lines = ["CHEM_FN void chem_tvals(double T, double* tv)", "{"]
    ops = {}
    for nid in sorted(live):
        n = d.nodes[nid]
        if n.op in LEAF_OPS:
            continue
        ops[n.op] = ops.get(n.op, 0) + 1
        nm = ("b" if d.is_bool(nid) else "v") + str(nid)
        em.name[nid] = nm
        ty = "int" if d.is_bool(nid) else "double"
        lines.append(f"    const {ty} {nm} = {em.expr(nid)};")
    for k, nid in enumerate(fr):
        lines.append(f"    tv[{k}] = {em.ref(nid)};")
    lines.append("}")
    return "\n".join(lines), ops


def emit_table_header(ir_path, out_path, info_path=None):
    d = emit.specialise(dagmod.read_ir(ir_path))
    f_roots = [d.outputs[("ydot", i)] for i in range(dagmod.NUM_SPEC)] + [d.outputs[("edot",)]]
    j_roots = {(r, c): d.outputs[("jac", r, c)] for r in range(dagmod.NEQS) for c in range(dagmod.NEQS)}
    fr = frontier(d, f_roots + list(j_roots.values()))
    slot = {nid: k for k, nid in enumerate(fr)}
    tvals, tv_ops = tvals_function(d, fr)
    fused, fused_ops = _TvEmitter(d, slot).tv_function("chem_fused_tv", f_roots, j_roots)
    rhs, rhs_ops = _TvEmitter(d, slot).tv_function("chem_rhs_tv", f_roots, None)
    sha = d.meta.get("source_sha256", "?")
    text = "\n".join([
        "/* Generated by warp-reactor_amd/gen/table.py from primordial_network.ir",
        f" * (network source sha256 {sha}). Do not edit.",
        " * The network split at its temperature-only frontier: chem_tvals computes",
        " * the CHEM_NTV values a rate table would hold; chem_fused_tv / chem_rhs_tv",
        " * are chem_fused / chem_rhs reading them from tv.  Same macros as the",
        " * network header. */",
        f"#define CHEM_NTV {len(fr)}",
        "", tvals, "", fused, "", rhs, ""])
    tmp = out_path + ".tmp"
    with open(tmp, "w") as f:
        f.write(text)
    os.replace(tmp, out_path)
    info = {"frontier": len(fr), "tvals_ops": tv_ops, "tvals_exp": tv_ops.get("exp", 0),
            "tvals_log": tv_ops.get("log", 0), "fused_tv_ops": fused_ops, "rhs_tv_ops": rhs_ops,
            "source_sha256": sha}
    if info_path:
        with open(info_path, "w") as f:
            json.dump(info, f, indent=1, sort_keys=True)
    return info


def main(argv):
    here = os.path.dirname(os.path.abspath(__file__))
    info = emit_table_header(os.path.join(here, "primordial_network.ir"), argv[1],
                             argv[2] if len(argv) > 2 else None)
    print(json.dumps(info))


if __name__ == "__main__":
    main(sys.argv)
