"""Register-budgeted emission of the network DAG.

The fused J + f evaluation has thousands of values alive at once in naive
order; left to the compiler it spills thousands of registers to scratch.  This
pass decides the evaluation order and the register contents itself:

1. schedule: a topological order chosen greedily to keep few values alive
   (prefer the ready node that frees the most operands; ties go to the node
   that a depth-first walk from the outputs reaches first);
2. allocate: walk the order with a register file of R doubles; when it is
   full, evict the value whose next use is furthest away (Belady), storing it
   once to a staging slot; reload it (as a new SSA name) before its next use;
3. emit straight-line C/HIP: `STAGE_STORE(slot, v)` / `STAGE_LOAD(slot)`
   for staged values, `OUT(key, v)` for every output as soon as it exists.

Only the evaluation order changes; every operation keeps its operands and
their order, so results are bit-identical to the unstaged code.  Staging
slots are reused once their value is dead (interval colouring).

The emitted functions take (T, X) and write outputs through macros the
includer defines, so the same text compiles as C (tests) and HIP.
"""
from __future__ import annotations

import bisect
import collections
import struct

import dag as dagmod


def _leaf(d, n):
    return d.nodes[n].op in ("const", "bconst", "T", "X", "in")


def cluster_roots(d, roots):
    """Order outputs so that consecutive ones share most of their cones
    (nearest neighbour on cone overlap per newly needed node)."""
    cones = {r: d.live_nodes([r]) for r in roots}
    rem = list(dict.fromkeys(roots))
    cur = rem.pop(0)
    order = [cur]
    acc = set(cones[cur])
    while rem:
        best = max(rem, key=lambda r: (len(cones[r] & cones[cur]) / (1 + len(cones[r] - acc)),
                                       -rem.index(r)))
        rem.remove(best)
        order.append(best)
        acc |= cones[best]
        cur = best
    return order


def schedule(d, roots):
    """Greedy low-pressure topological order of the non-leaf nodes under roots.
    Outputs are consumed as soon as they are computed."""
    roots = cluster_roots(d, roots)
    live = d.live_nodes(roots)
    nodes = [n for n in live if not _leaf(d, n)]
    users = {n: [] for n in nodes}
    for n in nodes:
        for a in set(d.nodes[n].args):
            if a in users:
                users[a].append(n)
    remaining = {n: len(users[n]) for n in nodes}
    npend = {n: sum(1 for a in set(d.nodes[n].args) if a in users) for n in nodes}
    order = {}

    def visit(n):
        stack = [(n, False)]
        while stack:
            m, done = stack.pop()
            if m in order or _leaf(d, m):
                continue
            if done:
                order[m] = len(order)
                continue
            stack.append((m, True))
            for a in reversed(d.nodes[m].args):
                if a not in order and not _leaf(d, a):
                    stack.append((a, False))

    for r in roots:
        visit(r)
    ready = set(n for n in nodes if npend[n] == 0)
    seq = []
    while ready:
        best, bs = None, None
        for n in ready:
            free = sum(1 for a in set(d.nodes[n].args) if a in remaining and remaining[a] == 1)
            keep = 1 if remaining[n] > 0 else 0
            s = (free - keep, -order[n])
            if bs is None or s > bs:
                bs, best = s, n
        ready.discard(best)
        seq.append(best)
        for a in set(d.nodes[best].args):
            if a in remaining:
                remaining[a] -= 1
        for u in users[best]:
            npend[u] -= 1
            if npend[u] == 0:
                ready.add(u)
    assert len(seq) == len(nodes)
    return seq


class Plan:
    """Result of allocation: a list of actions for the emitter."""

    def __init__(self):
        self.actions = []   # ("load", node, slot) | ("op", node) | ("store", node, slot)
        self.nslots = 0
        self.stores = 0
        self.loads = 0
        self.peak = 0
        self.pressure = []  # values held in registers at each action


def allocate(d, seq, R):
    uses = collections.defaultdict(list)
    for i, n in enumerate(seq):
        for a in set(d.nodes[n].args):
            if not _leaf(d, a):
                uses[a].append(i)

    def next_use(n, i):
        u = uses.get(n)
        if not u:
            return None
        k = bisect.bisect_left(u, i)
        return u[k] if k < len(u) else None

    plan = Plan()
    inreg = set()
    slot_of = {}
    free_slots = []
    slot_owner = {}
    for i, n in enumerate(seq):
        args = [a for a in dict.fromkeys(d.nodes[n].args) if not _leaf(d, a)]
        # bring operands in
        for a in args:
            if a not in inreg:
                plan.actions.append(("load", a, slot_of[a]))
                plan.loads += 1
                inreg.add(a)
                plan.pressure.append(len(inreg))
        # make room: evict furthest next use, never the operands of this op
        while len(inreg) + 1 > R:
            cands = [v for v in inreg if v not in args]
            if not cands:
                break
            far = max(cands, key=lambda v: (next_use(v, i) if next_use(v, i) is not None else 1 << 60))
            if next_use(far, i) is not None and far not in slot_of:
                s = free_slots.pop() if free_slots else plan.nslots
                if s == plan.nslots:
                    plan.nslots += 1
                slot_of[far] = s
                slot_owner[s] = far
                plan.actions.append(("store", far, s))
                plan.pressure.append(len(inreg))
                plan.stores += 1
            inreg.discard(far)
        plan.actions.append(("op", n))
        inreg.add(n)
        plan.pressure.append(len(inreg))
        plan.peak = max(plan.peak, len(inreg))
        # retire dead values (including this op if it has no later use)
        for v in args + [n]:
            if next_use(v, i + 1) is None:
                inreg.discard(v)
                if v in slot_of and slot_owner.get(slot_of[v]) == v:
                    free_slots.append(slot_of[v])
                    slot_owner.pop(slot_of[v])
    return plan


def hoist_loads(plan, distance, extra):
    """Move every staging reload up to `distance` actions earlier (never above
    the store that filled its slot), as long as no point then holds more than
    R + `extra` values: the load's latency is then covered by the work in
    between instead of stalling its first use.  Returns a new action list."""
    import numpy as np
    acts = plan.actions
    n = len(acts)
    press = np.array(plan.pressure, dtype=np.int64)
    cap = int(press.max()) + extra
    store_at = {}
    for i, a in enumerate(acts):
        if a[0] == "store":
            store_at[(a[1], a[2])] = i
    moved_to = {}
    for i, a in enumerate(acts):
        if a[0] != "load":
            continue
        lo = max(store_at[(a[1], a[2])] + 1, i - distance)
        if lo >= i:
            continue
        seg = press[lo:i]
        over = np.nonzero(seg + 1 > cap)[0]
        q = lo + (int(over[-1]) + 1 if len(over) else 0)
        if q >= i:
            continue
        press[q:i] += 1
        moved_to[i] = q
    at = collections.defaultdict(list)
    for i, q in moved_to.items():
        at[q].append(i)
    out = []
    for i, a in enumerate(acts):
        for j in at.get(i, ()):
            out.append(acts[j])
        if i not in moved_to:
            out.append(a)
    assert len(out) == n
    return out


# FP64 constants the gfx950 VALU takes as inline operands (no literal needed)
_INLINE_F64 = {0.0, 0.5, -0.5, 1.0, -1.0, 2.0, -2.0, 4.0, -4.0}


def emit_staged(d, fname, outputs, R, sig_extra="", out_macro="CHEM_OUT", prefetch=0,
                prefetch_extra=0, const_table=None, sync_every=0, params=None, in_macro=None):
    """const_table: name of a table the constants are read from through
    CHEM_K(name, i) (indices in order of first use; values in plan.ktab),
    except those the hardware encodes inline.  sync_every: emit CHEM_SYNC()
    every that many operations (keeps the waves of a block in lockstep so
    they share instruction fetches)."""
    """outputs: dict node -> list of output keys (str) handed to OUT macro.
    Returns (source text, plan)."""
    roots = list(outputs)
    seq = schedule(d, roots)
    plan = allocate(d, seq, R)
    if prefetch > 0:
        plan.actions = hoist_loads(plan, prefetch, prefetch_extra)
    cur = {}       # node -> current C name
    gen = collections.Counter()
    if params is None:
        params = f"double T, const double* X{sig_extra}"
    lines = [f"CHEM_FN void {fname}({params})", "{"]
    kidx = {}
    plan.ktab = []

    def ref(nid):
        n = d.nodes[nid]
        if n.op == "const":
            v = d.cval(nid)
            if const_table and v == v and abs(v) != float("inf") and v not in _INLINE_F64:
                bits_ = struct.pack("<d", v)
                if bits_ not in kidx:
                    kidx[bits_] = len(plan.ktab)
                    plan.ktab.append(v)
                return f"CHEM_K({const_table}, {kidx[bits_]})"
            s = v.hex() if v == v and abs(v) != float("inf") else (
                "(0.0/0.0)" if v != v else ("(1.0/0.0)" if v > 0 else "(-1.0/0.0)"))
            return f"({s})" if s.startswith("-") else s
        if n.op == "bconst":
            return "1" if n.val else "0"
        if n.op == "T":
            return "T"
        if n.op == "X":
            return f"X[{n.val}]"
        if n.op == "in":
            name, idx = n.val.split(":")
            return f"{in_macro}({name}, {idx})" if in_macro else f"{name}[{idx}]"
        return cur[nid]

    def expr(nid):
        n = d.nodes[nid]
        a = [ref(x) for x in n.args]
        op = n.op
        if op == "neg":
            return f"-{a[0]}"
        if op == "abs":
            return f"fabs({a[0]})"
        if op in ("exp", "log"):
            return f"CHEM_{op.upper()}({a[0]})"
        if op == "sqrt":
            return f"sqrt({a[0]})"
        bin_ = {"add": "+", "sub": "-", "mul": "*", "div": "/", "lt": "<", "le": "<=", "gt": ">",
                "ge": ">=", "eq": "==", "ne": "!=", "and": "&&", "or": "||"}
        if op in bin_:
            return f"{a[0]} {bin_[op]} {a[1]}"
        if op == "not":
            return f"!{a[0]}"
        if op == "sel":
            return f"{a[0]} ? {a[1]} : {a[2]}"
        raise ValueError(op)

    # outputs that are leaves (constants / inputs)
    for nid, keys in outputs.items():
        if _leaf(d, nid):
            for k in keys:
                lines.append(f"    {out_macro}({k}, {ref(nid)});")
    nops = 0
    for act in plan.actions:
        kind, nid = act[0], act[1]
        is_b = d.is_bool(nid)
        if kind == "op":
            nops += 1
            if sync_every and nops % sync_every == 0:
                lines.append("    CHEM_SYNC();")
            gen[nid] += 1
            nm = ("b" if is_b else "v") + f"{nid}"
            if gen[nid] > 1:
                nm += f"_{gen[nid]}"
            ty = "int" if is_b else "double"
            lines.append(f"    const {ty} {nm} = {expr(nid)};")
            cur[nid] = nm
            for k in outputs.get(nid, ()):
                lines.append(f"    {out_macro}({k}, {nm});")
        elif kind == "store":
            val = f"({cur[nid]} ? 1.0 : 0.0)" if is_b else cur[nid]
            lines.append(f"    STAGE_STORE({act[2]}, {val});")
        else:  # load
            gen[nid] += 1
            nm = ("b" if is_b else "v") + f"{nid}_r{gen[nid]}"
            if is_b:
                lines.append(f"    const int {nm} = STAGE_LOAD({act[2]}) != 0.0;")
            else:
                lines.append(f"    const double {nm} = STAGE_LOAD({act[2]});")
            cur[nid] = nm
    lines.append("}")
    return "\n".join(lines), plan
