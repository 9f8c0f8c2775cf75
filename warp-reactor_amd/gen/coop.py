"""Wave-cooperative evaluation of the network DAG: one 64-lane wavefront
evaluates one cell's fused J + f (or f alone).

The per-lane network code (stage.py) gives each lane its own cell: a cell that
needs tens of ROS2S attempts in an outer step is then a serial chain of
~27,000 instructions per evaluation on one lane.  Here the same DAG is
spread over the 64 lanes of one wave:

* *vector steps*: one operation class applied to up to 64 DAG nodes at once,
  one node per lane.  Operand and result locations are per-lane LDS byte
  addresses taken from a lane-interleaved table in global memory (word w of
  step s for lane l at tab[(base_s + w) * 64 + l]); lanes past the step's
  node count read a constant slot and write a trash slot, so every lane
  executes every step;
* *uniform items*: nodes on thin, latency-critical stretches of the DAG (long
  sequential sums, the tails of the piecewise rate fits) are computed by all
  lanes redundantly in registers, reading vector results from LDS at fixed
  addresses and storing their own results where vector steps need them.

Every node is still computed by exactly its own operation on exactly its own
operands, so the results are those of the plain code bit for bit.  The only
rewriting is exact: chains of boolean and/or are rebalanced into trees
(no rounding exists in boolean algebra), chains of selects are turned into
balanced priority-select trees (a select only moves a value), and `a < b`,
`a <= b` become `b > a`, `b >= a` (IEEE comparisons are symmetric, NaN
included), and `s != 0.0` of a select `s` between booleans (which the plain
code holds as the double 0.0 or 1.0) becomes `s`.  A subtraction is an
addition of the negated operand, as gfx950 executes every FP64 subtract.

Memory order is the only synchronisation: one wave owns its LDS region, and
LDS operations of a wave complete in issue order.  The emitted code issues
the LDS loads of step k+1 before the store of step k; the scheduler therefore
never lets a step read a value produced by the step right before it.  The
order of a slot's last load and the store that reuses the slot holds only
while the compiler sees the two accesses as possibly aliasing: booleans and
numbers live in separate slots (one access type per LDS byte), and the HIP
accessors use may_alias types.

The emitted header compiles as HIP (one lane per thread) and as C (every
vector statement looped over 64 lanes), so the CPU test suite checks the
emitted code itself against the plain network, bit for bit.
"""
from __future__ import annotations

import collections
import struct

import dag as dagmod

LANES = 64
ADDR_MASK = 0x3FFFF          # LDS byte addresses (< 256 KiB)
FLIP = 0x80000000            # word flag: flip the operand's sign (add) / xor sign (sgn)
CLEAR = 0x40000000           # word flag (sgn): clear the sign bit first

# operation classes of vector steps and their table words per lane
VCLASS = {"add": "add", "sub": "add", "mul": "mul", "div": "div", "exp": "exp", "log": "log",
          "sqrt": "sqrt", "neg": "sgn", "abs": "sgn", "ge": "ge", "gt": "gt", "eq": "eq",
          "ne": "ne", "and": "and", "or": "or", "not": "not", "sel": "sel"}
WORDS = {"add": 3, "mul": 3, "div": 3, "exp": 2, "log": 2, "sqrt": 2, "sgn": 2, "ge": 3,
         "gt": 3, "eq": 3, "ne": 3, "and": 3, "or": 3, "not": 2, "sel": 4, "selb": 4, "mov": 2}
# rough issue cost (instructions) of a vector step and of a uniform op
VCOST = {"add": 9, "mul": 6, "div": 16, "exp": 34, "log": 42, "sqrt": 14, "sgn": 7, "ge": 7,
         "gt": 7, "eq": 7, "ne": 7, "and": 6, "or": 6, "not": 5, "sel": 8, "selb": 8, "mov": 4}
UCOST = {"add": 1, "mul": 1, "div": 10, "exp": 28, "log": 36, "sqrt": 10, "sgn": 1, "ge": 2,
         "gt": 2, "eq": 2, "ne": 2, "and": 1, "or": 1, "not": 1, "sel": 2, "selb": 2}
ULAT = {"div": 40, "exp": 120, "log": 140, "sqrt": 40}
NO_REUSE = False             # debugging: give every value its own slot
TRACE = False                # debugging: COOP_TRACE(item) after every item
LDS_LAT = 110                # cycles from a store's issue to a dependent load's data


def _leaf(d, n):
    return d.nodes[n].op in ("const", "bconst", "T", "X")


# ---------------------------------------------------------------------------
# exact rewriting
# ---------------------------------------------------------------------------

def rebalance(d, roots):
    """Copy of d (restricted to roots' cone) with boolean and/or chains as
    balanced trees, select chains as balanced priority trees and lt/le turned
    into gt/ge.  Returns (new dag, new root ids in the order of `roots`)."""
    live = d.live_nodes(roots)
    nuse = collections.Counter()
    for n in live:
        for a in set(d.nodes[n].args):
            nuse[a] += 1
    for r in roots:
        nuse[r] += 1
    out = dagmod.Dag(meta=dict(d.meta))
    remap = {}

    def tree(op, items):
        items = list(items)
        while len(items) > 1:
            nxt = [out.add(op, (items[i], items[i + 1])) for i in range(0, len(items) - 1, 2)]
            if len(items) % 2:
                nxt.append(items[-1])
            items = nxt
        return items[0]

    def chain_leaves(n, op):
        leaves, stack = [], [n]
        while stack:
            m = stack.pop()
            for a in reversed(d.nodes[m].args):
                if d.nodes[a].op == op and nuse[a] == 1:
                    stack.append(a)
                else:
                    leaves.append(a)
        return leaves

    def prio(pairs, default):
        if not pairs:
            return default
        if len(pairs) == 1:
            return out.add("sel", (pairs[0][0], pairs[0][1], default))
        mid = len(pairs) // 2
        return out.add("sel", (tree("or", [c for c, _ in pairs[:mid]]), nodef(pairs[:mid]),
                               prio(pairs[mid:], default)))

    def nodef(pairs):
        if len(pairs) == 1:
            return pairs[0][1]
        mid = len(pairs) // 2
        return out.add("sel", (tree("or", [c for c, _ in pairs[:mid]]), nodef(pairs[:mid]),
                               nodef(pairs[mid:])))

    for n in sorted(live):
        node = d.nodes[n]
        op = node.op
        if op == "const":
            remap[n] = out.const(d.cval(n))
        elif op in ("bconst", "T", "X"):
            remap[n] = out.add(op, (), node.val)
        elif op in ("and", "or"):
            remap[n] = tree(op, [remap[a] for a in chain_leaves(n, op)])
        elif op == "sel":
            pairs, cur = [], n
            while True:
                c, a, b = d.nodes[cur].args
                if d.nodes[b].op == "sel" and nuse[b] == 1:
                    pairs.append((remap[c], remap[a]))
                    cur = b
                elif d.nodes[a].op == "sel" and nuse[a] == 1:
                    pairs.append((out.add("not", (remap[c],)), remap[b]))
                    cur = a
                else:
                    pairs.append((remap[c], remap[a]))
                    default = remap[b]
                    break
            remap[n] = prio(pairs, default) if len(pairs) >= 3 else None
            if remap[n] is None:
                c, a, b = node.args
                remap[n] = out.add("sel", (remap[c], remap[a], remap[b]))
        elif (op == "ne" and is_bool(d, node.args[0]) and d.nodes[node.args[1]].op == "const"
              and d.cval(node.args[1]) == 0.0):
            # truthy() of a select between booleans: the plain code holds the
            # select as the double 0.0 or 1.0, so `!= 0.0` is the boolean itself
            remap[n] = remap[node.args[0]]
        elif op == "lt":
            remap[n] = out.add("gt", (remap[node.args[1]], remap[node.args[0]]))
        elif op == "le":
            remap[n] = out.add("ge", (remap[node.args[1]], remap[node.args[0]]))
        else:
            remap[n] = out.add(op, tuple(remap[a] for a in node.args))
    return out, [remap[r] for r in roots]


# ---------------------------------------------------------------------------
# scheduling
# ---------------------------------------------------------------------------

def is_bool(d, n):
    """Whether node n is a boolean: a comparison or logical op, or a select
    between booleans (Dag.is_bool does not look through selects)."""
    memo = d.__dict__.setdefault("_coop_bool", {})
    got = memo.get(n)
    if got is None:
        node = d.nodes[n]
        got = is_bool(d, node.args[1]) if node.op == "sel" else d.is_bool(n)
        memo[n] = got
    return got


def vclass(d, n):
    node = d.nodes[n]
    if node.op == "sel":
        return "selb" if is_bool(d, node.args[1]) else "sel"
    return VCLASS[node.op]


class Program:
    """items: list of ("v", cls, [nodes]) vector steps and ("u", node) uniform ops."""

    def __init__(self, d, roots, name):
        self.d, self.roots, self.name = d, roots, name
        self.items = []
        self.where = {}      # node -> ("v"|"u", item index)


def schedule(d, roots, name, min_lanes=4, uniform_window=48):
    live = d.live_nodes(roots)
    nodes = sorted(n for n in live if not _leaf(d, n))
    users = collections.defaultdict(list)
    for n in nodes:
        for a in set(d.nodes[n].args):
            if not _leaf(d, a):
                users[a].append(n)
    bl = {}
    for n in reversed(nodes):
        c = VCOST[vclass(d, n)] * 4 + 60
        bl[n] = c + max([bl[u] for u in users[n]] or [0])
    npend = {n: sum(1 for a in set(d.nodes[n].args) if not _leaf(d, a)) for n in nodes}
    ready = set(n for n in nodes if npend[n] == 0)
    prog = Program(d, roots, name)
    avail_v = {}  # node -> cycle its value is usable by a vector step
    avail_u = {}  # node -> cycle its value is usable by a uniform op
    t = 0.0

    def ops(n):
        return [a for a in set(d.nodes[n].args) if not _leaf(d, a)]

    def vec_ok(n, k):
        # LDS operands must come from items <= k - 2 (loads of step k are
        # issued before the store of item k - 1)
        return all(prog.where[a][1] <= k - 2 for a in ops(n))

    def emit(kind, cls, take):
        nonlocal t
        k = len(prog.items)
        if kind == "v":
            start = max([t] + [avail_v[a] for n in take for a in ops(n)])
            cost = VCOST[cls] * 4
            t = start + cost
            for n in take:
                avail_v[n] = t + LDS_LAT
                avail_u[n] = t + LDS_LAT
                prog.where[n] = ("v", k)
            prog.items.append(("v", cls, list(take)))
        else:
            n = take[0]
            start = max([t] + [avail_u[a] for a in ops(n)])
            t = start + UCOST[cls] * 4
            avail_u[n] = t + ULAT.get(cls, 8)
            avail_v[n] = t + 60
            prog.where[n] = ("u", k)
            prog.items.append(("u", n))
        for n in take:
            ready.discard(n)
            for u in users[n]:
                npend[u] -= 1
                if npend[u] == 0:
                    ready.add(u)

    while ready:
        k = len(prog.items)
        vready = [n for n in ready if vec_ok(n, k) and
                  max([avail_v[a] for a in ops(n)] or [0]) <= t]
        bycls = collections.defaultdict(list)
        for n in vready:
            bycls[vclass(d, n)].append(n)
        big = {c: v for c, v in bycls.items()
               if len(v) >= min_lanes or (c in ("exp", "log") and len(v) >= 2)}
        if big:
            c = max(big, key=lambda c: (max(bl[n] for n in big[c]), len(big[c])))
            emit("v", c, sorted(big[c], key=lambda n: -bl[n])[:LANES])
            continue
        # thin region: the most critical ready node as a uniform op (expensive
        # transcendentals only when nothing else is ready)
        cand = sorted(ready, key=lambda n: (max([avail_u[a] for a in ops(n)] + [t]), -bl[n]))
        n = cand[0]
        cls = vclass(d, n)
        if cls in ("exp", "log") and bycls.get(cls):
            emit("v", cls, sorted(bycls[cls], key=lambda m: -bl[m])[:LANES])
            continue
        if cls in ("exp", "log", "div", "sqrt") and vec_ok(n, k):
            emit("v", cls, [n])
            continue
        emit("u", cls, [n])
    assert len(prog.where) == len(nodes)
    prog.cycles = max(list(avail_u.values()) + [t])
    prog.uniform_window = uniform_window
    return prog


# ---------------------------------------------------------------------------
# LDS slots and tables
# ---------------------------------------------------------------------------

class Layout:
    """LDS byte layout of a cooperative wave's region (the generator decides
    the constant and value-slot counts; the hand-written part uses the rest)."""

    def __init__(self, consts, nnum, nbool):
        self.consts = consts                  # list of 64-bit patterns
        off = 0
        self.const = off
        off += 8 * len(consts)
        self.trash = off
        off += 8
        self.inp = off                        # T, X[0..13]
        off += 8 * 15
        self.val = off                        # number slots (doubles)
        off += 8 * nnum
        self.valb = off                       # boolean slots (32-bit words, 8-byte pitch)
        off += 8 * nbool
        self.jac = off                        # J[15][15] row-major
        off += 8 * 225
        self.vec = off                        # y, f0, ak1, ak2, ak3, ynew, dy (15 each)
        off += 8 * 105
        self.lu = off                         # LINPACK factors [15][15]
        off += 8 * 225
        self.col = off                        # pivot-search column scratch
        off += 8 * 16
        self.total = off


VEC = {"y": 0, "f0": 15, "ak1": 30, "ak2": 45, "ak3": 60, "ynew": 75, "dy": 90}


def _bits(v):
    return struct.unpack("<Q", struct.pack("<d", v))[0]


def assign(prog, outputs, layout_fn):
    """Slots for every LDS-resident value.  outputs: node -> list of output
    byte offsets relative to the layout (callables of the layout).  Returns the
    per-node address map and the list of copy moves for extra outputs."""
    d = prog.d
    items = prog.items
    K = len(items)
    # which values need LDS: vector-produced; uniform-produced with a vector
    # consumer, a far uniform consumer, or an output
    cons = collections.defaultdict(list)  # node -> [(item, kind)]
    for k, it in enumerate(items):
        if it[0] == "v":
            for n in it[2]:
                for a in set(d.nodes[n].args):
                    if not _leaf(d, a):
                        cons[a].append((k, "v"))
        else:
            for a in set(d.nodes[it[1]].args):
                if not _leaf(d, a):
                    cons[a].append((k, "u"))
    lds_read = {}   # node -> last item reading it from LDS
    reg_read = {}   # (node, item) uniform reads from register
    in_lds = set()
    for n, (kind, k0) in prog.where.items():
        readers = cons.get(n, [])
        if kind == "v":
            in_lds.add(n)
            lds_read[n] = max([k for k, _ in readers] + [k0])
        else:
            far = [k for k, kk in readers if kk == "v" or k - k0 > prog.uniform_window]
            if far or n in outputs:
                in_lds.add(n)
                lds_read[n] = max(far + [k0])
            for k, kk in readers:
                if kk == "u" and k - k0 <= prog.uniform_window:
                    reg_read[(n, k)] = True
    prog.reg_read = reg_read
    # pinned output homes
    addr = {}
    extra = []  # (src node or leaf, dst addr)
    for n, dsts in outputs.items():
        if _leaf(d, n) or n not in prog.where:
            for a in dsts:
                extra.append((n, a))
            continue
        addr[n] = dsts[0]
        for a in dsts[1:]:
            extra.append((n, a))
    # interval colouring of the rest over item indices
    births = collections.defaultdict(list)
    for n in in_lds:
        if n not in addr:
            births[prog.where[n][1]].append(n)
    deaths = collections.defaultdict(list)
    for n in in_lds:
        if n not in addr:
            deaths[lds_read[n]].append(n)
    # Booleans (32-bit words) and numbers (doubles) draw slots from separate
    # pools, so no LDS byte is ever accessed with two types.  The HIP
    # compiler's type-based alias analysis takes a 32-bit load and a 64-bit
    # store (or the reverse) as independent: with a slot reused across types,
    # it may move the store of a slot's new value above the last load of its
    # old value.
    free = {False: [], True: []}
    count = {False: 0, True: 0}
    slot = {}
    for k in range(K):
        for n in births.get(k, ()):
            b = is_bool(d, n)
            pool = free[b]
            if pool:
                s = pool.pop()
            else:
                s = count[b]
                count[b] += 1
            slot[n] = (b, s)
        for n in deaths.get(k, ()):
            if not NO_REUSE:
                free[slot[n][0]].append(slot[n][1])
        for pool in free.values():
            pool.sort(reverse=True)
    prog.slot = slot
    prog.nslots_num = count[False]
    prog.nslots_bool = count[True]
    prog.nslots = count[False] + count[True]
    prog.outputs_home = addr
    prog.extra = extra
    prog.in_lds = in_lds
    return prog


def constants_of(progs):
    pats = []
    seen = set()
    for prog in progs:
        d = prog.d
        for it in prog.items:
            if it[0] != "v":
                continue
            for n in it[2]:
                for a in d.nodes[n].args:
                    node = d.nodes[a]
                    if node.op == "const":
                        b = node.val
                    elif node.op == "bconst":
                        b = 1 if node.val else 0
                    else:
                        continue
                    if b not in seen:
                        seen.add(b)
                        pats.append(b)
        for (n, _) in prog.extra:
            node = d.nodes[n]
            if node.op in ("const", "bconst"):
                b = node.val if node.op == "const" else (1 if node.val else 0)
                if b not in seen:
                    seen.add(b)
                    pats.append(b)
    one = _bits(1.0)
    if one not in seen:
        pats.append(one)
    return pats


class Emitted:
    pass


def finalize(prog, lay, outputs_addr):
    """Byte address of every operand kind, table words of every vector step."""
    d = prog.d
    cidx = {b: i for i, b in enumerate(lay.consts)}
    one = lay.const + 8 * cidx[_bits(1.0)]

    def a_of(n):
        node = d.nodes[n]
        if node.op == "const":
            return lay.const + 8 * cidx[node.val]
        if node.op == "bconst":
            return lay.const + 8 * cidx[1 if node.val else 0]
        if node.op == "T":
            return lay.inp
        if node.op == "X":
            return lay.inp + 8 * (1 + node.val)
        if n in prog.outputs_home:
            return outputs_addr(prog.outputs_home[n])
        b, s = prog.slot[n]
        return (lay.valb if b else lay.val) + 8 * s

    prog.a_of = a_of
    tab = []
    bases = []
    for it in prog.items:
        if it[0] != "v":
            bases.append(None)
            continue
        cls, take = it[1], it[2]
        W = WORDS[cls]
        words = [[0] * LANES for _ in range(W)]
        for l in range(LANES):
            if l < len(take):
                n = take[l]
                node = d.nodes[n]
                args = [a_of(a) for a in node.args]
                dst = a_of(n)
                if cls == "add":
                    w = [args[0], args[1] | (FLIP if node.op == "sub" else 0), dst]
                elif cls == "sgn":
                    w = [args[0] | (FLIP if node.op == "neg" else CLEAR), dst]
                elif cls in ("sel", "selb"):
                    w = [args[0], args[1], args[2], dst]
                else:
                    w = args + [dst]
            else:
                w = [one] * (W - 1) + [lay.trash]
                if cls in ("sel", "selb"):
                    w = [lay.const + 8 * cidx.get(0, cidx[_bits(1.0)]), one, one, lay.trash]
            for i in range(W):
                words[i][l] = w[i]
        bases.append(len(tab))
        tab.extend(words)
    prog.bases = bases
    prog.table = tab   # list of W-word rows, each LANES long
    return prog


# ---------------------------------------------------------------------------
# simulation (the semantics the emitted code implements)
# ---------------------------------------------------------------------------

def simulate(prog, lay, outputs_addr, T, X, exp, log, sqrt):
    """Run the program on a byte-addressed LDS image (Python floats)."""
    import math
    d = prog.d
    mem = {}
    for i, b in enumerate(lay.consts):
        mem[lay.const + 8 * i] = ("b", b)
    mem[lay.inp] = ("d", T)
    for i in range(14):
        mem[lay.inp + 8 * (1 + i)] = ("d", X[i])

    def rd(a):
        a &= ADDR_MASK
        kind, v = mem[a]
        if kind == "b":
            return struct.unpack("<d", struct.pack("<Q", v))[0]
        return v

    def rb(a):
        a &= ADDR_MASK
        kind, v = mem[a]
        if kind == "b":
            return v & 0xFFFFFFFF
        if kind == "u":
            return v
        raise AssertionError("bool read of a double slot")

    def flip(x, w):
        return -x if w & FLIP else x

    regs = {}
    for k, it in enumerate(prog.items):
        if it[0] == "m":
            base = prog.bases[k]
            src, dst = prog.table[base], prog.table[base + 1]
            for l in range(LANES):
                mem[dst[l] & ADDR_MASK] = mem[src[l] & ADDR_MASK]
            continue
        if it[0] == "u":
            n = it[1]
            node = d.nodes[n]

            def val(a, b_=False):
                an = d.nodes[a]
                if an.op == "const":
                    return d.cval(a)
                if an.op == "bconst":
                    return 1 if an.val else 0
                if an.op == "T":
                    return T
                if an.op == "X":
                    return X[an.val]
                if (a, k) in prog.reg_read:
                    return regs[a]
                return rb(prog.a_of(a)) if b_ else rd(prog.a_of(a))
            isb = [is_bool(d, a) for a in node.args]
            v = [val(a, b) for a, b in zip(node.args, isb)]
            r = _apply(node.op, v, exp, log, sqrt)
            regs[n] = r
            if n in prog.in_lds:
                mem[prog.a_of(n)] = ("u", r) if is_bool(d, n) else ("d", r)
            continue
        cls, take = it[1], it[2]
        base = prog.bases[k]
        rows = prog.table[base:base + WORDS[cls]]
        res = []
        for l in range(LANES):
            w = [row[l] for row in rows]
            if cls == "add":
                r = ("d", flip(rd(w[0]), w[0]) + flip(rd(w[1]), w[1]))
            elif cls == "mul":
                r = ("d", rd(w[0]) * rd(w[1]))
            elif cls == "div":
                r = ("d", _apply("div", [rd(w[0]), rd(w[1])], exp, log, sqrt))
            elif cls in ("exp", "log", "sqrt"):
                r = ("d", _apply(cls, [rd(w[0])], exp, log, sqrt))
            elif cls == "sgn":
                x = rd(w[0])
                r = ("d", -x if w[0] & FLIP else math.fabs(x))
            elif cls in ("ge", "gt", "eq", "ne"):
                r = ("u", int(_apply(cls, [rd(w[0]), rd(w[1])], exp, log, sqrt)))
            elif cls in ("and", "or"):
                a, b = rb(w[0]), rb(w[1])
                r = ("u", (a & b) if cls == "and" else (a | b))
            elif cls == "not":
                r = ("u", rb(w[0]) ^ 1)
            elif cls == "sel":
                r = ("d", rd(w[1]) if rb(w[0]) else rd(w[2]))
            elif cls == "selb":
                r = ("u", rb(w[1]) if rb(w[0]) else rb(w[2]))
            else:
                raise ValueError(cls)
            res.append((w[-1] & ADDR_MASK, r))
        for a, r in res:
            mem[a] = r
    for (n, dst) in ([] if getattr(prog, "moves_appended", False) else prog.extra):
        node = d.nodes[n]
        if node.op == "const":
            mem[outputs_addr(dst)] = ("d", d.cval(n))
        elif node.op == "X":
            mem[outputs_addr(dst)] = ("d", X[node.val])
        elif node.op == "T":
            mem[outputs_addr(dst)] = ("d", T)
        else:
            mem[outputs_addr(dst)] = mem[prog.a_of(n)]
    return mem


def _apply(op, v, exp, log, sqrt):
    import math
    if op == "add":
        return v[0] + v[1]
    if op == "sub":
        return v[0] - v[1]
    if op == "mul":
        return v[0] * v[1]
    if op == "div":
        if v[1] == 0:
            if v[0] == 0 or v[0] != v[0]:
                return math.nan
            return math.copysign(math.inf, v[0]) * math.copysign(1.0, v[1])
        return v[0] / v[1]
    if op == "exp":
        return exp(v[0])
    if op == "log":
        return log(v[0])
    if op == "sqrt":
        return sqrt(v[0])
    if op == "neg":
        return -v[0]
    if op == "abs":
        return math.fabs(v[0])
    if op == "ge":
        return v[0] >= v[1]
    if op == "gt":
        return v[0] > v[1]
    if op == "eq":
        return v[0] == v[1]
    if op == "ne":
        return v[0] != v[1]
    if op == "and":
        return int(bool(v[0]) and bool(v[1]))
    if op == "or":
        return int(bool(v[0]) or bool(v[1]))
    if op == "not":
        return int(not v[0])
    if op == "sel":
        return v[1] if v[0] else v[2]
    raise ValueError(op)


# ---------------------------------------------------------------------------
# emission
# ---------------------------------------------------------------------------

def _lit(v):
    if v != v:
        return "(0.0/0.0)"
    if v in (float("inf"), float("-inf")):
        return "(1.0/0.0)" if v > 0 else "(-1.0/0.0)"
    s = v.hex()
    return f"({s})" if s.startswith("-") else s


def _append_moves(prog, lay):
    """Extra output locations (leaves, repeated outputs) as vector copy steps
    appended after the program."""
    if not prog.extra:
        return
    d = prog.d
    prog.moves_appended = True
    srcs = [(prog.a_of(n), dst) for n, dst in prog.extra]
    for i in range(0, len(srcs), LANES):
        chunk = srcs[i:i + LANES]
        rows = [[s for s, _ in chunk] + [prog.a_of_one] * (LANES - len(chunk)),
                [dd for _, dd in chunk] + [lay.trash] * (LANES - len(chunk))]
        prog.bases.append(len(prog.table))
        prog.table.extend(rows)
        prog.items.append(("m", "mov", len(chunk)))


def emit_function(prog, lay, fname, tabname, outputs_addr, prefetch=8):
    """Straight-line code of one program (see the module doc for the macros)."""
    d = prog.d
    cidx = {b: i for i, b in enumerate(lay.consts)}
    prog.a_of_one = lay.const + 8 * cidx[_bits(1.0)]
    prog.extra = [(n, outputs_addr(dst)) for n, dst in prog.extra]
    _append_moves(prog, lay)
    items = prog.items
    K = len(items)
    L = [f"COOP_FN void {fname}(COOP_PARAMS)", "{"]
    vidx = [k for k, it in enumerate(items) if it[0] in ("v", "m")]

    def words(k):
        it = items[k]
        return WORDS[it[1]]

    def tw(k):
        w = words(k)
        names = [f"t{k}_{i}" for i in range(w)]
        out = [f"    COOP_DU({n});" for n in names]
        body = " ".join(f"COOP_V({n}) = COOP_TW({tabname}, {prog.bases[k] + i});"
                        for i, n in enumerate(names))
        out.append(f"    COOP_L({body})")
        return out

    def ld(k):
        it = items[k]
        cls = it[1]
        t = lambda i: f"COOP_V(t{k}_{i})"
        M = f"{ADDR_MASK:#x}u"
        # only the second word of an add (subtract flag) and the word of a
        # sign op carry flags above the address bits
        if cls == "add":
            decl = [f"COOP_DD(a{k})", f"COOP_DD(b{k})"]
            body = (f"COOP_V(a{k}) = COOP_LD({t(0)}); COOP_V(b{k}) = COOP_LD({t(1)} & {M});")
        elif cls in ("mul", "div", "ge", "gt", "eq", "ne"):
            decl = [f"COOP_DD(a{k})", f"COOP_DD(b{k})"]
            body = (f"COOP_V(a{k}) = COOP_LD({t(0)}); COOP_V(b{k}) = COOP_LD({t(1)});")
        elif cls == "sgn":
            decl = [f"COOP_DD(a{k})"]
            body = f"COOP_V(a{k}) = COOP_LD({t(0)} & {M});"
        elif cls in ("exp", "log", "sqrt", "mov"):
            decl = [f"COOP_DD(a{k})"]
            body = f"COOP_V(a{k}) = COOP_LD({t(0)});"
        elif cls in ("and", "or"):
            decl = [f"COOP_DU(a{k})", f"COOP_DU(b{k})"]
            body = f"COOP_V(a{k}) = COOP_LDU({t(0)}); COOP_V(b{k}) = COOP_LDU({t(1)});"
        elif cls == "not":
            decl = [f"COOP_DU(a{k})"]
            body = f"COOP_V(a{k}) = COOP_LDU({t(0)});"
        elif cls == "sel":
            decl = [f"COOP_DU(c{k})", f"COOP_DD(a{k})", f"COOP_DD(b{k})"]
            body = (f"COOP_V(c{k}) = COOP_LDU({t(0)}); COOP_V(a{k}) = COOP_LD({t(1)}); "
                    f"COOP_V(b{k}) = COOP_LD({t(2)});")
        elif cls == "selb":
            decl = [f"COOP_DU(c{k})", f"COOP_DU(a{k})", f"COOP_DU(b{k})"]
            body = (f"COOP_V(c{k}) = COOP_LDU({t(0)}); COOP_V(a{k}) = COOP_LDU({t(1)}); "
                    f"COOP_V(b{k}) = COOP_LDU({t(2)});")
        else:
            raise ValueError(cls)
        return [f"    {x};" for x in decl] + [f"    COOP_L({body})"]

    def cs(k):
        it = items[k]
        cls = it[1]
        a, b, c = f"COOP_V(a{k})", f"COOP_V(b{k})", f"COOP_V(c{k})"
        t = lambda i: f"COOP_V(t{k}_{i})"
        dst = t(WORDS[cls] - 1)
        if cls == "add":
            expr = f"coop_flip({a}, {t(0)}) + coop_flip({b}, {t(1)})"
        elif cls == "mul":
            expr = f"{a} * {b}"
        elif cls == "div":
            expr = f"{a} / {b}"
        elif cls == "exp":
            expr = f"CHEM_EXP({a})"
        elif cls == "log":
            expr = f"CHEM_LOG({a})"
        elif cls == "sqrt":
            expr = f"sqrt({a})"
        elif cls == "sgn":
            expr = f"coop_sgn({a}, {t(0)})"
        elif cls == "mov":
            expr = a
        elif cls in ("ge", "gt", "eq", "ne"):
            op = {"ge": ">=", "gt": ">", "eq": "==", "ne": "!="}[cls]
            return [f"    COOP_L(COOP_STU({dst}, ({a} {op} {b}) ? 1u : 0u);)"]
        elif cls in ("and", "or"):
            op = "&" if cls == "and" else "|"
            return [f"    COOP_L(COOP_STU({dst}, {a} {op} {b});)"]
        elif cls == "not":
            return [f"    COOP_L(COOP_STU({dst}, {a} ^ 1u);)"]
        elif cls == "sel":
            expr = f"{c} ? {a} : {b}"
        elif cls == "selb":
            return [f"    COOP_L(COOP_STU({dst}, {c} ? {a} : {b});)"]
        else:
            raise ValueError(cls)
        return [f"    COOP_L(COOP_ST({dst}, {expr});)"]

    def uop(k):
        n = items[k][1]
        node = d.nodes[n]
        isb = is_bool(d, n)

        def ref(a):
            an = d.nodes[a]
            if an.op == "const":
                return _lit(d.cval(a))
            if an.op == "bconst":
                return "1u" if an.val else "0u"
            if an.op in ("T", "X"):
                return f"COOP_LD({prog.a_of(a)})"
            if (a, k) in prog.reg_read:
                return f"u{a}"
            if is_bool(d, a):
                return f"COOP_LDU({prog.a_of(a)})"
            return f"COOP_LD({prog.a_of(a)})"
        r = [ref(a) for a in node.args]
        op = node.op
        if op in ("add", "sub", "mul", "div"):
            e = f"{r[0]} {dict(add='+', sub='-', mul='*', div='/')[op]} {r[1]}"
        elif op in ("exp", "log"):
            e = f"CHEM_{op.upper()}({r[0]})"
        elif op == "sqrt":
            e = f"sqrt({r[0]})"
        elif op == "neg":
            e = f"-{r[0]}"
        elif op == "abs":
            e = f"fabs({r[0]})"
        elif op in ("ge", "gt", "eq", "ne"):
            e = f"({r[0]} {dict(ge='>=', gt='>', eq='==', ne='!=')[op]} {r[1]}) ? 1u : 0u"
        elif op == "and":
            e = f"{r[0]} & {r[1]}"
        elif op == "or":
            e = f"{r[0]} | {r[1]}"
        elif op == "not":
            e = f"{r[0]} ^ 1u"
        elif op == "sel":
            e = f"{r[0]} ? {r[1]} : {r[2]}"
        else:
            raise ValueError(op)
        ty = "uint32_t" if isb else "double"
        out = [f"    const {ty} u{n} = {e};"]
        if n in prog.in_lds:
            out.append(f"    COOP_{'STU' if isb else 'ST'}({prog.a_of(n)}u, u{n});")
        return out

    # pipelined order: table words `prefetch` vector items ahead; the LDS
    # loads of item k+1 (a vector step) are issued just before the code of
    # item k, whatever its kind -- item k+1 never reads what item k writes
    # (the scheduler's rule) -- and the copies' loads after the last store.
    vpos = {k: i for i, k in enumerate(vidx)}
    done_tw = set()

    def ensure_tw(k):
        if k not in done_tw:
            done_tw.add(k)
            L.extend(tw(k))

    for k in vidx[:prefetch]:
        ensure_tw(k)
    if K and items[0][0] == "v":
        L += ld(0)
    for k in range(K):
        it = items[k]
        if it[0] in ("v", "m") and vpos[k] + prefetch < len(vidx):
            ensure_tw(vidx[vpos[k] + prefetch])
        nxt = items[k + 1] if k + 1 < K else None
        if nxt is not None and nxt[0] == "v":
            ensure_tw(k + 1)
            L += ld(k + 1)
        L.append(f"    /* item {k}: {it[0]} {it[1] if it[0] != 'u' else d.nodes[it[1]].op} */")
        L += cs(k) if it[0] in ("v", "m") else uop(k)
        if TRACE:
            L.append(f"    COOP_TRACE({k});")
        if nxt is not None and nxt[0] == "m":
            ensure_tw(k + 1)
            L += ld(k + 1)
    L.append("}")
    return "\n".join(L)


OUT_ADDR = {"J": lambda lay, i: lay.jac + 8 * i,
            "f0": lambda lay, i: lay.vec + 8 * (VEC["f0"] + i),
            "dy": lambda lay, i: lay.vec + 8 * (VEC["dy"] + i)}


def build_programs(d0):
    """The fused J + f program (outputs J and f0) and the f program (output
    dy), scheduled, with slots, constants and tables.  Returns (progs, layout)."""
    f_roots = [d0.outputs[("ydot", i)] for i in range(dagmod.NUM_SPEC)] + [d0.outputs[("edot",)]]
    j_roots = [d0.outputs[("jac", r, c)] for r in range(dagmod.NEQS) for c in range(dagmod.NEQS)]
    progs = []
    for name, roots, fkey in (("fused", f_roots + j_roots, "f0"), ("rhs", f_roots, "dy")):
        d, rr = rebalance(d0, roots)
        prog = schedule(d, rr, name)
        outs = {}
        for i, r in enumerate(rr[:15]):
            outs.setdefault(r, []).append((fkey, i))
        for k, r in enumerate(rr[15:]):
            outs.setdefault(r, []).append(("J", k))
        assign(prog, outs, None)
        progs.append(prog)
    lay = Layout(constants_of(progs), max(p.nslots_num for p in progs),
                 max(p.nslots_bool for p in progs))
    oaddr = lambda key: OUT_ADDR[key[0]](lay, key[1])
    for prog in progs:
        finalize(prog, lay, oaddr)
    return progs, lay, oaddr


def emit_header(ir_path, out_path, json_path=None):
    import json
    import os
    import emit as emitmod
    d0 = emitmod.specialise(dagmod.read_ir(ir_path))
    progs, lay, oaddr = build_programs(d0)
    funcs, tabs, stats = [], [], {}
    for prog in progs:
        tab = f"coop_tab_{prog.name}"
        nv = sum(1 for it in prog.items if it[0] == "v")
        nu = sum(1 for it in prog.items if it[0] == "u")
        code = emit_function(prog, lay, f"coop_{prog.name}", tab, oaddr)
        funcs.append(code)
        flat = [w for row in prog.table for w in row]
        body = ",\n".join(", ".join(f"{w:#x}" for w in flat[i:i + 16]) for i in range(0, len(flat), 16))
        tabs.append(f"COOP_DATA uint32_t {tab}[{len(flat)}] = {{\n{body}\n}};")
        stats[prog.name] = {"items": len(prog.items), "vector_steps": nv, "uniform_ops": nu,
                            "copy_steps": len(prog.items) - nv - nu, "value_slots": prog.nslots,
                            "table_words": len(flat), "model_cycles": prog.cycles}
    consts = ", ".join(f"{b:#x}ull" for b in lay.consts)
    hdr = [
        "/* Generated by warp-reactor_amd/gen/coop.py from primordial_network.ir. Do not edit.",
        " * Wave-cooperative programs of the network: coop_fused (J and f at the state in",
        " * the LDS input slots) and coop_rhs (f).  The includer defines COOP_FN, COOP_PARAMS,",
        " * COOP_DATA, COOP_DU/DD (lane variable declarations), COOP_V (lane variable),",
        " * COOP_L (a lane statement), COOP_TW (table word), COOP_LD/LDU/ST/STU (LDS access",
        " * at a byte address), coop_flip, coop_sgn, CHEM_EXP and CHEM_LOG. */",
        "#ifndef COOP_LAYOUT_DEFINED",
        "#define COOP_LAYOUT_DEFINED",
        f"#define COOP_OFF_CONST {lay.const}",
        f"#define COOP_OFF_TRASH {lay.trash}",
        f"#define COOP_OFF_IN {lay.inp}",
        f"#define COOP_OFF_VAL {lay.val}",
        f"#define COOP_OFF_J {lay.jac}",
        f"#define COOP_OFF_VEC {lay.vec}",
        f"#define COOP_OFF_LU {lay.lu}",
        f"#define COOP_OFF_COL {lay.col}",
        f"#define COOP_LDS_BYTES {lay.total}",
        f"#define COOP_NCONST {len(lay.consts)}",
        f"#define COOP_ADDR_MASK {ADDR_MASK:#x}u",
        "#endif /* COOP_LAYOUT_DEFINED */",
        "/* with COOP_LAYOUT_ONLY defined, the header stops after the layout */",
        "#ifndef COOP_LAYOUT_ONLY",
        "#ifndef COOP_CODE_DEFINED",
        "#define COOP_CODE_DEFINED",
        f"COOP_DATA uint64_t coop_const_bits[{len(lay.consts)}] = {{{consts}}};",
    ] + tabs + funcs + ["#endif /* COOP_CODE_DEFINED */", "#endif /* COOP_LAYOUT_ONLY */"]
    text = "\n".join(hdr) + "\n"
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    tmp = out_path + ".tmp"
    with open(tmp, "w") as f:
        f.write(text)
    os.replace(tmp, out_path)
    stats["lds_bytes"] = lay.total
    stats["constants"] = len(lay.consts)
    if json_path:
        with open(json_path, "w") as f:
            json.dump(stats, f, indent=1, sort_keys=True)
    return stats


def main(argv):
    import os
    global NO_REUSE
    global TRACE
    if "--trace" in argv:
        TRACE = True
        argv = [a for a in argv if a != "--trace"]
    if "--no-reuse" in argv:
        NO_REUSE = True
        argv = [a for a in argv if a != "--no-reuse"]
    here = os.path.dirname(os.path.abspath(__file__))
    stats = emit_header(os.path.join(here, "primordial_network.ir"), argv[1],
                        argv[2] if len(argv) > 2 else None)
    print("coop:", stats)


if __name__ == "__main__":
    import sys
    main(sys.argv)
