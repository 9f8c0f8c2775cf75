"""Expression DAG of the primordial network: parser, hash-consed IR, IR file I/O.

The network's right-hand side (14 species rates + specific-energy rate) and its
15x15 analytic Jacobian are machine-generated straight-line expressions in the
reference (``reproducer.mojo:388-2321``; the same expressions in C++ at
``reproducer.cpp:341-6994``).  This module turns that text into a DAG of
IEEE-754 double operations with a fixed evaluation order, so that the code
emitted from it (CPU C for the oracle, HIP for gfx950) rounds every operation
exactly as the reference's compiled C++ does.

Grammar of the source (the subset ``tools/slice_dag.py:92-217`` accepts):
``var xK_V = <expr>``, ``vset(ydot, i, <expr>)``, ``return <expr>`` and
``mset(jac, r, c, <expr>)``; expressions use ``portable_mul(a, b)`` for a
separately rounded product, ``+ - /``, unary ``-``, ``exp log abs
portable_sqrt powi_m3 powi_m5 powi_m7``, comparisons, ``and or``,
``a if truthy(c) else b`` and ``vget(X, i)``.

Transformations applied while building the DAG are exactly those that cannot
change a single bit of any result:
  * structural hashing (common-subexpression elimination) across all three
    functions -- the same operation on the same operands rounds identically;
  * folding of ``+ - * /``, negation, ``abs`` and ``sqrt`` on constants (all
    correctly rounded in IEEE-754, so Python's float arithmetic reproduces
    them), and of ``exp``/``log`` of constants with the host libm, which is what
    the reference's CPU build evaluates;
  * ``x*1 -> x``, ``x/1 -> x``, ``x-0 -> x``, selects on constant conditions.
"""
from __future__ import annotations

import hashlib
import math
import re
import struct
from dataclasses import dataclass, field

NUM_SPEC = 14
NEQS = 15

# Operation codes.  "f" ops produce a double, "b" ops a boolean.
FLOAT_OPS = {"const", "T", "X", "z", "in", "neg", "abs", "exp", "log", "sqrt",
             "add", "sub", "mul", "div", "sel"}
BOOL_OPS = {"bconst", "lt", "le", "gt", "ge", "eq", "ne", "and", "or", "not"}
ARITY = {"neg": 1, "abs": 1, "exp": 1, "log": 1, "sqrt": 1, "not": 1,
         "add": 2, "sub": 2, "mul": 2, "div": 2, "lt": 2, "le": 2, "gt": 2,
         "ge": 2, "eq": 2, "ne": 2, "and": 2, "or": 2, "sel": 3}

# Pinned manifest of the source functions: (definitions, outputs)
# (tools/slice_dag.py:26-30).
MANIFEST = {"rhs_specie": (119, 14), "rhs_eint": (162, 1), "jac_nuc": (1394, 225)}


@dataclass(frozen=True)
class Node:
    op: str
    args: tuple = ()
    val: object = None  # float for const, int index for X, bool for bconst


@dataclass
class Dag:
    nodes: list = field(default_factory=list)      # list[Node], topological
    index: dict = field(default_factory=dict)      # Node -> id
    outputs: dict = field(default_factory=dict)    # ("ydot", i) / ("edot",) / ("jac", r, c) -> id
    meta: dict = field(default_factory=dict)

    # -- construction with hashing and safe folding -------------------------
    def add(self, op, args=(), val=None):
        if op == "const":
            val = float(val)
            key = Node(op, (), struct.unpack("<Q", struct.pack("<d", val))[0])
        else:
            key = Node(op, tuple(args), val)
        got = self.index.get(key)
        if got is not None:
            return got
        folded = self._fold(op, args, val)
        if folded is not None:
            return folded
        nid = len(self.nodes)
        self.nodes.append(key)
        self.index[key] = nid
        return nid

    def const(self, v):
        return self.add("const", (), v)

    def cval(self, nid):
        n = self.nodes[nid]
        if n.op == "const":
            return struct.unpack("<d", struct.pack("<Q", n.val))[0]
        return None

    def bval(self, nid):
        n = self.nodes[nid]
        return n.val if n.op == "bconst" else None

    def _fold(self, op, args, val):
        if op in ("const", "T", "X", "z", "bconst", "in"):
            return None
        cv = [self.cval(a) for a in args] if op not in ("and", "or", "not", "sel") else None
        if cv is not None and all(c is not None for c in cv):
            a = cv
            try:
                if op == "neg":
                    return self.const(-a[0])
                if op == "abs":
                    return self.const(abs(a[0]))
                if op == "sqrt":
                    return self.const(math.sqrt(a[0]))
                if op == "exp":
                    return self.const(math.exp(a[0]))
                if op == "log":
                    return self.const(math.log(a[0]))
                if op == "add":
                    return self.const(a[0] + a[1])
                if op == "sub":
                    return self.const(a[0] - a[1])
                if op == "mul":
                    return self.const(a[0] * a[1])
                if op == "div":
                    return self.const(a[0] / a[1])
                if op in ("lt", "le", "gt", "ge", "eq", "ne"):
                    r = {"lt": a[0] < a[1], "le": a[0] <= a[1], "gt": a[0] > a[1],
                         "ge": a[0] >= a[1], "eq": a[0] == a[1], "ne": a[0] != a[1]}[op]
                    return self.add("bconst", (), bool(r))
            except (ValueError, ZeroDivisionError, OverflowError):
                return None
        if op == "mul":
            if self.cval(args[1]) == 1.0:
                return args[0]
            if self.cval(args[0]) == 1.0:
                return args[1]
        if op == "div" and self.cval(args[1]) == 1.0:
            return args[0]
        if op == "sub" and self.cval(args[1]) == 0.0 and \
                math.copysign(1.0, self.cval(args[1])) > 0:
            return args[0]
        if op == "sel":
            b = self.bval(args[0])
            if b is not None:
                return args[1] if b else args[2]
            if args[1] == args[2]:
                return args[1]
        if op in ("and", "or"):
            b0, b1 = self.bval(args[0]), self.bval(args[1])
            if op == "and":
                if b0 is False or b1 is False:
                    return self.add("bconst", (), False)
                if b0 is True:
                    return args[1]
                if b1 is True:
                    return args[0]
            else:
                if b0 is True or b1 is True:
                    return self.add("bconst", (), True)
                if b0 is False:
                    return args[1]
                if b1 is False:
                    return args[0]
        if op == "not":
            b = self.bval(args[0])
            if b is not None:
                return self.add("bconst", (), not b)
        return None

    # -- queries --------------------------------------------------------------
    def is_bool(self, nid):
        return self.nodes[nid].op in BOOL_OPS

    def live_nodes(self, roots=None):
        """Ids reachable from the given output ids (default: all outputs)."""
        if roots is None:
            roots = list(self.outputs.values())
        seen = set()
        stack = list(roots)
        while stack:
            n = stack.pop()
            if n in seen:
                continue
            seen.add(n)
            stack.extend(self.nodes[n].args)
        return seen

    def op_counts(self, roots=None):
        live = self.live_nodes(roots)
        counts = {}
        for nid in live:
            op = self.nodes[nid].op
            counts[op] = counts.get(op, 0) + 1
        return counts


# ---------------------------------------------------------------------------
# Source parser
# ---------------------------------------------------------------------------

_TOKEN_RE = re.compile(r"""
    (?P<num>\d+\.\d*(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+|\d+)
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op><=|>=|==|!=|[-+/(),<>])
  | (?P<ws>\s+)
""", re.VERBOSE)


def _tokenize(text):
    toks = []
    pos = 0
    while pos < len(text):
        m = _TOKEN_RE.match(text, pos)
        if not m:
            raise SyntaxError(f"bad character at {text[pos:pos + 30]!r}")
        pos = m.end()
        if m.lastgroup == "ws":
            continue
        toks.append((m.lastgroup, m.group(m.lastgroup)))
    return toks


class _ExprParser:
    """Recursive-descent parser for one right-hand side, Python precedence."""

    def __init__(self, dag, env, text):
        self.dag = dag
        self.env = env
        self.toks = _tokenize(text)
        self.i = 0

    def peek(self, k=0):
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else (None, None)

    def take(self, value=None):
        tok = self.peek()
        if value is not None and tok[1] != value:
            raise SyntaxError(f"expected {value!r}, got {tok[1]!r}")
        self.i += 1
        return tok

    def parse(self):
        e = self.ternary()
        if self.i != len(self.toks):
            raise SyntaxError(f"trailing tokens {self.toks[self.i:self.i + 5]}")
        return e

    def ternary(self):
        a = self.lor()
        if self.peek()[1] == "if":
            self.take("if")
            c = self.lor()
            self.take("else")
            b = self.ternary()
            return self.dag.add("sel", (self._as_bool(c), a, b))
        return a

    def _as_bool(self, nid):
        if self.dag.is_bool(nid):
            return nid
        # truthy(Float64) is x != 0.0 (reproducer.mojo:75-76)
        return self.dag.add("ne", (nid, self.dag.const(0.0)))

    def lor(self):
        a = self.land()
        while self.peek()[1] == "or":
            self.take()
            b = self.land()
            a = self.dag.add("or", (self._as_bool(a), self._as_bool(b)))
        return a

    def land(self):
        a = self.lnot()
        while self.peek()[1] == "and":
            self.take()
            b = self.lnot()
            a = self.dag.add("and", (self._as_bool(a), self._as_bool(b)))
        return a

    def lnot(self):
        if self.peek()[1] == "not":
            self.take()
            return self.dag.add("not", (self._as_bool(self.lnot()),))
        return self.cmp()

    def cmp(self):
        a = self.arith()
        t = self.peek()[1]
        ops = {"<": "lt", "<=": "le", ">": "gt", ">=": "ge", "==": "eq", "!=": "ne"}
        if t in ops:
            self.take()
            b = self.arith()
            return self.dag.add(ops[t], (a, b))
        return a

    def arith(self):
        a = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            b = self.term()
            a = self.dag.add("add" if op == "+" else "sub", (a, b))
        return a

    def term(self):
        a = self.unary()
        while self.peek()[1] == "/":
            self.take()
            b = self.unary()
            a = self.dag.add("div", (a, b))
        return a

    def unary(self):
        if self.peek()[1] == "-":
            self.take()
            return self.dag.add("neg", (self.unary(),))
        if self.peek()[1] == "+":
            self.take()
            return self.unary()
        return self.primary()

    def args(self):
        self.take("(")
        out = [self.ternary()]
        while self.peek()[1] == ",":
            self.take(",")
            out.append(self.ternary())
        self.take(")")
        return out

    def primary(self):
        kind, tok = self.peek()
        if kind == "num":
            self.take()
            return self.dag.const(float(tok))
        if tok == "(":
            self.take("(")
            e = self.ternary()
            self.take(")")
            return e
        if kind == "name":
            if tok in ("True", "False"):
                self.take()
                return self.dag.add("bconst", (), tok == "True")
            if self.peek(1)[1] == "(":
                return self.call()
            self.take()
            if tok not in self.env:
                raise SyntaxError(f"use before definition: {tok}")
            return self.env[tok]
        raise SyntaxError(f"unexpected token {tok!r}")

    def call(self):
        name = self.take()[1]
        d = self.dag
        if name == "vget":
            self.take("(")
            vec = self.take()[1]
            self.take(",")
            idx = int(self.take()[1])
            self.take(")")
            if vec != "X" or not 0 <= idx < NUM_SPEC:
                raise SyntaxError(f"unsupported vget({vec}, {idx})")
            return self.env[("X", idx)]
        a = self.args()
        if name == "portable_mul":
            return d.add("mul", (a[0], a[1]))
        if name in ("exp", "log", "abs"):
            return d.add(name, (a[0],))
        if name == "portable_sqrt":
            return d.add("sqrt", (a[0],))
        if name == "truthy":
            return self._as_bool(a[0])
        # math::powi<-n> (reproducer.cpp:64-72; reproducer.mojo:79-89)
        x = a[0]
        if name == "powi_m3":
            return d.add("div", (d.const(1.0), d.add("mul", (x, d.add("mul", (x, x))))))
        if name == "powi_m5":
            x2 = d.add("mul", (x, x))
            return d.add("div", (d.const(1.0), d.add("mul", (x, d.add("mul", (x2, x2))))))
        if name == "powi_m7":
            x3 = d.add("mul", (x, d.add("mul", (x, x))))
            return d.add("div", (d.const(1.0), d.add("mul", (x, d.add("mul", (x3, x3))))))
        raise SyntaxError(f"unsupported call {name}")


_DEF_RE = re.compile(r"^var (x\d+_\d+) = (.*)$")
_VSET_RE = re.compile(r"^vset\(ydot, (\d+), (.*)\)$")
_MSET_RE = re.compile(r"^mset\(jac, (\d+), (\d+), (.*)\)$")
_RET_RE = re.compile(r"^return (.*)$")


def _function_lines(text, name):
    lines = text.split("\n")
    start = None
    for i, line in enumerate(lines):
        if line.startswith(f"def {name}("):
            start = i
            break
    if start is None:
        raise SyntaxError(f"function {name} not found")
    body = []
    pending, depth = "", 0
    for line in lines[start + 1:]:
        if line and not line.startswith(" ") and depth == 0:
            break
        if not line.strip():
            continue
        # a statement continues on the next lines while its parentheses are open
        pending = (pending + " " + line.strip()) if pending else line.strip()
        depth += line.count("(") - line.count(")")
        if depth <= 0:
            body.append((start + 1 + len(body), pending))
            pending, depth = "", 0
    if pending:
        raise SyntaxError(f"unbalanced parentheses in {name}: {pending[:80]}")
    return start + 1, body


def parse_network(text, dag=None):
    """Parse rhs_specie, rhs_eint and jac_nuc from the Mojo source text."""
    dag = dag or Dag()
    # Named constants of the source (reproducer.mojo:17-18): ``Pi`` and
    # ``Log10`` are M_PI and M_LN10 in the C++ text (e.g. reproducer.cpp:457).
    base = {"T": dag.add("T"), "z": dag.add("z"),
            "Pi": dag.const(float("3.141592653589793238462643383279502884")),
            "Log10": dag.const(float("2.3025850929940456840179914546843642"))}
    for i in range(NUM_SPEC):
        base[("X", i)] = dag.add("X", (), i)
    counts = {}
    for fname in ("rhs_specie", "rhs_eint", "jac_nuc"):
        line0, body = _function_lines(text, fname)
        env = dict(base)
        ndef = nout = 0
        for _, stmt in body:
            if stmt == "var T = state.T":
                continue
            m = _DEF_RE.match(stmt)
            if m:
                if m.group(1) in env:
                    raise SyntaxError(f"duplicate definition {m.group(1)} in {fname}")
                env[m.group(1)] = _ExprParser(dag, env, m.group(2)).parse()
                ndef += 1
                continue
            m = _VSET_RE.match(stmt)
            if m and fname == "rhs_specie":
                key = ("ydot", int(m.group(1)))
                dag.outputs[key] = _ExprParser(dag, env, m.group(2)).parse()
                nout += 1
                continue
            m = _RET_RE.match(stmt)
            if m and fname == "rhs_eint":
                dag.outputs[("edot",)] = _ExprParser(dag, env, m.group(1)).parse()
                nout += 1
                continue
            m = _MSET_RE.match(stmt)
            if m and fname == "jac_nuc":
                key = ("jac", int(m.group(1)), int(m.group(2)))
                dag.outputs[key] = _ExprParser(dag, env, m.group(3)).parse()
                nout += 1
                continue
            raise SyntaxError(f"unsupported statement in {fname}: {stmt[:80]}")
        counts[fname] = (ndef, nout)
    if counts != MANIFEST:
        raise ValueError(f"manifest drift: {counts} != {MANIFEST}")
    for key, nid in dag.outputs.items():
        if dag.is_bool(nid):
            raise ValueError(f"boolean output {key}")
    dag.meta["manifest"] = counts
    return dag


# ---------------------------------------------------------------------------
# IR file format (own, line oriented; constants as exact hex floats)
# ---------------------------------------------------------------------------

def _compact(dag):
    """Renumber to live nodes only, keeping topological order."""
    live = dag.live_nodes()
    order = sorted(live)
    remap = {old: new for new, old in enumerate(order)}
    out = Dag(meta=dict(dag.meta))
    for old in order:
        n = dag.nodes[old]
        nn = Node(n.op, tuple(remap[a] for a in n.args), n.val)
        out.index[nn] = len(out.nodes)
        out.nodes.append(nn)
    out.outputs = {k: remap[v] for k, v in dag.outputs.items()}
    return out


def write_ir(dag, path, source_sha256):
    dag = _compact(dag)
    lines = ["# primordial network DAG IR v1",
             f"# source-sha256 {source_sha256}",
             "# manifest " + " ".join(f"{k}={v[0]}/{v[1]}" for k, v in dag.meta["manifest"].items()),
             f"# nodes {len(dag.nodes)}"]
    for nid, n in enumerate(dag.nodes):
        if n.op == "const":
            v = struct.unpack("<d", struct.pack("<Q", n.val))[0]
            lines.append(f"{nid} const {v.hex()}")
        elif n.op == "bconst":
            lines.append(f"{nid} bconst {int(n.val)}")
        elif n.op == "X":
            lines.append(f"{nid} X {n.val}")
        else:
            lines.append(f"{nid} {n.op} " + " ".join(str(a) for a in n.args))
            lines[-1] = lines[-1].rstrip()
    for key, nid in sorted(dag.outputs.items(), key=lambda kv: (kv[0][0], kv[0][1:])):
        lines.append("out " + " ".join(str(k) for k in key) + f" {nid}")
    text = "\n".join(lines) + "\n"
    with open(path, "w") as f:
        f.write(text)
    return hashlib.sha256(text.encode()).hexdigest()


def read_ir(path):
    dag = Dag()
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith("#"):
                parts = line[1:].split()
                if parts and parts[0] == "source-sha256":
                    dag.meta["source_sha256"] = parts[1]
                if parts and parts[0] == "manifest":
                    man = {}
                    for p in parts[1:]:
                        k, v = p.split("=")
                        a, b = v.split("/")
                        man[k] = (int(a), int(b))
                    dag.meta["manifest"] = man
                continue
            parts = line.split()
            if parts[0] == "out":
                key = (parts[1],) + tuple(int(p) for p in parts[2:-1])
                dag.outputs[key] = int(parts[-1])
                continue
            nid, op = int(parts[0]), parts[1]
            if nid != len(dag.nodes):
                raise ValueError(f"IR node ids not dense at {nid}")
            if op == "const":
                v = float.fromhex(parts[2])
                node = Node("const", (), struct.unpack("<Q", struct.pack("<d", v))[0])
            elif op == "bconst":
                node = Node("bconst", (), bool(int(parts[2])))
            elif op == "X":
                node = Node("X", (), int(parts[2]))
            else:
                args = tuple(int(p) for p in parts[2:])
                if op not in ARITY and op not in ("T", "z"):
                    raise ValueError(f"unknown op {op}")
                if op in ARITY and len(args) != ARITY[op]:
                    raise ValueError(f"bad arity for {op} at {nid}")
                if any(a >= nid for a in args):
                    raise ValueError(f"IR not topological at {nid}")
                node = Node(op, args, None)
            dag.index[node] = nid
            dag.nodes.append(node)
    return dag


def evaluate(dag, T, X, z=30.0, exp=math.exp, log=math.log):
    """Reference-order evaluation in Python floats (used only by checks)."""
    vals = [None] * len(dag.nodes)
    for nid, n in enumerate(dag.nodes):
        op, a = n.op, n.args
        if op == "const":
            v = dag.cval(nid)
        elif op == "bconst":
            v = n.val
        elif op == "T":
            v = T
        elif op == "z":
            v = z
        elif op == "X":
            v = X[n.val]
        elif op == "sel":
            v = vals[a[1]] if vals[a[0]] else vals[a[2]]
        else:
            x = [vals[i] for i in a]
            try:
                v = {"neg": lambda: -x[0], "abs": lambda: abs(x[0]),
                     "exp": lambda: exp(x[0]), "log": lambda: log(x[0]),
                     "sqrt": lambda: math.sqrt(x[0]),
                     "add": lambda: x[0] + x[1], "sub": lambda: x[0] - x[1],
                     "mul": lambda: x[0] * x[1],
                     "div": lambda: x[0] / x[1] if x[1] != 0 else math.copysign(math.inf, x[0]) * math.copysign(1.0, x[1]) if x[0] != 0 else math.nan,
                     "lt": lambda: x[0] < x[1], "le": lambda: x[0] <= x[1],
                     "gt": lambda: x[0] > x[1], "ge": lambda: x[0] >= x[1],
                     "eq": lambda: x[0] == x[1], "ne": lambda: x[0] != x[1],
                     "and": lambda: x[0] and x[1], "or": lambda: x[0] or x[1],
                     "not": lambda: not x[0]}[op]()
            except (ValueError, OverflowError):
                v = math.nan if op == "log" else math.inf
        vals[nid] = v
    return vals
