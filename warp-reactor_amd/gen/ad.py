"""Tangent-linear and adjoint code of the frozen-step collapse map, generated
by source transformation of the network DAG.

The differentiated map is one cell's outer step of the gridwide-independent
collapse advance (reproducer.cpp:7736-7785) with every controller decision
held at the value the forward run took: dt_grid of each outer step and the
step sizes h of the accepted ROS2S attempts of each cell (rejected attempts
leave y unchanged, so only accepted ones enter the map).  Its pieces:

  pre   (xn, T, rho, driver; dt) -> (y0, rho1, driver1)
        the advance kernel's density update (collapse_timestep, driver +=
        dt driver / tff, ratio scaling of xn and rho) and eos_rt
        (reproducer.cpp:7633-7644, 7750-7766, 7598-7604);
  f     f_theta(y) = the network RHS at (T(e, X), X = max(y, 1e-100))
        (reproducer.cpp:7029-7050) with each reaction flux term of the 14
        species equations multiplied by its rate multiplier M_r = exp(theta_r);
  J     the analytic Jacobian of the reference at y (reproducer.cpp:7052-7060,
        1538-6994), which does not depend on theta;
  post  (y, rho1) -> (xn, T): floor/normalise, charge balance, floor/normalise,
        eos_re (reproducer.cpp:7773-7776, 311-339).

Every builder below restates the C++ operation order exactly (left-associated
sums and products, the same constants), so the forward code emitted from these
DAGs is bit-identical to the hand-written forward path; the derivative code is
ordinary forward-mode (jvp) and reverse-mode (vjp) differentiation of the same
DAG nodes.  Reaction flux terms: expanding each species rate ydot_i through
+, -, unary - and products with constants leaves its terms; each distinct term
node is one flux (rate coefficient x abundances), 50 in all, 44 of them shared
by two or more species equations.  The energy equation's rates are separate
expressions in the reference and are not scaled.
"""
from __future__ import annotations

import collections
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dag as dm  # noqa: E402
import emit as em  # noqa: E402
import stage  # noqa: E402

NSPEC, NEQ = 14, 15
# Network constants (reproducer.cpp:233-252): species masses and gamma.
MASS = [9.10938188e-28, 1.67262158e-24, 1.67353251819e-24, 1.67444345638e-24,
        3.34512158e-24, 3.34603251819e-24, 3.34615409819e-24, 3.34694345638e-24,
        3.34706503638e-24, 5.01865409819e-24, 5.01956503638e-24, 6.69024316e-24,
        6.69115409819e-24, 6.69206503638e-24]
GAMMA = [1.4 if n in (6, 8, 9, 10) else 5.0 / 3.0 for n in range(NSPEC)]
GINV = [1.0 / (g - 1.0) for g in GAMMA]
MP = 1.67262192595e-24
RGAS = 6.02214076e+23 * 1.3806490000000002e-16
FLOOR = 1.0e-100
GRAV = 6.674e-8
PI = 3.141592653589793238462643383279502884
LEAF = ("const", "bconst", "T", "X", "in", "z")


class Ops:
    """Node builders on a Dag (hash-consed, folded as dag.Dag.add folds)."""

    def __init__(self, d):
        self.d = d

    def k(self, v):
        return self.d.const(v)

    def inp(self, name, i):
        return self.d.add("in", (), f"{name}:{i}")

    def add(self, a, b):
        return self.d.add("add", (a, b))

    def sub(self, a, b):
        return self.d.add("sub", (a, b))

    def mul(self, a, b):
        return self.d.add("mul", (a, b))

    def div(self, a, b):
        return self.d.add("div", (a, b))

    def neg(self, a):
        return self.d.add("neg", (a,))

    def sqrt(self, a):
        return self.d.add("sqrt", (a,))

    def lt(self, a, b):
        return self.d.add("lt", (a, b))

    def sel(self, c, a, b):
        return self.d.add("sel", (c, a, b))

    def rmax(self, a, b):  # std::max as the reference evaluates it: (a < b) ? b : a
        return self.sel(self.lt(a, b), b, a)


# ---- EOS and driver pieces (wr_eos.hpp restates reproducer.cpp:270-339, 7633-7644) -------

def density_of(o, xn):
    rho = o.k(0.0)
    for n in range(NSPEC):
        rho = o.add(rho, o.mul(xn[n], o.k(MASS[n])))
    return rho


def eos_sums(o, xn):
    rhotot = density_of(o, xn)
    sa, sg = o.k(0.0), o.k(0.0)
    for n in range(NSPEC):
        sa = o.add(sa, xn[n])
        sg = o.add(sg, o.mul(o.div(o.mul(xn[n], o.k(MP)), rhotot), o.k(GINV[n])))
    sa = o.mul(sa, o.div(o.k(MP), rhotot))
    sg = o.div(sg, sa)
    return sa, sg


def eos_temperature(o, xn, e):
    sa, sg = eos_sums(o, xn)
    return o.div(e, o.mul(o.mul(sg, o.k(RGAS)), sa))


def eos_energy(o, xn, T):
    sa, sg = eos_sums(o, xn)
    return o.mul(o.mul(o.mul(sg, sa), o.k(RGAS)), T)


def normalize(o, xn, rho):
    mf = [o.div(o.mul(o.k(MASS[n]), xn[n]), rho) for n in range(NSPEC)]
    s = o.k(0.0)
    for n in range(NSPEC):
        s = o.add(s, mf[n])
    return [o.div(o.mul(o.div(mf[n], s), rho), o.k(MASS[n])) for n in range(NSPEC)]


def floor_normalize(o, xn, rho):
    return normalize(o, [o.rmax(x, o.k(FLOOR)) for x in xn], rho)


def balance_charge(o, x):
    v = o.sub(o.neg(x[3]), x[7])
    for i in (1, 12, 6, 4, 9):
        v = o.add(v, x[i])
    v = o.add(v, o.mul(o.k(2.0), x[11]))
    return [v] + list(x[1:])


def post_step(o, xn, rho, e):
    xn = floor_normalize(o, xn, rho)
    xn = balance_charge(o, xn)
    xn = floor_normalize(o, xn, rho)
    return xn, eos_temperature(o, xn, e)


def pre_step(o, xn, T, rho, driver, dt):
    """advance kernel preamble + eos_rt for a cell that integrates."""
    rhox = density_of(o, xn)
    tffx = o.sqrt(o.div(o.k(PI * 3.0), o.mul(o.mul(o.k(32.0), rhox), o.k(GRAV))))
    local_dt = o.mul(o.k(1.0e-1), tffx)
    tff = o.div(local_dt, o.k(1.0e-1))
    drv = o.add(driver, o.mul(dt, o.div(driver, tff)))
    ratio = o.div(drv, driver)
    xn1 = [o.mul(x, ratio) for x in xn]
    rho1 = o.mul(rho, ratio)
    e0 = eos_energy(o, xn1, T)
    return xn1 + [e0], rho1, drv


def state_from_y(o, y):
    X = [o.rmax(y[n], o.k(FLOOR)) for n in range(NSPEC)]
    return X, eos_temperature(o, X, y[NSPEC])


# ---- the network with rate multipliers ---------------------------------------------------

def network():
    return em.specialise(dm.read_ir(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                 "primordial_network.ir")))


def flux_terms(net):
    """Terms of the 14 species rates (expansion through + - neg and products
    with constants), in order of first appearance; with, per term, the species
    equations it enters and its coefficient there."""
    f_roots = [net.outputs[("ydot", i)] for i in range(NSPEC)]
    where = collections.OrderedDict()

    def walk(n, coef, i):
        nd = net.nodes[n]
        if nd.op in ("add", "sub"):
            walk(nd.args[0], coef, i)
            walk(nd.args[1], coef if nd.op == "add" else -coef, i)
        elif nd.op == "neg":
            walk(nd.args[0], -coef, i)
        elif nd.op == "mul" and net.cval(nd.args[0]) is not None:
            walk(nd.args[1], coef * net.cval(nd.args[0]), i)
        elif nd.op == "mul" and net.cval(nd.args[1]) is not None:
            walk(nd.args[0], coef * net.cval(nd.args[1]), i)
        else:
            where.setdefault(n, []).append((i, coef))

    for i, r in enumerate(f_roots):
        walk(r, 1.0, i)
    return where


class Model:
    """The DAGs of one cell step, in one hash-consed Dag with named inputs:
    y[15], M[R] (rate multipliers), xn[14], T, rho, driver, dt."""

    def __init__(self):
        self.net = network()
        self.terms = flux_terms(self.net)
        self.R = len(self.terms)
        self.d = dm.Dag()
        o = self.o = Ops(self.d)
        self.y = [o.inp("y", i) for i in range(NEQ)]
        self.M = [o.inp("M", r) for r in range(self.R)]
        X, T = state_from_y(o, self.y)
        f_roots = [self.net.outputs[("ydot", i)] for i in range(NSPEC)] + \
                  [self.net.outputs[("edot",)]]
        j_roots = [self.net.outputs[("jac", r, c)] for r in range(NEQ) for c in range(NEQ)]
        remap = self._graft(X, T, f_roots + j_roots)
        term_id = {n: r for r, n in enumerate(self.terms)}
        memo = {}

        def scaled(n):
            if n in memo:
                return memo[n]
            nd = self.net.nodes[n]
            if n in term_id:
                v = o.mul(remap[n], self.M[term_id[n]])
            elif nd.op in ("add", "sub"):
                v = self.d.add(nd.op, (scaled(nd.args[0]), scaled(nd.args[1])))
            elif nd.op == "neg":
                v = o.neg(scaled(nd.args[0]))
            elif nd.op == "mul" and self.net.cval(nd.args[0]) is not None:
                v = o.mul(remap[nd.args[0]], scaled(nd.args[1]))
            elif nd.op == "mul" and self.net.cval(nd.args[1]) is not None:
                v = o.mul(scaled(nd.args[0]), remap[nd.args[1]])
            else:
                raise AssertionError("term expansion mismatch")
            memo[n] = v
            return v

        self.f = [scaled(f_roots[i]) for i in range(NSPEC)] + [remap[f_roots[NSPEC]]]
        self.J = [remap[n] for n in j_roots]
        # pre / post pieces
        self.xn = [o.inp("xn", i) for i in range(NSPEC)]
        self.Tin = o.inp("s", 0)
        self.rho = o.inp("s", 1)
        self.drv = o.inp("s", 2)
        self.dt = o.inp("s", 3)
        self.pre_y, self.pre_rho, self.pre_drv = pre_step(o, self.xn, self.Tin, self.rho,
                                                          self.drv, self.dt)
        self.post_xn, self.post_T = post_step(o, self.y[:NSPEC], self.rho, self.y[NSPEC])

    def _graft(self, X, T, roots):
        net, d = self.net, self.d
        remap = {}
        for nid in sorted(net.live_nodes(roots)):
            n = net.nodes[nid]
            if n.op == "T":
                remap[nid] = T
            elif n.op == "X":
                remap[nid] = X[n.val]
            elif n.op == "const":
                remap[nid] = d.const(net.cval(nid))
            elif n.op == "bconst":
                remap[nid] = d.add("bconst", (), n.val)
            else:
                remap[nid] = d.add(n.op, tuple(remap[a] for a in n.args))
        return remap

    def term_table(self):
        names = ["e-", "H+", "H", "H-", "D+", "D", "H2+", "D-", "H2", "HD+", "HD", "He++",
                 "He+", "He"]
        out = []
        for r, (n, w) in enumerate(self.terms.items()):
            out.append({"r": r, "species": [[names[i], c] for i, c in w]})
        return out


# ---- differentiation -----------------------------------------------------------------------

def _sum(o, a, b):
    if a is None:
        return b
    if b is None:
        return a
    return o.add(a, b)


def jvp(d, roots, seed):
    """Forward mode: tangent node of each root given leaf -> tangent node."""
    o = Ops(d)
    zero = o.k(0.0)
    t = {}
    for n in sorted(d.live_nodes(roots)):
        nd = d.nodes[n]
        if nd.op in LEAF:
            t[n] = seed.get(n)
            continue
        if d.is_bool(n):
            continue
        a = nd.args
        ta = [t.get(x) for x in a]
        op = nd.op
        if op == "add":
            r = _sum(o, ta[0], ta[1])
        elif op == "sub":
            r = ta[0] if ta[1] is None else (o.neg(ta[1]) if ta[0] is None else o.sub(ta[0], ta[1]))
        elif op == "neg":
            r = None if ta[0] is None else o.neg(ta[0])
        elif op == "mul":
            r = _sum(o, None if ta[0] is None else o.mul(ta[0], a[1]),
                     None if ta[1] is None else o.mul(a[0], ta[1]))
        elif op == "div":
            # d(a/b) = (da - (a/b) db) / b
            num = ta[0]
            if ta[1] is not None:
                q = o.mul(n, ta[1])
                num = o.neg(q) if num is None else o.sub(num, q)
            r = None if num is None else o.div(num, a[1])
        elif op == "abs":
            r = None if ta[0] is None else o.sel(o.lt(a[0], zero), o.neg(ta[0]), ta[0])
        elif op == "exp":
            r = None if ta[0] is None else o.mul(ta[0], n)
        elif op == "log":
            r = None if ta[0] is None else o.div(ta[0], a[0])
        elif op == "sqrt":
            r = None if ta[0] is None else o.div(o.mul(ta[0], o.k(0.5)), n)
        elif op == "sel":
            if ta[1] is None and ta[2] is None:
                r = None
            else:
                r = o.sel(a[0], ta[1] if ta[1] is not None else zero,
                          ta[2] if ta[2] is not None else zero)
        else:
            raise ValueError(op)
        t[n] = r
    return [t.get(r) if t.get(r) is not None else zero for r in roots]


def vjp(d, roots, seeds, wrt):
    """Reverse mode: adjoint node of each leaf in `wrt` given one seed node per
    root.  Contributions to a node are summed in the order they arrive.

    Selects route a zero adjoint into their untaken operand, whose forward
    values may be infinite or NaN there (exp(0 log 0) and the like, which the
    forward code discards); a product or quotient of that zero with a
    non-constant partial is therefore zero-preserving: (g == 0) ? 0 : g p."""
    o = Ops(d)
    zero = o.k(0.0)
    contrib = collections.defaultdict(list)

    def guard(g, v):
        return o.sel(d.add("eq", (g, zero)), zero, v)

    def gmul(g, p):  # g * p, zero when g is zero
        return o.mul(g, p) if d.cval(p) is not None else guard(g, o.mul(g, p))

    def gdiv(g, p):  # g / p, zero when g is zero
        return o.div(g, p) if d.cval(p) is not None else guard(g, o.div(g, p))
    for r, s in zip(roots, seeds):
        contrib[r].append(s)

    def push(n, g):
        nd = d.nodes[n]
        if nd.op in ("const", "bconst") or d.is_bool(n):
            return
        contrib[n].append(g)

    for n in sorted(d.live_nodes(roots), reverse=True):
        nd = d.nodes[n]
        if nd.op in LEAF or d.is_bool(n) or n not in contrib:
            continue
        g = None
        for c in contrib[n]:
            g = _sum(o, g, c)
        a = nd.args
        op = nd.op
        if op == "add":
            push(a[0], g)
            push(a[1], g)
        elif op == "sub":
            push(a[0], g)
            push(a[1], o.neg(g))
        elif op == "neg":
            push(a[0], o.neg(g))
        elif op == "mul":
            push(a[0], gmul(g, a[1]))
            push(a[1], gmul(g, a[0]))
        elif op == "div":
            push(a[0], gdiv(g, a[1]))
            if d.cval(a[1]) is None:
                push(a[1], o.neg(gmul(g, o.div(n, a[1]))))
        elif op == "abs":
            push(a[0], o.sel(o.lt(a[0], zero), o.neg(g), g))
        elif op == "exp":
            push(a[0], gmul(g, n))
        elif op == "log":
            push(a[0], gdiv(g, a[0]))
        elif op == "sqrt":
            push(a[0], gdiv(o.mul(g, o.k(0.5)), n))
        elif op == "sel":
            push(a[1], o.sel(a[0], g, zero))
            push(a[2], o.sel(a[0], zero, g))
        else:
            raise ValueError(op)
    out = []
    for w in wrt:
        g = None
        for c in contrib.get(w, []):
            g = _sum(o, g, c)
        out.append(g if g is not None else zero)
    return out


# ---- emission ----------------------------------------------------------------------------

class _PlainEmitter(em._Emitter):
    def ref(self, nid):
        n = self.d.nodes[nid]
        if n.op == "in":
            name, idx = n.val.split(":")
            return f"{name}[{idx}]"
        return super().ref(nid)


def emit_plain(d, fname, params, outputs):
    """Straight-line function; outputs: list of (C lvalue, node)."""
    e = _PlainEmitter(d)
    roots = [n for _, n in outputs]
    lines = [f"CHEM_FN void {fname}({params})", "{"]
    ops = 0
    for nid in sorted(d.live_nodes(roots)):
        n = d.nodes[nid]
        if n.op in LEAF:
            continue
        ops += 1
        is_b = d.is_bool(nid)
        nm = ("b" if is_b else "v") + str(nid)
        e.name[nid] = nm
        lines.append(f"    const {'int' if is_b else 'double'} {nm} = {e.expr(nid)};")
    for lv, n in outputs:
        lines.append(f"    {lv} = {e.ref(n)};")
    lines.append("}")
    return "\n".join(lines), ops


def emit_staged_fn(d, fname, params, outputs, R, prefetch=500, extra=32):
    """Register-budgeted version (stage.py); outputs leave through
    AD_OUT(lvalue, v) as soon as computed."""
    outs = collections.OrderedDict()
    for lv, n in outputs:
        outs.setdefault(n, []).append(lv)
    text, plan = stage.emit_staged(d, fname, outs, R, out_macro="AD_OUT", prefetch=prefetch,
                                   prefetch_extra=extra, params=params, in_macro="AD_IN")
    return text, {"slots": plan.nslots, "stores": plan.stores, "loads": plan.loads}


def build(out_path, info_path=None, dev_path=None):
    m = Model()
    d, o, R = m.d, m.o, m.R
    dy = [o.inp("dy", i) for i in range(NEQ)]
    dM = [o.inp("dM", r) for r in range(R)]
    w = [o.inp("w", i) for i in range(NEQ)]
    W = [o.inp("W", k) for k in range(NEQ * NEQ)]
    dxn = [o.inp("dxn", i) for i in range(NSPEC)]
    ds = [o.inp("ds", i) for i in range(4)]
    wy = [o.inp("wy", i) for i in range(NEQ)]
    wx = [o.inp("wx", i) for i in range(NSPEC)]
    ws = [o.inp("ws", i) for i in range(3)]

    # tangents
    f_t = jvp(d, m.f, dict(zip(m.y + m.M, dy + dM)))
    J_t = jvp(d, m.J, dict(zip(m.y, dy)))
    pre_roots = m.pre_y + [m.pre_rho, m.pre_drv]
    pre_t = jvp(d, pre_roots, dict(zip(m.xn + [m.Tin, m.rho, m.drv], dxn + ds[:3])))
    post_roots = m.post_xn + [m.post_T]
    post_t = jvp(d, post_roots, dict(zip(m.y + [m.rho], dy + [ds[1]])))
    # adjoints
    f_a = vjp(d, m.f, w, m.y + m.M)
    J_a = vjp(d, m.J, W, m.y)
    pre_a = vjp(d, pre_roots, wy + ws[:2], m.xn + [m.Tin, m.rho, m.drv])
    post_a = vjp(d, post_roots, wx + [ws[0]], m.y + [m.rho])

    P_Y = "const double* y"
    fns, info = [], {"R": R, "terms": m.term_table()}
    plain = [
        ("ad_f", f"{P_Y}, const double* M, double* f", [(f"f[{i}]", m.f[i]) for i in range(NEQ)]),
        ("ad_fj", f"{P_Y}, const double* M, double* f, double* J",
         [(f"f[{i}]", m.f[i]) for i in range(NEQ)] + [(f"J[{k}]", m.J[k]) for k in range(225)]),
        ("ad_f_jvp", f"{P_Y}, const double* M, const double* dy, const double* dM, double* df",
         [(f"df[{i}]", f_t[i]) for i in range(NEQ)]),
        ("ad_jac_jvp", f"{P_Y}, const double* dy, double* dJ",
         [(f"dJ[{k}]", J_t[k]) for k in range(225)]),
        ("ad_f_vjp", f"{P_Y}, const double* M, const double* w, double* ybar, double* Mbar",
         [(f"ybar[{i}]", f_a[i]) for i in range(NEQ)] +
         [(f"Mbar[{r}]", f_a[NEQ + r]) for r in range(R)]),
        ("ad_jac_vjp", f"{P_Y}, const double* W, double* ybar",
         [(f"ybar[{i}]", J_a[i]) for i in range(NEQ)]),
        ("ad_pre", "const double* xn, const double* s, double* y0, double* s1",
         [(f"y0[{i}]", m.pre_y[i]) for i in range(NEQ)] +
         [("s1[0]", m.pre_rho), ("s1[1]", m.pre_drv)]),
        ("ad_pre_jvp", "const double* xn, const double* s, const double* dxn, const double* ds, "
                       "double* dy0, double* ds1",
         [(f"dy0[{i}]", pre_t[i]) for i in range(NEQ)] +
         [("ds1[0]", pre_t[NEQ]), ("ds1[1]", pre_t[NEQ + 1])]),
        ("ad_pre_vjp", "const double* xn, const double* s, const double* wy, const double* ws, "
                       "double* xnbar, double* sbar",
         [(f"xnbar[{i}]", pre_a[i]) for i in range(NSPEC)] +
         [(f"sbar[{k}]", pre_a[NSPEC + k]) for k in range(3)]),
        ("ad_post", f"{P_Y}, const double* s, double* xn, double* T",
         [(f"xn[{i}]", m.post_xn[i]) for i in range(NSPEC)] + [("T[0]", m.post_T)]),
        ("ad_post_jvp", f"{P_Y}, const double* s, const double* dy, const double* ds, "
                        "double* dxn, double* dT",
         [(f"dxn[{i}]", post_t[i]) for i in range(NSPEC)] + [("dT[0]", post_t[NSPEC])]),
        ("ad_post_vjp", f"{P_Y}, const double* s, const double* wx, const double* ws, "
                        "double* ybar, double* rhobar",
         [(f"ybar[{i}]", post_a[i]) for i in range(NEQ)] + [("rhobar[0]", post_a[NEQ])]),
    ]
    counts = {}
    for name, params, outs in plain:
        text, ops = emit_plain(d, name, params, outs)
        fns.append(text)
        counts[name] = ops
    info["ops"] = counts
    staged = []
    # register budgets (doubles) and reload distances (operations, extra registers)
    for name, params, outs, regs, pf, extra in [
            ("ad_fj_st", "AD_STAGE_IN(y), AD_STAGE_IN(M) AD_STAGE_PARAMS",
             [(f"F, {i}", m.f[i]) for i in range(NEQ)] +
             [(f"J, {k}", m.J[k]) for k in range(225)], 128, 500, 32),
            ("ad_f_st", "AD_STAGE_IN(y), AD_STAGE_IN(M) AD_STAGE_PARAMS",
             [(f"F, {i}", m.f[i]) for i in range(NEQ)], 96, 500, 32),
            ("ad_f_vjp_st", "AD_STAGE_IN(y), AD_STAGE_IN(M), AD_STAGE_IN(w) AD_STAGE_PARAMS",
             [(f"Y, {i}", f_a[i]) for i in range(NEQ)] +
             [(f"MB, {r}", f_a[NEQ + r]) for r in range(R)], 112, 500, 32),
            ("ad_jac_vjp_st", "AD_STAGE_IN(y), AD_STAGE_IN(W) AD_STAGE_PARAMS",
             [(f"Y, {i}", J_a[i]) for i in range(NEQ)], 128, 500, 32)]:
        text, st = emit_staged_fn(d, name, params, outs, regs, prefetch=pf, extra=extra)
        staged.append(text)
        info.setdefault("staged", {})[name] = st
    nslot = max(v["slots"] for v in info["staged"].values())
    fwd_slots = max(info["staged"]["ad_fj_st"]["slots"], info["staged"]["ad_f_st"]["slots"])
    hdr = [
        "/* Generated by warp-reactor_amd/gen/ad.py from primordial_network.ir. Do not edit.",
        " * Forward, tangent-linear and adjoint code of one cell's frozen-step collapse",
        " * map: f_theta, J, the advance preamble and the post-step sequence.",
        " * Compile with -ffp-contract=off. The includer defines CHEM_FN, CHEM_EXP,",
        " * CHEM_LOG; for the staged forms AD_STAGE_IN, AD_IN, AD_STAGE_PARAMS,",
        " * STAGE_STORE/LOAD and AD_OUT. */",
        f"#define AD_NTHETA {R}",
        f"#define AD_NSLOT {max(nslot, 1)}",
        f"#define AD_FWD_NSLOT {max(fwd_slots, 1)}",
        "",
    ]
    # the CPU header: every plain function; the device header: the staged
    # network pieces and the small plain ones (pre/post and their adjoints)
    dev_plain = [t for (name, _, _), t in zip(plain, fns)
                 if name in ("ad_pre", "ad_post", "ad_pre_vjp", "ad_post_vjp")]
    for path, body in ((out_path, hdr + fns),
                       (dev_path, hdr + dev_plain + [""] + staged)):
        if not path:
            continue
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            f.write("\n".join(body + [""]))
        os.replace(tmp, path)
    if info_path:
        with open(info_path, "w") as f:
            json.dump(info, f, indent=1)
    return info


if __name__ == "__main__":
    # ad.py CPU_HEADER [INFO_JSON [DEVICE_HEADER]]
    inf = build(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else None,
                sys.argv[3] if len(sys.argv) > 3 else None)
    print("theta terms:", inf["R"], "ops:", inf["ops"], "staged:", inf.get("staged"))
