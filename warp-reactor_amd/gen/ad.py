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
        (reproducer.cpp:7029-7050) with each reaction's rate multiplied by its
        multiplier M_r = exp(theta_r) in every species rate it enters and in
        the energy rate where the network books its heat;
  J     J_theta: the analytic Jacobian of the reference at y
        (reproducer.cpp:7052-7060, 1538-6994) plus, per reaction, (M_r - 1)
        times the Jacobian of that reaction's own part of f, so that J_theta
        is the Jacobian of f_theta in the reference Jacobian's variables;
  post  (y, rho1) -> (xn, T): floor/normalise, charge balance, floor/normalise,
        eos_re (reproducer.cpp:7773-7776, 311-339).

Every builder below restates the C++ operation order exactly (left-associated
sums and products, the same constants), so the forward code emitted from these
DAGs is bit-identical to the hand-written forward path (with every M_r = 1,
f_theta and J_theta are f and J bit for bit); the derivative code is ordinary
forward-mode (jvp) and reverse-mode (vjp) differentiation of the same DAG
nodes.  Reactions: the reference's rate code is common-subexpression output
with no reaction structure; Reactions recovers the 45 reactions (each balanced
in H, D, He nuclei and charge) from the monomials of the species rates, and the
heat terms of seven of them (H2 formation through H-, H2+ and three bodies;
collisional dissociation by e-, H and H2) from the energy rate.
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


SPECIES = ["e-", "H+", "H", "H-", "D+", "D", "H2+", "D-", "H2", "HD+", "HD", "He++",
           "He+", "He"]
CHARGE = [-1, 1, 0, -1, 1, 0, 1, -1, 0, 1, 0, 2, 1, 0]


def _abundance_dependent(net):
    dep = [False] * len(net.nodes)
    for nid, n in enumerate(net.nodes):
        dep[nid] = n.op == "X" or any(dep[a] for a in n.args)
    return dep


def _product(a, b):
    return [(x * y, tuple(sorted(m1 + m2))) for x, m1 in a for y, m2 in b]


class Expansion:
    """Sum-of-products form of network nodes.  A node becomes a list of
    monomials (coefficient, atoms); atoms are a sorted tuple of ("a", node)
    for a node that is not a sum, difference, negation, product or quotient
    (an abundance X, a rate fit, a select ...) and ("i", node) for the
    reciprocal of a non-constant divisor.  Constant factors and divisors go
    into the coefficient; equal monomials are merged."""

    def __init__(self, net):
        self.net = net
        self.xdep = _abundance_dependent(net)
        self.memo = {}

    def __call__(self, n):
        got = self.memo.get(n)
        if got is not None:
            return got
        net = self.net
        nd, c = net.nodes[n], net.cval(n)
        if c is not None:
            r = [(c, ())]
        elif nd.op in ("add", "sub"):
            b = self(nd.args[1])
            r = self(nd.args[0]) + ([(-x, m) for x, m in b] if nd.op == "sub" else b)
        elif nd.op == "neg":
            r = [(-x, m) for x, m in self(nd.args[0])]
        elif nd.op == "mul":
            r = _product(self(nd.args[0]), self(nd.args[1]))
        elif nd.op == "div" and net.cval(nd.args[1]) is not None:
            q = net.cval(nd.args[1])
            r = [(x / q, m) for x, m in self(nd.args[0])]
        elif nd.op == "div":
            r = _product(self(nd.args[0]), [(1.0, (("i", nd.args[1]),))])
        else:
            r = [(1.0, (("a", n),))]
        acc = collections.OrderedDict()
        for x, m in r:
            acc[m] = acc.get(m, 0.0) + x
        r = [(x, m) for m, x in acc.items() if x != 0.0]
        self.memo[n] = r
        return r

    def leaves(self, m):
        """Species indices of the abundance factors of monomial m."""
        return tuple(sorted(self.net.nodes[a].val for k, a in m
                            if k == "a" and self.net.nodes[a].op == "X"))


def _sample_states(net, n=16, seed=20240611):
    """Node values of the network at n states spread over the collapse's
    range (T 20 K .. 3e4 K, hydrogen densities 1e-1 .. 1e12, abundances
    down to 1e-8 of it)."""
    import random
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        T = 10.0 ** rng.uniform(1.3, 4.5)
        nh = 10.0 ** rng.uniform(-1.0, 12.0)
        X = [nh * 10.0 ** rng.uniform(-8.0, 0.0) for _ in range(NSPEC)]
        out.append(dm.evaluate(net, T, X))
    return out


def _proportional(u, v, rtol=1e-9, need=4):
    """u and v (value lists over the sample states) differ by a constant factor."""
    import math
    q = [a / b for a, b in zip(u, v)
         if math.isfinite(a) and math.isfinite(b) and a != 0.0 and b != 0.0 and
         math.isfinite(a / b)]
    if len(q) < need:
        return False
    q.sort()
    mid = q[len(q) // 2]
    return all(abs(x - mid) <= rtol * abs(mid) for x in q)


class Reactions:
    """The network's reactions, recovered from the generated rate code.

    The reference's rhs_specie (reproducer.cpp:342-672) is common-subexpression
    code: one reaction's rate can be split over several temporaries and
    temporaries can mix reactions.  Expanding the 14 species rates into
    monomials (Expansion) undoes that: a monomial is c x (rate-fit factors) x
    (abundances), and each monomial enters the species equations with a vector
    of coefficients.  Monomials whose abundance-dependent factors are equal and
    whose coefficient vectors are parallel belong to one reaction (a rate fit
    that is a sum of terms gives several monomials of one reaction); the vector
    is the reaction's stoichiometry, and every one of them conserves mass and
    charge.

    rhs_eint (reproducer.cpp:679-1512) recomputes some of these rates in its
    own expressions (the heat of H2 formation and dissociation).  A monomial
    of the energy rate is that reaction's thermal channel when its abundance
    factors are the reaction's and its rate factors are proportional, as
    functions of (T, X), to one of the reaction's rate terms, and the
    energy-rate monomials so matched to a reaction sum to a multiple of its
    whole rate (checked on sample states).  The rest of the energy rate (line
    cooling, collisional ionisation and recombination cooling fits,
    bremsstrahlung, Compton) is not a reaction of the network."""

    def __init__(self, net):
        self.net = net
        ex = self.ex = Expansion(net)
        table = collections.OrderedDict()
        for i in range(NSPEC):
            for c, m in ex(net.outputs[("ydot", i)]):
                table.setdefault(m, [0.0] * NSPEC)[i] += c
        self.label = {}          # species-rate monomial -> reaction
        self.list = []           # per reaction: dict
        index = {}
        for m, v in table.items():
            xpart = tuple(a for a in m if ex.xdep[a[1]])
            lead = next(i for i in range(NSPEC) if v[i] != 0.0)
            key = (xpart, tuple(round(v[i] / v[lead], 9) for i in range(NSPEC)))
            r = index.get(key)
            if r is None:
                r = index[key] = len(self.list)
                self.list.append({"xpart": xpart, "dir": [v[i] / v[lead] for i in range(NSPEC)],
                                  "terms": [], "leaves": ex.leaves(m)})
            self.list[r]["terms"].append((v[lead], m))
            self.label[m] = r
        self.R = len(self.list)
        for rx in self.list:
            rx["nu"] = self._integer_nu(rx)
        self.thermal = self._thermal_channels()

    @staticmethod
    def _integer_nu(rx):
        """Smallest integer stoichiometry along the coefficient vector, signed so
        that every species the reaction consumes is one of its abundance
        factors, and of two such signs the one that consumes more of them (for
        a rate with no plain abundance factors, so that the first nonzero entry
        is a loss)."""
        d = rx["dir"]
        for k in (1, 2, 3, 4, 6):
            if all(abs(k * x - round(k * x)) < 1e-9 for x in d):
                nu = [int(round(k * x)) for x in d]
                break
        else:
            raise ValueError("non-integer stoichiometry")
        leaves = rx["leaves"]

        def covered(v):
            return all(leaves.count(s) >= -v[s] for s in range(NSPEC) if v[s] < 0)
        neg = [-x for x in nu]
        ok = [v for v in (nu, neg) if covered(v)]
        if len(ok) == 2:   # the orientation that consumes more of the abundance factors
            used = [sum(-x for x in v if x < 0) for v in ok]
            if used[0] == used[1]:
                raise ValueError("ambiguous stoichiometry sign")
            ok = [ok[used.index(max(used))]]
        if ok:
            return ok[0]
        if leaves:
            raise ValueError("stoichiometry not covered by the abundance factors")
        return nu if next(x for x in nu if x) < 0 else neg

    def _value(self, vals, m, coef=1.0, skip_leaves=False, skip_inv=False):
        v = coef
        for k, a in m:
            if k == "i":
                if not (skip_inv and self.ex.xdep[a]):
                    v /= vals[a]
            elif not (skip_leaves and self.net.nodes[a].op == "X"):
                v *= vals[a]
        return v

    def _thermal_channels(self):
        net, ex = self.net, self.ex
        states = _sample_states(net)
        groups = collections.OrderedDict()
        for c, m in ex(net.outputs[("edot",)]):
            inv = tuple(a for a in m if a[0] == "i" and ex.xdep[a[1]])
            groups.setdefault((ex.leaves(m), inv), []).append((c, m))
        thermal = {}
        for (leaves, _), mons in groups.items():
            for r, rx in enumerate(self.list):
                if rx["leaves"] != leaves:
                    continue
                terms = [[self._value(s, m, c, skip_leaves=True) for s in states]
                         for c, m in rx["terms"]]
                rate = [sum(t[k] for t in terms) for k in range(len(states))]
                hit = []
                for c, m in mons:
                    if m in thermal:
                        continue
                    u = [self._value(s, m, c, skip_leaves=True, skip_inv=True) for s in states]
                    if any(_proportional(u, t) for t in terms):
                        hit.append((m, u))
                if not hit:
                    continue
                tot = [sum(u[k] for _, u in hit) for k in range(len(states))]
                if _proportional(tot, rate):
                    for m, _ in hit:
                        thermal[m] = r
        return thermal

    def _monomial_node(self, c, m):
        net = self.net
        v = None
        for k, a in m:
            f = a if k == "a" else net.add("div", (net.const(1.0), a))
            v = f if v is None else net.add("mul", (v, f))
        return net.add("mul", (net.const(c), v)) if v is not None else net.const(c)

    def rate_nodes(self):
        """Per reaction, network nodes of its flux (the rate of the reaction,
        so that its part of species rate i is dir_i * flux) with the per-species
        factors dir, and of its part of the energy rate (None without one)."""
        out = []
        chan = collections.defaultdict(list)
        for c, m in self.ex(self.net.outputs[("edot",)]):
            r = self.thermal.get(m)
            if r is not None:
                chan[r].append((c, m))
        for r, rx in enumerate(self.list):
            flux = None
            for c, m in rx["terms"]:
                t = self._monomial_node(c, m)
                flux = t if flux is None else self.net.add("add", (flux, t))
            heat = None
            for c, m in chan.get(r, []):
                t = self._monomial_node(c, m)
                heat = t if heat is None else self.net.add("add", (heat, t))
            out.append((flux, rx["dir"], heat))
        return out

    def species_label(self, m):
        return self.label[m]

    def energy_label(self, m):
        return self.thermal.get(m)

    def name(self, r):
        nu, leaves = self.list[r]["nu"], self.list[r]["leaves"]
        lhs = list(leaves)
        if any(nu[s] < 0 and lhs.count(s) < -nu[s] for s in range(NSPEC)):
            lhs = [s for s in range(NSPEC) for _ in range(max(0, -nu[s]))]
        cnt = [lhs.count(s) + nu[s] for s in range(NSPEC)]
        rhs = [s for s in range(NSPEC) for _ in range(cnt[s])]
        return " + ".join(SPECIES[s] for s in lhs) + " -> " + " + ".join(SPECIES[s] for s in rhs)

    def table(self):
        chan = collections.Counter(self.thermal.values())
        return [{"r": r, "reaction": self.name(r), "nu": rx["nu"],
                 "rate_terms": len(rx["terms"]), "thermal_monomials": chan.get(r, 0)}
                for r, rx in enumerate(self.list)]


class MixedLabels(Exception):
    pass


class Tagger:
    """Multiplies each labelled monomial of a network expression by its
    reaction's multiplier without changing the expression's operation order:
    a subtree all of whose monomials (in the product context it sits in) carry
    one label r becomes M_r * subtree; sums and products of mixed subtrees are
    rebuilt around tagged parts.  With every M_r = 1 the result is
    bit-identical to the untagged expression (x * 1 = x)."""

    def __init__(self, net, ex, d, remap, M):
        self.net, self.ex, self.d, self.remap, self.M = net, ex, d, remap, M
        self.inserted = 0

    def labels(self, n, outer, label):
        return {label(tuple(sorted(m1 + m2))) for _, m1 in outer for _, m2 in self.ex(n)}

    def build(self, n, label, outer=((1.0, ()),)):
        net, d, remap = self.net, self.d, self.remap
        L = self.labels(n, outer, label)
        if L == {None}:
            return remap[n]
        if len(L) == 1:
            self.inserted += 1
            return d.add("mul", (remap[n], self.M[L.pop()]))
        nd = net.nodes[n]
        a = nd.args
        if nd.op in ("add", "sub"):
            return d.add(nd.op, (self.build(a[0], label, outer), self.build(a[1], label, outer)))
        if nd.op == "neg":
            return d.add("neg", (self.build(a[0], label, outer),))
        if nd.op == "mul":
            for k in (0, 1):
                try:
                    new = self.build(a[k], label, _product(outer, self.ex(a[1 - k])))
                except MixedLabels:
                    continue
                args = (new, remap[a[1]]) if k == 0 else (remap[a[0]], new)
                return d.add("mul", args)
            raise MixedLabels(n)
        if nd.op == "div" and net.cval(a[1]) is not None:
            return d.add("div", (self.build(a[0], label, outer), remap[a[1]]))
        if nd.op == "div":
            inner = _product(outer, [(1.0, (("i", a[1]),))])
            return d.add("div", (self.build(a[0], label, inner), remap[a[1]]))
        raise MixedLabels(n)


class Model:
    """The DAGs of one cell step, in one hash-consed Dag with named inputs:
    y[15], M[R] (reaction rate multipliers), xn[14], T, rho, driver, dt."""

    def __init__(self):
        self.net = network()
        self.rx = Reactions(self.net)
        self.R = self.rx.R
        self.d = dm.Dag()
        o = self.o = Ops(self.d)
        self.y = [o.inp("y", i) for i in range(NEQ)]
        self.M = [o.inp("M", r) for r in range(self.R)]
        X, T = state_from_y(o, self.y)
        f_roots = [self.net.outputs[("ydot", i)] for i in range(NSPEC)] + \
                  [self.net.outputs[("edot",)]]
        j_roots = [self.net.outputs[("jac", r, c)] for r in range(NEQ) for c in range(NEQ)]
        remap = self._graft(X, T, f_roots + j_roots)
        tag = Tagger(self.net, self.rx.ex, self.d, remap, self.M)
        self.f = [tag.build(f_roots[i], self.rx.species_label) for i in range(NSPEC)]
        self.f.append(tag.build(f_roots[NSPEC], self.rx.energy_label))
        self.inserted = tag.inserted
        self.J_ref = [remap[n] for n in j_roots]
        self.J = self._reaction_jacobian(X, T, self.J_ref)
        # pre / post pieces
        self.xn = [o.inp("xn", i) for i in range(NSPEC)]
        self.Tin = o.inp("s", 0)
        self.rho = o.inp("s", 1)
        self.drv = o.inp("s", 2)
        self.dt = o.inp("s", 3)
        self.pre_y, self.pre_rho, self.pre_drv = pre_step(o, self.xn, self.Tin, self.rho,
                                                          self.drv, self.dt)
        self.post_xn, self.post_T = post_step(o, self.y[:NSPEC], self.rho, self.y[NSPEC])

    def _reaction_jacobian(self, X, T, J):
        """J_theta = J - D with D = sum_r (1 - M_r) dG_r/dz: the reference's
        Jacobian J (reproducer.cpp:1539-6994; species columns at fixed T, the
        energy column dF/dT dT/de) plus, per reaction, its own part G_r of f
        (dir * flux in the species rows, its heat in the energy row) scaled by
        M_r - 1 and differentiated the same way (forward mode over the
        network).  Per reaction and column the factor (1 - M_r) dflux/dz_j is
        formed once and shared by the rows.  D is summed from +0, so that with
        every M_r = 1 it is +0 and J_theta is J bit for bit."""
        net, d, o = self.net, self.d, self.o
        one = net.const(1.0)
        zero = d.const(0.0)
        xleaf = {net.nodes[n].val: n for n in range(len(net.nodes)) if net.nodes[n].op == "X"}
        tleaf = next(n for n in range(len(net.nodes)) if net.nodes[n].op == "T")
        dTde = jvp(d, [T], {self.y[NSPEC]: d.const(1.0)})[0]

        def columns(node):
            live = net.live_nodes([node])
            cols = [(j, xleaf[j]) for j in range(NSPEC) if xleaf.get(j) in live]
            return cols + ([(NSPEC, tleaf)] if tleaf in live else [])

        def scaled_derivative(r, node, leaf):
            g = jvp(net, [node], {leaf: one})[0]
            if net.cval(g) == 0.0:
                return None
            return o.mul(o.sub(d.const(1.0), self.M[r]), self._graft(X, T, [g])[g])

        terms = collections.defaultdict(list)     # (row, column) -> [(factor, node)]
        for r, (flux, dirv, heat) in enumerate(self.rx.rate_nodes()):
            for j, leaf in columns(flux):
                a = scaled_derivative(r, flux, leaf)
                if a is not None:
                    for i in range(NSPEC):
                        if dirv[i] != 0.0:
                            terms[(i, j)].append((dirv[i], a))
            if heat is not None:
                for j, leaf in columns(heat):
                    a = scaled_derivative(r, heat, leaf)
                    if a is not None:
                        terms[(NSPEC, j)].append((1.0, a))
        out = list(J)
        roots = []
        self.D = {}      # J_theta[k] = J[k] - D[k]
        for (i, j), lst in sorted(terms.items()):
            acc = zero
            for c, a in lst:
                if c == 1.0:
                    acc = o.add(acc, a)
                elif c == -1.0:
                    acc = o.sub(acc, a)
                else:
                    acc = o.add(acc, o.mul(d.const(c), a))
            if j == NSPEC:
                acc = o.add(zero, o.mul(acc, dTde))
            roots.append(acc)
            self.D[i * NEQ + j] = acc
            out[i * NEQ + j] = o.sub(J[i * NEQ + j], acc)
        self.jac_terms = sum(len(v) for v in terms.values())
        # Divisions in D: their adjoints can overflow where a rate fit has
        # underflowed to a subnormal (the derivative of exp(c log|k|) by way of
        # 1/|k|); those contributions are taken as 0.
        self.finite_div = frozenset(n for n in d.live_nodes(roots) if d.nodes[n].op == "div")
        return out

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

    def reaction_table(self):
        return self.rx.table()


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


def vjp(d, roots, seeds, wrt, finite_div=frozenset()):
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

    inf = o.k(float("inf"))

    def finite(v):  # v, or zero where v overflowed
        return o.sel(o.lt(d.add("abs", (v,)), inf), v, zero)
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
        elif op == "div" and n in finite_div:
            push(a[0], finite(gdiv(g, a[1])))
            push(a[1], o.neg(gmul(g, finite(o.div(n, a[1])))))
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
    J_t = jvp(d, m.J, dict(zip(m.y + m.M, dy + dM)))
    pre_roots = m.pre_y + [m.pre_rho, m.pre_drv]
    pre_t = jvp(d, pre_roots, dict(zip(m.xn + [m.Tin, m.rho, m.drv], dxn + ds[:3])))
    post_roots = m.post_xn + [m.post_T]
    post_t = jvp(d, post_roots, dict(zip(m.y + [m.rho], dy + [ds[1]])))
    # adjoints
    f_a = vjp(d, m.f, w, m.y + m.M)
    # J_theta's adjoint in two sweeps that run one after the other (sharing
    # their staging memory): the reference Jacobian's, then D's with seeds -W.
    J_a = vjp(d, m.J_ref, W, m.y)
    dk = sorted(m.D)
    D_a = vjp(d, [m.D[k] for k in dk], [o.neg(W[k]) for k in dk], m.y + m.M,
              finite_div=m.finite_div)
    pre_a = vjp(d, pre_roots, wy + ws[:2], m.xn + [m.Tin, m.rho, m.drv])
    post_a = vjp(d, post_roots, wx + [ws[0]], m.y + [m.rho])

    P_Y = "const double* y"
    fns, info = [], {"R": R, "reactions": m.reaction_table(), "multipliers_inserted": m.inserted}
    plain = [
        ("ad_f", f"{P_Y}, const double* M, double* f", [(f"f[{i}]", m.f[i]) for i in range(NEQ)]),
        ("ad_fj", f"{P_Y}, const double* M, double* f, double* J",
         [(f"f[{i}]", m.f[i]) for i in range(NEQ)] + [(f"J[{k}]", m.J[k]) for k in range(225)]),
        ("ad_f_jvp", f"{P_Y}, const double* M, const double* dy, const double* dM, double* df",
         [(f"df[{i}]", f_t[i]) for i in range(NEQ)]),
        ("ad_jac_jvp", f"{P_Y}, const double* M, const double* dy, const double* dM, double* dJ",
         [(f"dJ[{k}]", J_t[k]) for k in range(225)]),
        ("ad_f_vjp", f"{P_Y}, const double* M, const double* w, double* ybar, double* Mbar",
         [(f"ybar[{i}]", f_a[i]) for i in range(NEQ)] +
         [(f"Mbar[{r}]", f_a[NEQ + r]) for r in range(R)]),
        ("ad_jac_vjp", f"{P_Y}, const double* W, double* ybar",
         [(f"ybar[{i}]", J_a[i]) for i in range(NEQ)]),
        ("ad_jacd_vjp", f"{P_Y}, const double* M, const double* W, double* ybar, double* Mbar",
         [(f"ybar[{i}]", D_a[i]) for i in range(NEQ)] +
         [(f"Mbar[{r}]", D_a[NEQ + r]) for r in range(R)]),
        # D is linear in M: its M-adjoint (the same nodes as ad_jacd_vjp's
        # Mbar) does not depend on M, and its y-adjoint carries the factor
        # 1 - M_r, so with every M_r = 1 the y-adjoint is zero and this is all
        # of D's adjoint
        ("ad_jacd_mbar", f"{P_Y}, const double* W, double* Mbar",
         [(f"Mbar[{r}]", D_a[NEQ + r]) for r in range(R)]),
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
             [(f"Y, {i}", J_a[i]) for i in range(NEQ)], 128, 500, 32),
            ("ad_jacd_vjp_st", "AD_STAGE_IN(y), AD_STAGE_IN(M), AD_STAGE_IN(W) AD_STAGE_PARAMS",
             [(f"Y, {i}", D_a[i]) for i in range(NEQ)] +
             [(f"MB, {r}", D_a[NEQ + r]) for r in range(R)], 128, 500, 32),
            ("ad_jacd_mbar_st", "AD_STAGE_IN(y), AD_STAGE_IN(W) AD_STAGE_PARAMS",
             [(f"MB, {r}", D_a[NEQ + r]) for r in range(R)], 128, 500, 32)]:
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
