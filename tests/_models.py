"""Tiny UPPAAL models with known answers, built as XML strings."""
from __future__ import annotations

import math

import numpy as np

HEAD = '<?xml version="1.0" encoding="utf-8"?>\n<nta>\n'


def single_exponential(rate: float) -> str:
    """One component: l0 (rate) -> Goal.  Pr[<=T] = 1 - 2^(-rate T)."""
    return HEAD + f"""<declaration>int dummy = 0;</declaration>
<template><name>T</name><declaration></declaration>
<location id="id0"><name>l0</name><label kind="exponentialrate">{rate}</label></location>
<location id="id1"><name>Goal</name></location>
<init ref="id0"/>
<transition><source ref="id0"/><target ref="id1"/></transition>
</template><system>system T;</system></nta>"""


def single_uniform(upper: int) -> str:
    """One component: l0 with invariant x <= U -> Goal.  Pr[<=T] = T/U."""
    return HEAD + f"""<declaration>int dummy = 0;</declaration>
<template><name>T</name><declaration>clock x;</declaration>
<location id="id0"><name>l0</name><label kind="invariant">x &lt;= {upper}</label></location>
<location id="id1"><name>Goal</name></location>
<init ref="id0"/>
<transition><source ref="id0"/><target ref="id1"/></transition>
</template><system>system T;</system></nta>"""


def first_winner(rates: list[float], urgent: list[bool] | None = None) -> str:
    """Components racing once: the first to fire sets w := index + 1 (guard
    w == 0 keeps the others from firing).  The distribution of w is the
    distribution of the race winner."""
    urgent = urgent or [False] * len(rates)
    tpls = []
    for k, r in enumerate(rates):
        a, b = 2 * k, 2 * k + 1
        lab = "<urgent/>" if urgent[k] else f'<label kind="exponentialrate">{r}</label>'
        tpls.append(f"""<template><name>C{k}</name><declaration></declaration>
<location id="id{a}"><name>go</name>{lab}</location>
<location id="id{b}"><name>done</name></location>
<init ref="id{a}"/>
<transition><source ref="id{a}"/><target ref="id{b}"/><label kind="guard">w == 0</label>
<label kind="assignment">w = {k + 1}</label></transition>
</template>""")
    return HEAD + "<declaration>int w = 0;</declaration>\n" + "\n".join(tpls) + "<system>system;</system></nta>"


def urgent_order() -> str:
    """comp0 exponential, comp1 and comp2 urgent; each fires once, appending
    its number to w.  Zero delays win, lowest index first: comp1 then comp2
    fire at time 0, so w starts 2, 23."""
    tpls = []
    for k, urg in enumerate([False, True, True]):
        a, b = 2 * k, 2 * k + 1
        lab = "<urgent/>" if urg else '<label kind="exponentialrate">1</label>'
        tpls.append(f"""<template><name>C{k}</name><declaration></declaration>
<location id="id{a}"><name>go</name>{lab}</location>
<location id="id{b}"><name>done</name></location>
<init ref="id{a}"/>
<transition><source ref="id{a}"/><target ref="id{b}"/>
<label kind="assignment">w = w * 10 + {k + 1}</label></transition>
</template>""")
    return HEAD + "<declaration>int w = 0;</declaration>\n" + "\n".join(tpls) + "<system>system;</system></nta>"


def zeno() -> str:
    """An urgent location whose only edge is never enabled: zero delays
    forever, time never reaches the bound; runs are cut off at max_steps."""
    return HEAD + """<declaration>int w = 0;</declaration>
<template><name>Z</name><declaration></declaration>
<location id="id0"><name>stuck</name><urgent/></location>
<location id="id1"><name>never</name></location>
<init ref="id0"/>
<transition><source ref="id0"/><target ref="id1"/><label kind="guard">w == 1</label></transition>
</template><system>system Z;</system></nta>"""


def sir_exact(n: int, beta: float, gamma: float, i0: int, k: int, T: float, p_inf: float = 0.1) -> float:
    """Pr[max I >= k within T] for the agent SIR network of
    cusmc_amd.generators.sir_network under the step semantics, from the
    lumped CTMC on (S, I): rates are scaled by ln 2 (delays are -log2(u)/rate),
    a spontaneous infection is followed by a Binomial(S-1, p_inf) wave of
    receivers, recovered agents race at rate 1 without changing the state, and
    the transition that crosses the time bound still counts (it is fired
    before time is compared; its receivers are not processed)."""
    from scipy.linalg import expm
    from scipy.stats import binom
    ln2 = math.log(2.0)
    states = [(s, i) for i in range(0, k) for s in range(0, n - i + 1)]
    idx = {st: j for j, st in enumerate(states)}
    HIT = len(states)
    Q = np.zeros((HIT + 1, HIT + 1))
    spont, total = np.zeros(HIT), np.zeros(HIT)
    for (s, i), j in idx.items():
        a = s * beta * i / n * ln2
        g = i * gamma * ln2
        spont[j] = a
        total[j] = a + g + (n - s - i) * ln2
        if a > 0:
            for m in range(0, s):
                pm = binom.pmf(m, s - 1, p_inf)
                s2, i2 = s - 1 - m, i + 1 + m
                dst = HIT if i + 1 >= k or i2 >= k else idx[(s2, i2)]
                Q[j, dst] += a * pm
        if g > 0:
            Q[j, idx[(s, i - 1)]] += g
        Q[j, j] -= a + g
    p0 = np.zeros(HIT + 1)
    p0[idx[(n - i0, i0)]] = 1.0
    pT = p0 @ expm(Q * T)
    over = sum(pT[idx[(s, k - 1)]] * spont[idx[(s, k - 1)]] / total[idx[(s, k - 1)]]
               for s in range(0, n - (k - 1) + 1) if total[idx[(s, k - 1)]] > 0)
    return float(pT[HIT] + over)


def receiver_goal(live_rate: float | None = None, clock_rate: bool = False) -> str:
    """A sender S: a (rate 1) --go!--> b, a receiver R: r0 --go?--> Goal, and
    optionally a live component L looping at `live_rate`.  The goal is
    entered by a broadcast receiver, so it counts only if the step after the
    receivers (every component racing, S in b and R in Goal at the default
    rate 1) keeps t < T:
      without L:  Pr = (1 - 2^-T)^2
      with L:     Pr = 1 - 2^-T - (2^-T - 2^-(2+r)T) / (1+r),  r = live_rate.
    clock_rate: L's rate is `live_rate + x` with x a clock never reset (a
    rate that changes at every step)."""
    live = ""
    system = "system S, R;"
    if live_rate is not None:
        rate = f"{live_rate} + x" if clock_rate else f"{live_rate}"
        live = f"""<template><name>L</name><declaration>clock x;</declaration>
<location id="id4"><name>l</name><label kind="exponentialrate">{rate}</label></location>
<init ref="id4"/>
<transition><source ref="id4"/><target ref="id4"/></transition>
</template>"""
        system = "system S, R, L;"
    return HEAD + f"""<declaration>broadcast chan go;
int dummy = 0;</declaration>
<template><name>S</name><declaration></declaration>
<location id="id0"><name>a</name><label kind="exponentialrate">1</label></location>
<location id="id1"><name>b</name></location>
<init ref="id0"/>
<transition><source ref="id0"/><target ref="id1"/><label kind="synchronisation">go!</label></transition>
</template>
<template><name>R</name><declaration></declaration>
<location id="id2"><name>r0</name></location>
<location id="id3"><name>Goal</name></location>
<init ref="id2"/>
<transition><source ref="id2"/><target ref="id3"/><label kind="synchronisation">go?</label></transition>
</template>
{live}<system>{system}</system></nta>"""


def receiver_goal_p(T: float, live_rate: float | None = None) -> float:
    """Closed form of receiver_goal (see there)."""
    a = 2.0 ** -T
    if live_rate is None:
        return (1 - a) ** 2
    r = live_rate
    return 1 - a - (a - 2.0 ** (-(2 + r) * T)) / (1 + r)


def clock_rate_race() -> str:
    """A: a (rate 0.1 + x, x a clock) -> Goal beside D, a component with no
    edge (default rate 1).  D's null steps re-evaluate A's rate; the law of
    A's firing time depends on them, so the two race samplings agree only
    if D races."""
    return HEAD + """<declaration>int dummy = 0;</declaration>
<template><name>A</name><declaration>clock x;</declaration>
<location id="id0"><name>a</name><label kind="exponentialrate">0.1 + x</label></location>
<location id="id1"><name>Goal</name></location>
<init ref="id0"/>
<transition><source ref="id0"/><target ref="id1"/></transition>
</template>
<template><name>D</name><declaration></declaration>
<location id="id2"><name>d</name></location>
<init ref="id2"/>
</template><system>system A, D;</system></nta>"""


def name_locations(xml: str, names: dict[str, str]) -> str:
    """Give unnamed locations a name: `names` maps a location id to its name
    (test fixtures of the reference leave most locations unnamed)."""
    import re
    for lid, nm in names.items():
        pat = re.compile(r'(<location id="' + re.escape(lid) + r'"[^>]*>)')
        xml, k = pat.subn(lambda m: m.group(1) + f"<name>{nm}</name>", xml, count=1)
        assert k == 1, lid
    return xml


class RawNetwork:
    """A network given directly as smc_network arrays, without the XML front
    end (whose grammar cannot express every operator of the evaluator, e.g.
    `||` or `==` outside a `?:` condition).  Exposes .net_ptr / .net like
    cusmc_amd.ParsedNetwork, so the oracle and the GPU simulator take it."""

    def __init__(self):
        self.var_kind, self.var_init = [], []
        self.exprs = []          # (op, a, b, c, var, value)
        self.comps = []          # (first location, init location, n locations)
        self.locs = []           # dict(type, rate, invs=[(op, var, rhs)])
        self.edges = []          # dict(src, dst, channel, weight, guards=[(op, var, rhs)], upds=[(var, expr)])

    def var(self, init: float = 0.0, clock: bool = False) -> int:
        self.var_kind.append(1 if clock else 0)
        self.var_init.append(float(init))
        return len(self.var_kind) - 1

    def expr(self, op: int, a: int = -1, b: int = -1, c: int = -1, var: int = -1, value: float = 0.0) -> int:
        self.exprs.append((op, a, b, c, var, float(value)))
        return len(self.exprs) - 1

    def lit(self, v: float) -> int:
        return self.expr(0, value=v)

    def ref(self, v: int) -> int:
        return self.expr(1, var=v)

    def component(self, locs: list[dict], init: int = 0) -> int:
        """Locations of one component (dicts with type, rate, invs); returns
        the global index of its first location."""
        first = len(self.locs)
        for L in locs:
            self.locs.append({"type": L.get("type", 0), "rate": L.get("rate", -1), "invs": L.get("invs", [])})
        self.comps.append((first, first + init, len(locs)))
        return first

    def edge(self, src: int, dst: int, *, channel: int = 0, weight: int = -1, guards=(), upds=()) -> None:
        self.edges.append({"src": src, "dst": dst, "channel": channel, "weight": weight,
                           "guards": list(guards), "upds": list(upds)})

    def build(self) -> "RawNetwork":
        import ctypes as C
        from cusmc_amd import _capi as A
        nl = len(self.locs)
        by_src = [[e for e in self.edges if e["src"] == l] for l in range(nl)]
        cons, upds, edges, locs = [], [], [], []
        for l, L in enumerate(self.locs):
            inv_first = len(cons)
            for op, v, rhs in L["invs"]:
                cons.append(A.Constraint(op, 1, v, -1, rhs))
            edge_first = len(edges)
            for e in by_src[l]:
                g_first = len(cons)
                for op, v, rhs in e["guards"]:
                    cons.append(A.Constraint(op, 1, v, -1, rhs))
                u_first = len(upds)
                for v, x in e["upds"]:
                    upds.append(A.Update(v, x))
                edges.append(A.Edge(l, e["dst"], e["channel"], e["weight"], g_first, len(e["guards"]), u_first,
                                    len(e["upds"])))
            locs.append(A.Location(L["type"], L["rate"], inv_first, len(L["invs"]), edge_first, len(by_src[l])))

        def arr(ctype, items):
            a = (ctype * max(1, len(items)))()
            for i, x in enumerate(items):
                a[i] = x
            return a

        self._keep = [
            arr(C.c_int32, self.var_kind), arr(C.c_double, self.var_init),
            arr(C.c_int32, [c[0] for c in self.comps] + [nl]), arr(C.c_int32, [c[1] for c in self.comps]),
            arr(A.Location, locs), arr(A.Edge, edges), arr(A.Constraint, cons), arr(A.Update, upds),
            arr(A.ExprNode, [A.ExprNode(op, a, b, c, v, 0, val) for op, a, b, c, v, val in self.exprs]),
        ]
        k = self._keep
        self._net = A.Network(len(self.comps), len(self.var_kind), nl, len(edges), len(cons), len(upds),
                              len(self.exprs), 0, k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7], k[8])
        self.net_ptr = C.pointer(self._net)
        return self

    @property
    def net(self):
        return self._net


# Expression operators (smc_network.h, cuSMC expr::operators) and their
# meaning in expressions.cu:10-95, restated in Python for brute-force checks.
E_LIT, E_VAR, E_RANDOM, E_ADD, E_SUB, E_MUL, E_DIV, E_POW, E_NEG, E_SQRT, E_MOD = range(11)
E_AND, E_OR, E_LE, E_GE, E_LT, E_GT, E_EQ, E_NE, E_NOT, E_COND = range(11, 21)
UNARY = (E_NEG, E_SQRT, E_NOT)
BINARY = (E_ADD, E_SUB, E_MUL, E_DIV, E_POW, E_MOD, E_AND, E_OR, E_LE, E_GE, E_LT, E_GT, E_EQ, E_NE)
TOL = 1e-5


def _c_int(x: float) -> int:
    """static_cast<int> with saturation and NaN -> 0 (C's cast is undefined
    outside int's range; the simulator saturates)."""
    if math.isnan(x):
        return 0
    if x >= 2147483647.0:
        return 2147483647
    if x <= -2147483648.0:
        return -2147483648
    return int(x)  # truncation toward zero


def py_eval(exprs, e: int, vars_, pow_fn=None) -> float:
    """expressions.cu:10-95 on a tree of (op, a, b, c, var, value) tuples.
    pow_fn replaces math.pow (the simulator's own pow is checked against
    math.pow separately, to a tolerance)."""
    op, a, b, c, v, val = exprs[e]
    if op == E_LIT:
        return val
    if op == E_VAR:
        return vars_[v]
    if op == E_RANDOM:
        return 1.0
    if op == E_COND:
        cond = py_eval(exprs, a, vars_, pow_fn)
        return py_eval(exprs, b, vars_, pow_fn) if abs(cond) > TOL else py_eval(exprs, c, vars_, pow_fn)
    x = py_eval(exprs, a, vars_, pow_fn)
    if op == E_NEG:
        return -x
    if op == E_SQRT:
        return math.sqrt(x) if x >= 0 else (x if math.isnan(x) else float("nan"))
    if op == E_NOT:
        return 1.0 if abs(x) < TOL else 0.0
    y = py_eval(exprs, b, vars_, pow_fn)
    if op == E_ADD:
        return x + y
    if op == E_SUB:
        return x - y
    if op == E_MUL:
        return x * y
    if op == E_DIV:
        if y == 0.0:
            if x == 0.0 or math.isnan(x):
                return float("nan")
            return math.copysign(math.inf, x) * math.copysign(1.0, y)
        return x / y
    if op == E_POW:
        if pow_fn is not None:
            return pow_fn(x, y)
        try:
            return math.pow(x, y)
        except (OverflowError, ValueError):
            return float("nan") if x < 0 else math.inf
    if op == E_MOD:
        ia, ib = _c_int(x), _c_int(y + TOL)
        if ib == 0 or (ia == -2147483648 and ib == -1):
            return 0.0
        r = abs(ia) % abs(ib)
        return float(-r if ia < 0 else r)
    if op == E_AND:
        return 1.0 if abs(x) > TOL and abs(y) > TOL else 0.0
    if op == E_OR:
        return 1.0 if abs(x) > TOL or abs(y) > TOL else 0.0
    if op == E_LE:
        return 1.0 if x <= y else 0.0
    if op == E_GE:
        return 1.0 if x >= y else 0.0
    if op == E_LT:
        return 1.0 if x < y else 0.0
    if op == E_GT:
        return 1.0 if x > y else 0.0
    if op == E_EQ:
        return 1.0 if abs(x - y) <= TOL else 0.0
    if op == E_NE:
        return 1.0 if abs(x - y) > TOL else 0.0
    raise ValueError(op)


def random_tree(rng, net: RawNetwork, n_vars: int, depth: int) -> int:
    """A random expression over every operator, with leaves drawn so that
    comparisons sit near their tolerance (equal values, values 1e-6 apart)."""
    if depth == 0 or rng.random() < 0.25:
        k = rng.integers(0, 5)
        if k < 2:
            return net.ref(int(rng.integers(0, n_vars)))
        if k == 2:
            return net.lit(float(rng.integers(-3, 8)))
        if k == 3:
            return net.lit(float(rng.choice([0.0, 1e-6, -1e-6, 2e-5, 0.5, 2.0])))
        return net.lit(round(float(rng.uniform(-4, 4)), 4))
    r = rng.random()
    if r < 0.15:
        return net.expr(int(rng.choice(UNARY + (E_RANDOM,))), random_tree(rng, net, n_vars, depth - 1))
    if r < 0.25:
        return net.expr(E_COND, random_tree(rng, net, n_vars, depth - 1), random_tree(rng, net, n_vars, depth - 1),
                        random_tree(rng, net, n_vars, depth - 1))
    return net.expr(int(rng.choice(BINARY)), random_tree(rng, net, n_vars, depth - 1),
                    random_tree(rng, net, n_vars, depth - 1))


def large_model(name: str) -> str:
    """A model of the reference's large-model corpus (tests/golden/models/
    large, gzipped) as XML text."""
    import gzip
    from pathlib import Path
    p = Path(__file__).resolve().parent / "golden" / "models" / "large" / (name + ".gz")
    return gzip.decompress(p.read_bytes()).decode()


def guarded_uniform(n_dead: int = 0) -> str:
    """U: l0 with invariant x <= 5 and one edge to Goal guarded by x >= 2,
    beside n_dead components with no edge (default rate 1).  cuSMC redraws
    the uniform delay on (0, 5 - x] at every step (after U's own null wins
    with x < 2 and after every null step of the others); its hazard
    1/(5 - tau) depends on the absolute time only, so for 2 <= T <= 5
    Pr[<=T](<> Goal) = 1 - (5 - T)/3 = (T - 2)/3, whatever n_dead."""
    dead = "".join(f'<template><name>D{k}</name><declaration></declaration>'
                   f'<location id="idd{k}"><name>d</name></location><init ref="idd{k}"/></template>'
                   for k in range(n_dead))
    system = "system U" + "".join(f", D{k}" for k in range(n_dead)) + ";"
    return HEAD + f"""<declaration>int dummy = 0;</declaration>
<template><name>U</name><declaration>clock x;</declaration>
<location id="id0"><name>l0</name><label kind="invariant">x &lt;= 5</label></location>
<location id="id1"><name>Goal</name></location>
<init ref="id0"/>
<transition><source ref="id0"/><target ref="id1"/><label kind="guard">x &gt;= 2</label></transition>
</template>{dead}<system>{system}</system></nta>"""
