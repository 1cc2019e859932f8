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
