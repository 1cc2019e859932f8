"""Seeded generators of the benchmark networks, emitted as UPPAAL XML so they
go through the same front end as any user model.

  race_network(n, seed)        n single-clock 4-location cycles (race-8)
  fischer_network(n)           Fischer mutual exclusion with n processes
  sir_network(n, beta, gamma, initial_infected | seed)  agent-based SIR
"""
from __future__ import annotations

import math
import random
from xml.sax.saxutils import escape

DEFAULT_SEED = 20261202

_HEADER = ("<?xml version=\"1.0\" encoding=\"utf-8\"?>\n"
           "<!DOCTYPE nta PUBLIC '-//Uppaal Team//DTD Flat System 1.6//EN' "
           "'http://www.it.uu.se/research/group/darts/uppaal/flat-1_6.dtd'>\n")


def _label(kind: str, text: str) -> str:
    return f'<label kind="{kind}">{escape(text)}</label>'


def _nta(global_decl: str, templates: list[str], system: str) -> str:
    return (_HEADER + "<nta>\n<declaration>" + escape(global_decl) + "</declaration>\n"
            + "\n".join(templates) + f"\n<system>{escape(system)}</system>\n</nta>\n")


class _Ids:
    def __init__(self):
        self.n = 0

    def __call__(self) -> str:
        s = f"id{self.n}"
        self.n += 1
        return s


def race_network(n: int = 8, seed: int = DEFAULT_SEED) -> tuple[str, dict]:
    """n components, each the cycle L0 -> L1 -> L2 -> L3 -> L0 on its own clock
    x.  L0 and L2 carry invariant x <= U and edge guard x >= L with
    L ~ U{0..5}, U = L + U{1..10}; L1 and L3 have an exponential rate drawn
    from U[0.5, 2.0].  Every edge resets x and increments the shared `count`.
    Query: Pr[<=100](<> count >= 40)."""
    rng = random.Random(seed)
    ids = _Ids()
    tpls, params = [], []
    for k in range(n):
        L0, L1, L2, L3 = ids(), ids(), ids(), ids()
        lo0 = rng.randint(0, 5); up0 = lo0 + rng.randint(1, 10)
        lo2 = rng.randint(0, 5); up2 = lo2 + rng.randint(1, 10)
        r1 = round(rng.uniform(0.5, 2.0), 6); r3 = round(rng.uniform(0.5, 2.0), 6)
        params.append(dict(L0=(lo0, up0), L2=(lo2, up2), r1=r1, r3=r3))
        locs = [
            f'<location id="{L0}"><name>L0</name>{_label("invariant", f"x <= {up0}")}</location>',
            f'<location id="{L1}"><name>L1</name>{_label("exponentialrate", str(r1))}</location>',
            f'<location id="{L2}"><name>L2</name>{_label("invariant", f"x <= {up2}")}</location>',
            f'<location id="{L3}"><name>L3</name>{_label("exponentialrate", str(r3))}</location>',
        ]
        upd = _label("assignment", "x = 0, count = count + 1")
        trans = [
            f'<transition><source ref="{L0}"/><target ref="{L1}"/>{_label("guard", f"x >= {lo0}")}{upd}</transition>',
            f'<transition><source ref="{L1}"/><target ref="{L2}"/>{upd}</transition>',
            f'<transition><source ref="{L2}"/><target ref="{L3}"/>{_label("guard", f"x >= {lo2}")}{upd}</transition>',
            f'<transition><source ref="{L3}"/><target ref="{L0}"/>{upd}</transition>',
        ]
        tpls.append(f"<template><name>C{k}</name><declaration>clock x;</declaration>\n"
                    + "\n".join(locs) + f'\n<init ref="{L0}"/>\n' + "\n".join(trans) + "\n</template>")
    system = "system " + ", ".join(f"C{k}" for k in range(n)) + ";"
    return _nta("int count = 0;", tpls, system), {"components": params}


def fischer_network(n: int = 32, delay_bound: int = 2) -> tuple[str, dict]:
    """Fischer's protocol: shared `id`, per process clock x and pid = k+1.
    A -> req (id == 0, x := 0); req -> wait (invariant x <= 2, id := pid,
    x := 0); wait -> cs (x > 2 && id == pid); wait -> A (id != pid);
    cs -> A (id := 0).  A and cs leave at rate 1.  Query Pr[<=300](<> P0.cs)."""
    ids = _Ids()
    tpls = []
    for k in range(n):
        A, Rq, W, Cs = ids(), ids(), ids(), ids()
        locs = [
            f'<location id="{A}"><name>A</name>{_label("exponentialrate", "1")}</location>',
            f'<location id="{Rq}"><name>req</name>{_label("invariant", f"x <= {delay_bound}")}</location>',
            f'<location id="{W}"><name>wait</name></location>',
            f'<location id="{Cs}"><name>cs</name>{_label("exponentialrate", "1")}</location>',
        ]
        trans = [
            f'<transition><source ref="{A}"/><target ref="{Rq}"/>{_label("guard", "id == 0")}{_label("assignment", "x = 0")}</transition>',
            f'<transition><source ref="{Rq}"/><target ref="{W}"/>{_label("assignment", "id = pid, x = 0")}</transition>',
            f'<transition><source ref="{W}"/><target ref="{Cs}"/>{_label("guard", f"x > {delay_bound} && id == pid")}</transition>',
            f'<transition><source ref="{W}"/><target ref="{A}"/>{_label("guard", "id != pid")}</transition>',
            f'<transition><source ref="{Cs}"/><target ref="{A}"/>{_label("assignment", "id = 0")}</transition>',
        ]
        tpls.append(f"<template><name>P{k}</name><declaration>clock x;\nconst int pid = {k + 1};</declaration>\n"
                    + "\n".join(locs) + f'\n<init ref="{A}"/>\n' + "\n".join(trans) + "\n</template>")
    system = "system " + ", ".join(f"P{k}" for k in range(n)) + ";"
    return _nta("int id = 0;", tpls, system), {"n": n}


def sir_network(n: int = 350, beta: float = 0.3, gamma: float = 0.1, *, initial_infected: int | None = None,
                seed: int = DEFAULT_SEED, weights: tuple[int, int] = (1, 9)) -> tuple[str, dict]:
    """Agent-based SIR with shared S, I, R.  An S agent becomes infected at
    rate beta*I/N and announces it on broadcast channel `infect`; every other
    S agent then takes one of two weighted receive edges: with weight
    weights[0] it is infected too, with weights[1] it stays susceptible.  I
    agents recover at rate gamma.  By default 1% of the agents (at least
    one), chosen from the seed, start infected."""
    rng = random.Random(seed)
    if initial_infected is None:
        initial_infected = max(1, round(0.01 * n))
    infected = set(rng.sample(range(n), initial_infected))
    ids = _Ids()
    tpls = []
    w_inf, w_stay = weights
    for k in range(n):
        Sx, Ix, Rx = ids(), ids(), ids()
        locs = [
            f'<location id="{Sx}"><name>S</name>{_label("exponentialrate", "beta * I / N")}</location>',
            f'<location id="{Ix}"><name>I</name>{_label("exponentialrate", "gamma")}</location>',
            f'<location id="{Rx}"><name>R</name></location>',
        ]
        inf = _label("assignment", "S = S - 1, I = I + 1")
        trans = [
            f'<transition><source ref="{Sx}"/><target ref="{Ix}"/>{_label("synchronisation", "infect!")}{inf}</transition>',
            f'<transition><source ref="{Sx}"/><target ref="{Ix}"/>{_label("synchronisation", "infect?")}'
            f'{_label("probability", str(w_inf))}{inf}</transition>',
            f'<transition><source ref="{Sx}"/><target ref="{Sx}"/>{_label("synchronisation", "infect?")}'
            f'{_label("probability", str(w_stay))}</transition>',
            f'<transition><source ref="{Ix}"/><target ref="{Rx}"/>{_label("assignment", "I = I - 1, R = R + 1")}</transition>',
        ]
        init = Ix if k in infected else Sx
        tpls.append(f"<template><name>A{k}</name><declaration></declaration>\n"
                    + "\n".join(locs) + f'\n<init ref="{init}"/>\n' + "\n".join(trans) + "\n</template>")
    i0 = len(infected)
    decl = (f"broadcast chan infect;\nint S = {n - i0};\nint I = {i0};\nint R = 0;\n"
            f"const double beta = {beta!r};\nconst double gamma = {gamma!r};\nconst int N = {n};")
    system = "system " + ", ".join(f"A{k}" for k in range(n)) + ";"
    return _nta(decl, tpls, system), {"n": n, "beta": beta, "gamma": gamma, "initial_infected": sorted(infected)}


def log_uniform_int(rng: random.Random, lo: int, hi: int) -> int:
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def batch_instances(n_sir: int = 256, n_fischer: int = 256, *, seed: int = DEFAULT_SEED) -> list[dict]:
    """Uneven instances for one batched launch: SIR networks with N
    log-uniform in [16, 512], beta uniform in [0.1, 0.5], gamma uniform in
    [0.05, 0.2] (query Pr[<=100](<> I > N/5)), and Fischer networks with N
    uniform in {4..64} (query Pr[<=300](<> P0.cs)).  Each entry holds the
    XML, the location query and the query's parameters."""
    rng = random.Random(seed)
    out = []
    for k in range(n_sir):
        n = log_uniform_int(rng, 16, 512)
        beta = rng.uniform(0.1, 0.5)
        gamma = rng.uniform(0.05, 0.2)
        xml, meta = sir_network(n, beta, gamma, seed=seed + 1 + k)
        out.append({"kind": "sir", "xml": xml, "loc_query": "", "var": "I", "threshold": n // 5,
                    "time_bound": 100.0, **meta})
    for k in range(n_fischer):
        n = rng.randint(4, 64)
        xml, _ = fischer_network(n)
        out.append({"kind": "fischer", "xml": xml, "loc_query": "P0.cs", "time_bound": 300.0, "n": n})
    return out
