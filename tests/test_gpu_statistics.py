"""GPU estimates vs exact answers and vs the reference algorithm.

(a) closed forms and the exact SIR CTMC, at 10^6 runs (tight intervals);
(b) every estimate lies inside the Chernoff-Hoeffding interval of the
    reference's own algorithm (the oracle in LITERAL mode: each component
    draws its own delay every step, as simulation.cu:409-541 does);
(c) edge cases: a Zeno cycle is cut off identically on GPU and oracle, and
    random update expressions are evaluated bit for bit like the oracle.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cusmc_amd as cs
from cusmc_amd import generators as G

import oracle

from _models import (E_ADD, E_MUL, E_NOT, RawNetwork, random_tree, receiver_goal, receiver_goal_p, single_exponential,
                     single_uniform, sir_exact, zeno)

pytestmark = pytest.mark.gpu


def _within(est, p, n, k=5.0):
    return abs(est - p) <= k * math.sqrt(max(p * (1 - p), 1e-12) / n) + 1e-12


def _gpu(net, q, seed, n, **kw):
    sim = cs.GpuSimulator(net)
    try:
        return sim.run(q, seed=seed, n_runs=n, **kw)
    finally:
        sim.close()


@pytest.mark.parametrize("rate,T", [(0.7, 2.0), (2.5, 0.4)])
def test_gpu_exponential_closed_form(rate, T):
    net = cs.ParsedNetwork.from_string(single_exponential(rate), "")
    n = 1_000_000
    r = _gpu(net, cs.reach_query(T), 21, n)
    assert _within(r.estimate, 1 - 2 ** (-rate * T), n), r.estimate


def test_gpu_uniform_closed_form():
    net = cs.ParsedNetwork.from_string(single_uniform(8), "")
    n = 1_000_000
    r = _gpu(net, cs.reach_query(3.0), 22, n)
    assert _within(r.estimate, 3 / 8, n), r.estimate


def test_gpu_condtest_and_racetest(models_dir):
    n = 1_000_000
    net = cs.ParsedNetwork.from_file(models_dir / "condTest.xml", "Template.l3")
    r = _gpu(net, cs.reach_query(4.0), 23, n)
    assert _within(r.estimate, 0.75 - 2 * math.log(2) * 2 ** -4.0, n), r.estimate
    net = cs.ParsedNetwork.from_file(models_dir / "RaceTest.xml", "")
    T = 2.0
    r = _gpu(net, cs.var_query(0, T, is_max=True, estimate=True, threshold=-1), 24, n)
    mean = 8 * (1 - 2 ** -T) + (4 / 9) * 2 ** -T
    sd = math.sqrt(max(r.sumsq / n - (r.sum / n) ** 2, 0))
    assert abs(r.estimate - mean) <= 5 * sd / math.sqrt(n), (r.estimate, mean)


@pytest.mark.parametrize("live", [None, 0.2], ids=["absorbing", "live"])
def test_gpu_receiver_goal_closed_form(live):
    """A goal entered by a broadcast receiver counts only if the next step of
    every component keeps t < T (simulation.cu:409-415, 835-849)."""
    net = cs.ParsedNetwork.from_string(receiver_goal(live), "")
    n = 1_000_000
    T = 2.0
    r = _gpu(net, cs.reach_query(T), 26, n)
    p = receiver_goal_p(T, live)
    assert _within(r.estimate, p, n), (r.estimate, p)


@pytest.mark.parametrize("n_agents,beta,gamma,k,T", [(12, 0.9, 0.3, 5, 4.0), (20, 0.6, 0.2, 8, 6.0)])
def test_gpu_sir_exact_ctmc(n_agents, beta, gamma, k, T):
    xml, _ = G.sir_network(n_agents, beta, gamma, initial_infected=1, seed=1)
    net = cs.ParsedNetwork.from_string(xml, "")
    q = cs.var_query(net.variable_id("I"), T, is_max=True, threshold=k - 1)
    n = 1_000_000
    r = _gpu(net, q, 25, n)
    p = sir_exact(n_agents, beta, gamma, 1, k, T)
    assert _within(r.estimate, p, n), (r.estimate, p)


CONFIGS = [
    ("race8", lambda: G.race_network(8), "", lambda net: cs.var_query(net.variable_id("count"), 30.0, threshold=14)),
    ("fischer32", lambda: G.fischer_network(32), "P0.cs", lambda net: cs.reach_query(300.0)),
    ("sir100", lambda: G.sir_network(100, 0.3, 0.1), "", lambda net: cs.var_query(net.variable_id("I"), 30.0, threshold=40)),
]


@pytest.mark.parametrize("name,gen,locq,qf", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_gpu_inside_ch_interval_of_reference_algorithm(name, gen, locq, qf):
    xml, _ = gen()
    net = cs.ParsedNetwork.from_string(xml, locq)
    q = qf(net)
    n_ref = 4000
    ref, _ = oracle.run(net, q, seed=31, n_runs=n_ref, mode=oracle.LITERAL, threads=16)
    r = _gpu(net, q, 32, 200_000)
    eps = cs.ch_epsilon(n_ref) + cs.ch_epsilon(200_000)
    print(f"[ch] {name}: gpu {r.estimate:.4f} reference-algorithm {ref.estimate:.4f} eps {eps:.4f}", flush=True)
    assert abs(r.estimate - ref.estimate) <= eps


def test_gpu_zeno_cutoff_matches_oracle():
    net = cs.ParsedNetwork.from_string(zeno(), "Z.never")
    q = cs.reach_query(10.0, max_steps=3000)
    r = _gpu(net, q, 1, 64, per_run=("steps", "time", "digest"))
    o, ro = oracle.run(net, q, seed=1, n_runs=64, per_run=("steps", "time", "digest"))
    assert r.truncated == 64 and (r.per_run["steps"] == 3000).all()
    assert np.array_equal(r.per_run["digest"], ro["digest"])


def _random_update_model(rng, n_vars=4, n_upd=6):
    ops = ["+", "-", "*", "/", "%"]
    def ex(d):
        if d == 0 or rng.random() < 0.3:
            return rng.choice([f"v{rng.integers(0, n_vars)}", str(int(rng.integers(1, 7))),
                               f"{rng.uniform(0.5, 3):.4f}"])
        o = rng.choice(ops + ["neg"])
        if o == "neg":
            return f"-({ex(d - 1)})"
        return f"({ex(d - 1)}) {o} ({ex(d - 1)})"
    upds = []
    for _ in range(n_upd):
        if rng.random() < 0.3:  # `c ? a : b` is only accepted at the top of an assignment
            rhs = f"v{rng.integers(0, n_vars)} > {int(rng.integers(0, 5))} ? {ex(2)} : {ex(2)}"
        else:
            rhs = ex(3)
        upds.append(f"v{rng.integers(0, n_vars)} = {rhs}")
    decl = "\n".join(f"double v{i} = {i + 1};" for i in range(n_vars))
    body = ('<location id="id0"><name>a</name><label kind="exponentialrate">2</label></location>'
            '<init ref="id0"/><transition><source ref="id0"/><target ref="id0"/>'
            f'<label kind="assignment">{", ".join(upds)}</label></transition>')
    return (f'<nta><declaration>{decl}</declaration><template><name>T</name><declaration></declaration>'
            f'{body}</template><system>system T;</system></nta>')


def test_gpu_bytecode_matches_tree_evaluation_on_random_expressions():
    """Chains of random assignments (?:, %, /, unary minus) fired many
    times: variable bits enter the digest, so GPU bytecode and oracle tree
    evaluation must agree exactly (NaNs compare as one value: their payload
    bits are not portable)."""
    rng = np.random.default_rng(7)
    checked = 0
    for k in range(24):
        xml = _random_update_model(rng)
        try:
            net = cs.ParsedNetwork.from_string(xml, "")
        except cs.SmcError:
            continue
        q = cs.var_query(0, 3.0, is_max=True, estimate=True)
        o, ro = oracle.run(net, q, seed=k, n_runs=32, per_run=("digest", "value"))
        r = _gpu(net, q, k, 32, per_run=("digest", "value"))
        assert np.array_equal(r.per_run["digest"], ro["digest"]), xml
        checked += 1
    assert checked >= 20


def _full_operator_model(rng):
    """Two components over four variables, built as network arrays so that
    every operator of expressions.cu:10-95 appears: component 0 loops at rate
    2 through two edges whose guards, weights and updates are random trees
    (both edges enabled at once exercises the weighted pick); component 1
    loops at rate 1 + 0.5 !(random tree), a rate expression that changes
    with the variables."""
    net = RawNetwork()
    for k in range(4):
        net.var(float(k + 1))
    t = lambda d=3: random_tree(rng, net, 4, d)
    l0 = net.component([{"rate": net.lit(2.0)}])
    for _ in range(2):
        guards = [(int(rng.choice([0, 1, 2, 3, 4])), int(rng.integers(0, 4)), t(2))] if rng.random() < 0.6 else []
        upds = [(int(rng.integers(0, 4)), t()) for _ in range(int(rng.integers(1, 4)))]
        net.edge(l0, l0, weight=t(2), guards=guards, upds=upds)
    rate = net.expr(E_ADD, net.lit(1.0), net.expr(E_MUL, net.lit(0.5), net.expr(E_NOT, t())))
    l1 = net.component([{"rate": rate}])
    net.edge(l1, l1, upds=[(int(rng.integers(0, 4)), t())])
    return net.build()


def test_gpu_bytecode_matches_oracle_on_every_operator():
    """Random trees over the whole operator set (sqrt, ^, &&, ||, !, ==, !=,
    <, <=, >, >=, %, ?:, random, unary minus) in updates, guards, weights and
    a rate: the trajectory digest holds every fired edge and the final
    variable bits, so the GPU bytecode must agree with the oracle's tree
    evaluation exactly."""
    rng = np.random.default_rng(11)
    checked = 0
    for k in range(40):
        net = _full_operator_model(rng)
        try:
            sim = cs.GpuSimulator(net)
        except cs.SmcError:
            continue  # beyond the bytecode's register or length limit
        try:
            q = cs.var_query(0, 3.0, is_max=True, estimate=True)
            print(f"[gpu ops] model {k}", flush=True)
            r = sim.run(q, seed=k, n_runs=64, per_run=("digest", "value", "transitions"))
        finally:
            sim.close()
        o, ro = oracle.run(net, q, seed=k, n_runs=64, per_run=("digest", "value", "transitions"))
        assert np.array_equal(r.per_run["transitions"], ro["transitions"]), k
        assert np.array_equal(r.per_run["digest"], ro["digest"]), k
        checked += 1
    assert checked >= 30
