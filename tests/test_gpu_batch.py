"""Batched multi-instance launches, graph-replayed rounds and the autotuner.

A batch runs instances with uneven component counts in one persistent
launch (one slots-per-lane value, lanes per run per instance) and reduces
every instance on the GPU; each replay of its captured graph simulates the
next run range of every instance.  Per-run results do not depend on the
launch, so every per-instance count must equal the count of an eager
smc_run over the same run range.
"""
from __future__ import annotations

import numpy as np
import pytest

import cusmc_amd as cs
from cusmc_amd import generators as G

pytestmark = pytest.mark.gpu


def _instances():
    specs = []
    for n in (4, 9, 16, 33):
        xml, _ = G.fischer_network(n)
        specs.append((cs.ParsedNetwork.from_string(xml, "P0.cs"), lambda net: cs.reach_query(60.0)))
    for n, beta in ((16, 0.4), (40, 0.3), (100, 0.25)):
        xml, _ = G.sir_network(n, beta, 0.1)
        specs.append((cs.ParsedNetwork.from_string(xml, ""),
                      lambda net, n=n: cs.var_query(net.variable_id("I"), 30.0, threshold=n // 3)))
    xml, _ = G.race_network(8)
    specs.append((cs.ParsedNetwork.from_string(xml, ""),
                  lambda net: cs.var_query(net.variable_id("count"), 30.0, threshold=14)))
    return [(net, qf(net)) for net, qf in specs]


@pytest.fixture(scope="module")
def sims():
    inst = _instances()
    sims = [(cs.GpuSimulator(net), q) for net, q in inst]
    yield sims
    for s, _ in sims:
        s.close()


def test_batch_matches_eager_runs(sims):
    n = 3000
    items = [(s, q, 100 + k, n) for k, (s, q) in enumerate(sims)]
    b = cs.Batch(items)
    try:
        res, st = b.run(1)
    finally:
        b.close()
    print(f"[batch] groups={st.groups} runs={st.runs} graph_ms={st.graph_ms:.2f}", flush=True)
    assert st.runs == n * len(sims) and st.groups == 1  # one persistent launch
    for (s, q), r, k in zip(sims, res, range(len(sims))):
        e = s.run(q, seed=100 + k, n_runs=n)
        assert r.hits == e.hits and r.steps == e.steps and r.transitions == e.transitions
        assert r.sum == pytest.approx(e.sum, rel=1e-12)


def test_graph_rounds_interleaved_with_eager_work(sims):
    """Three replays of the captured round with eager runs and syncs between
    them: replay r must reproduce the eager result of run range r."""
    n = 1500
    items = [(s, q, 7, n) for s, q in sims[:4]]
    b = cs.Batch(items)
    try:
        for rnd in range(3):
            res, _ = b.run(1)
            for (s, q), r in zip(sims[:4], res):
                e = s.run(q, seed=7, n_runs=n, run_begin=rnd * n)
                assert r.hits == e.hits and r.steps == e.steps, (rnd, r.hits, e.hits)
        res, st = b.run(2)  # rounds 3 and 4 in one call
        assert st.rounds == 2
        for (s, q), r in zip(sims[:4], res):
            e = s.run(q, seed=7, n_runs=2 * n, run_begin=3 * n)
            assert r.hits == e.hits and r.transitions == e.transitions
    finally:
        b.close()


def test_autotune_picks_a_shape_and_keeps_results(sims):
    s, q = sims[3]  # fischer-33 -> 64 leaves
    base = s.run(q, seed=5, n_runs=2000, lanes_per_run=64, per_run=("digest",))
    best = s.autotune(q, seed=1, sample_runs=8192)
    assert best.lanes_per_run in (64, 32, 16, 8, 4)
    tuned = s.run(q, seed=5, n_runs=2000, per_run=("digest",))
    assert tuned.lanes_per_run == best.lanes_per_run
    assert np.array_equal(base.per_run["digest"], tuned.per_run["digest"])


def test_batch_with_autotune(sims):
    items = [(s, q, 11, 2000) for s, q in sims]
    b = cs.Batch(items, tune=True)
    try:
        res, st = b.run(1)
    finally:
        b.close()
    for (s, q), r in zip(sims, res):
        e = s.run(q, seed=11, n_runs=2000)
        assert r.hits == e.hits and r.steps == e.steps


def test_c4_batch_bit_exact_vs_oracle():
    """The batched workload at its real shape: 256 SIR networks (N
    log-uniform in [16, 512], beta and gamma from the seed, query
    Pr[<=100](<> I > N/5)) and 256 Fischer networks (N in {4..64},
    Pr[<=300](<> P0.cs)) in one batch round; each instance's hits, steps,
    transitions and value sums over its runs equal the canonical oracle's on
    the same runs (integer counts: any difference in any trajectory shows)."""
    import oracle

    specs = G.batch_instances()
    assert len(specs) == 512
    ns = [d["n"] for d in specs]
    assert max(ns[:256]) >= 400 and min(ns[:256]) <= 20 and max(ns[256:]) >= 60
    n_runs = 48
    nets, qs, sims = [], [], []
    try:
        for d in specs:
            net = cs.ParsedNetwork.from_string(d["xml"], d["loc_query"])
            if d["kind"] == "sir":
                q = cs.var_query(net.variable_id(d["var"]), d["time_bound"], threshold=d["threshold"])
            else:
                q = cs.reach_query(d["time_bound"])
            nets.append(net)
            qs.append(q)
            sims.append(cs.GpuSimulator(net))
        items = [(s, q, 900 + k, n_runs) for k, (s, q) in enumerate(zip(sims, qs))]
        print(f"[c4] 512 instances x {n_runs} runs, one batch round", flush=True)
        b = cs.Batch(items)
        try:
            res, st = b.run(1)
        finally:
            b.close()
        assert st.runs == 512 * n_runs
        print(f"[c4] groups={st.groups} graph_ms={st.graph_ms:.1f}; oracle", flush=True)
        for k, (net, q, r) in enumerate(zip(nets, qs, res)):
            o, _ = oracle.run(net, q, seed=900 + k, n_runs=n_runs, threads=16)
            assert (r.hits, r.steps, r.transitions) == (o.hits, o.steps, o.transitions), (k, specs[k]["kind"], ns[k])
            assert r.sum == o.sum, k
    finally:
        for s in sims:
            s.close()


def test_c4_instances_inside_ch_interval_of_reference_algorithm():
    """The largest SIR and Fischer instances of the batch and two others at
    1e5 runs: each estimate lies inside the Chernoff-Hoeffding interval of
    the reference algorithm (the oracle's LITERAL sampling) on 3000 runs."""
    import oracle

    specs = G.batch_instances()
    sir = max(range(256), key=lambda k: specs[k]["n"])
    fis = 256 + max(range(256), key=lambda k: specs[256 + k]["n"])
    for k in (sir, fis, 17, 300):
        d = specs[k]
        net = cs.ParsedNetwork.from_string(d["xml"], d["loc_query"])
        q = (cs.var_query(net.variable_id(d["var"]), d["time_bound"], threshold=d["threshold"])
             if d["kind"] == "sir" else cs.reach_query(d["time_bound"]))
        sim = cs.GpuSimulator(net)
        try:
            r = sim.run(q, seed=41 + k, n_runs=100_000)
        finally:
            sim.close()
        n_ref = 3000
        ref, _ = oracle.run(net, q, seed=42 + k, n_runs=n_ref, mode=oracle.LITERAL, threads=16)
        eps = cs.ch_epsilon(n_ref) + cs.ch_epsilon(100_000)
        print(f"[c4 ch] instance {k} ({d['kind']} N={d['n']}): gpu {r.estimate:.4f} "
              f"reference-algorithm {ref.estimate:.4f} eps {eps:.4f}", flush=True)
        assert abs(r.estimate - ref.estimate) <= eps


def test_autotune_choice_persists(tmp_path, monkeypatch):
    """The tuned shape is written to the cache file (SMC_TUNE_CACHE) and a
    second model of the same network and query takes it from there without
    timing anything; the waves-per-CU axis is part of the search."""
    import time
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SMC_TUNE_CACHE", str(cache))
    xml, _ = G.fischer_network(16)
    q = cs.reach_query(60.0)
    a = cs.GpuSimulator(cs.ParsedNetwork.from_string(xml, "P0.cs"))
    b = cs.GpuSimulator(cs.ParsedNetwork.from_string(xml, "P0.cs"))
    try:
        t0 = time.perf_counter()
        first = a.autotune(q, seed=1, sample_runs=8192)
        t_first = time.perf_counter() - t0
        lines = cache.read_text().split("\n")
        assert len([l for l in lines if l]) == 1
        t0 = time.perf_counter()
        second = b.autotune(q, seed=1, sample_runs=8192)
        t_second = time.perf_counter() - t0
        assert (second.lanes_per_run, second.waves_per_cu, second.occupancy, second.tables) == \
               (first.lanes_per_run, first.waves_per_cu, first.occupancy, first.tables)
        assert t_second < t_first / 5, (t_first, t_second)
        assert len([l for l in cache.read_text().split("\n") if l]) == 1
        r = b.run(q, seed=3, n_runs=4000, per_run=("digest",))
        e = a.run(q, seed=3, n_runs=4000, lanes_per_run=16, per_run=("digest",))
        assert np.array_equal(r.per_run["digest"], e.per_run["digest"])
    finally:
        a.close()
        b.close()
