"""The reference's large models on the GPU (xml_files/LargeModels and the
agent-based covid models, committed gzipped as fixtures): per-run results
bit-exact with the canonical oracle, and estimates inside the
Chernoff-Hoeffding interval of the reference algorithm (LITERAL sampling).

CSMA_100: 100 components, 900 variables (100 clocks), 5500 edges.
fischer_100_29: 100 processes. aloha_50: 50 stations, broadcasts.
agentBaseCovid_350 / _100: 350 / 100 agents."""
from __future__ import annotations

import numpy as np
import pytest

import cusmc_amd as cs

import oracle

from _models import large_model

pytestmark = pytest.mark.gpu

PER_RUN = ("flag", "value", "steps", "transitions", "digest", "time")

CASES = [
    # (model, location query, query variable or None, time bound, runs, lanes per run values)
    ("CSMA_100.xml", "Process0.SUCCESS", None, 100.0, 400, (0, 16)),
    ("fischer_100_29.xml", "P0.cs", None, 100.0, 200, (0, 32)),
    ("aloha_50.xml", "", "nt", 100.0, 100, (0,)),
    ("agentBaseCovid_350.xml", "", "inf", 100.0, 200, (0, 32)),
    ("agentBaseCovid_100_1.0.xml", "", "inf", 100.0, 400, (0,)),
]


def _query(net, var, bound):
    if var is None:
        return cs.reach_query(bound)
    return cs.var_query(net.variable_id(var), bound, is_max=True, estimate=True)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_large_model_bit_exact_vs_oracle(case):
    name, locq, var, bound, runs, lane_opts = case
    net = cs.ParsedNetwork.from_string(large_model(name), locq)
    q = _query(net, var, bound)
    seed = 0xB16 + len(name)
    ref, ro = oracle.run(net, q, seed=seed, n_runs=runs, run_begin=3, threads=16, per_run=PER_RUN)
    sim = cs.GpuSimulator(net)
    try:
        for S in lane_opts:
            print(f"[gpu large] {name} lanes_per_run={S} runs={runs}", flush=True)
            res = sim.run(q, seed=seed, n_runs=runs, run_begin=3, lanes_per_run=S, per_run=PER_RUN)
            for k in PER_RUN:
                a, b = res.per_run[k], ro[k]
                same = np.array_equal(a.view(np.uint64), b.view(np.uint64)) if k in ("value", "time") \
                    else np.array_equal(a, b)
                assert same, f"{name} S={S}: per-run {k} differs from the oracle"
            assert res.steps == ref.steps and res.transitions == ref.transitions and res.hits == ref.hits
    finally:
        sim.close()


@pytest.mark.parametrize("name,locq,var,thr", [("fischer_100_29.xml", "P0.cs", None, 0),
                                               ("agentBaseCovid_100_1.0.xml", "", "inf", 17)])
def test_large_model_inside_ch_interval_of_reference_algorithm(name, locq, var, thr):
    net = cs.ParsedNetwork.from_string(large_model(name), locq)
    q = cs.reach_query(100.0) if var is None else cs.var_query(net.variable_id(var), 100.0, threshold=thr)
    sim = cs.GpuSimulator(net)
    try:
        r = sim.run(q, seed=51, n_runs=100_000)
    finally:
        sim.close()
    n_ref = 3000
    ref, _ = oracle.run(net, q, seed=52, n_runs=n_ref, mode=oracle.LITERAL, threads=16)
    eps = cs.ch_epsilon(n_ref) + cs.ch_epsilon(100_000)
    print(f"[large ch] {name}: gpu {r.estimate:.4f} reference-algorithm {ref.estimate:.4f} eps {eps:.4f}", flush=True)
    assert abs(r.estimate - ref.estimate) <= eps
