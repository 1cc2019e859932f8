"""GPU run kernel vs the CPU oracle: bit-exact per-run results.

Each case simulates the same runs (same seed, same run indices) on the GPU
through the C ABI and on the oracle in its canonical sampling mode, and
requires every per-run output to be identical: the trajectory digest (a hash
of every fired edge with its step and component, the final time bits, every
variable's bits and every final location), the flag, the query value, the
step and transition counts, and the end time.
"""
from __future__ import annotations

import numpy as np
import pytest

import cusmc_amd as cs
from cusmc_amd import generators as G

import oracle

from _models import name_locations, receiver_goal

pytestmark = pytest.mark.gpu

PER_RUN = ("flag", "value", "steps", "transitions", "digest", "time")


def _net(models_dir, src, locq=""):
    if src.endswith(".xml"):
        return cs.ParsedNetwork.from_file(models_dir / src, locq)
    kind, _, arg = src.partition(":")
    if kind == "race":
        xml, _ = G.race_network(int(arg))
    elif kind == "fischer":
        xml, _ = G.fischer_network(int(arg))
    elif kind == "sir":
        xml, _ = G.sir_network(int(arg))
    elif kind == "sirsmall":
        xml, _ = G.sir_network(int(arg), 0.6, 0.2, initial_infected=2, seed=5)
    elif kind == "recv":
        # receiver-entered goal: alone (absorbing), beside a live component,
        # beside a component whose rate reads a clock
        live = None if arg == "" else float(arg.rstrip("c"))
        xml = receiver_goal(live, clock_rate=arg.endswith("c"))
    elif kind == "racetest_named":
        xml = name_locations((models_dir / "RaceTest.xml").read_text(), {"id10": "sent", "id1": "got"})
    else:
        raise ValueError(src)
    return cs.ParsedNetwork.from_string(xml, locq)


def _query(net, q):
    kind = q[0]
    if kind == "reach":
        return cs.reach_query(q[1])
    _, var, bound, is_max, est, thr = q
    vid = net.variable_id(var) if isinstance(var, str) else var
    return cs.var_query(vid, bound, is_max=is_max, estimate=est, threshold=thr)


CASES = [
    # (id, model, location query, query, runs, lanes_per_run values)
    ("condTest", "condTest.xml", "Template.l3", ("reach", 10.0), 4000, (0,)),
    ("RaceTest", "RaceTest.xml", "", ("var", 0, 10.0, True, True, -1.0), 4000, (0, 4)),
    ("UppaalWeightTest", "UppaalWeightTest.xml", "Template.l3", ("reach", 60.0), 2000, (0,)),
    ("fischer_10_29", "fischer_10_29.xml", "P0.cs", ("reach", 300.0), 500, (0, 4)),
    ("aloha_25", "aloha_25.xml", "", ("var", "nt", 100.0, True, True, 0.0), 300, (0, 8)),
    ("covid_32", "agentBaseCovid_32_1.xml", "", ("var", "inf", 100.0, True, False, 3.0), 300, (0, 8)),
    ("bluetooth", "bluetoothSmacc.xml", "", ("var", 0, 50.0, True, True, 0.0), 300, (0,)),
    ("fischermulti", "fischermulti.xml", "", ("var", 0, 100.0, False, True, 0.0), 300, (0,)),
    ("race8", "race:8", "", ("var", "count", 100.0, True, False, 39.0), 2000, (0, 2)),
    ("fischer32", "fischer:32", "P0.cs", ("reach", 300.0), 400, (0, 8, 16)),
    ("sir64", "sir:64", "", ("var", "I", 100.0, True, False, 20.0), 100, (0, 16)),
    ("sir350", "sir:350", "", ("var", "I", 100.0, True, False, 104.0), 24, (0,)),
    # goals entered by broadcast receivers (the step after them decides)
    ("recvgoal", "recv:", "", ("reach", 2.0), 4000, (0, 2)),
    ("recvgoal_live", "recv:0.2", "", ("reach", 2.0), 4000, (0, 4)),
    ("recvgoal_clockrate", "recv:0.3c", "", ("reach", 2.0), 4000, (0,)),
    ("RaceTest_receiver_loc", "racetest_named:", "Process7.got", ("reach", 3.0), 4000, (0, 16)),
    ("RaceTest_sender_loc", "racetest_named:", "Process4.sent", ("reach", 3.0), 2000, (0,)),
    ("sir24_loc", "sirsmall:24", "A3.I", ("reach", 8.0), 2000, (0, 8)),
    ("behavior3", "UppaalBehaviorTest3.xml", "c2.g2", ("reach", 20.0), 2000, (0,)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_bit_exact_vs_oracle(case, models_dir):
    name, src, locq, qspec, runs, lane_opts = case
    net = _net(models_dir, src, locq)
    q = _query(net, qspec)
    seed = 0x5EED0000 + len(name)
    ref, ro = oracle.run(net, q, seed=seed, n_runs=runs, run_begin=7, threads=8, per_run=PER_RUN)
    sim = cs.GpuSimulator(net)
    try:
        for S in lane_opts:
            print(f"[gpu parity] {name} lanes_per_run={S} runs={runs}", flush=True)
            res = sim.run(q, seed=seed, n_runs=runs, run_begin=7, lanes_per_run=S, per_run=PER_RUN)
            for k in PER_RUN:
                a, b = res.per_run[k], ro[k]
                if k in ("value", "time"):
                    same = np.array_equal(a.view(np.uint64), b.view(np.uint64))
                else:
                    same = np.array_equal(a, b)
                if not same:
                    bad = np.nonzero(a.view(np.uint64) != b.view(np.uint64))[0] if a.dtype.itemsize == 8 \
                        else np.nonzero(a != b)[0]
                    i = int(bad[0])
                    pytest.fail(f"{name} S={S}: per-run {k} differs at {len(bad)} runs; first run {i}: "
                                f"gpu={a[i]!r} oracle={b[i]!r}")
            assert res.hits == ref.hits
            assert res.steps == ref.steps and res.transitions == ref.transitions
    finally:
        sim.close()


# Runs without digests may stage the model tables in LDS (workgroups of four
# waves): the same per-run results as the oracle, for both register policies.
TABLE_CASES = ["RaceTest", "UppaalWeightTest", "fischer_10_29", "aloha_25", "covid_32", "race8", "fischer32",
               "sir64", "recvgoal_live", "RaceTest_receiver_loc", "sir24_loc"]


@pytest.mark.parametrize("name", TABLE_CASES)
def test_lds_tables_bit_exact_vs_oracle(name, models_dir):
    case = next(c for c in CASES if c[0] == name)
    _, src, locq, qspec, runs, lane_opts = case
    net = _net(models_dir, src, locq)
    q = _query(net, qspec)
    seed = 0x7AB0000 + len(name)
    keys = ("flag", "value", "steps", "transitions", "time")
    ref, ro = oracle.run(net, q, seed=seed, n_runs=runs, run_begin=3, threads=8, per_run=keys)
    sim = cs.GpuSimulator(net)
    try:
        for S in lane_opts:
            for occ, tab in ((0, 1), (3, 1), (3, 2)):
                print(f"[gpu parity, tables in LDS] {name} lanes_per_run={S} occupancy={occ} tables={tab} "
                      f"runs={runs}", flush=True)
                res = sim.run(q, seed=seed, n_runs=runs, run_begin=3, lanes_per_run=S, occupancy=occ, tables=tab,
                              per_run=keys)
                for k in keys:
                    a, b = res.per_run[k], ro[k]
                    same = np.array_equal(a.view(np.uint64), b.view(np.uint64)) if k in ("value", "time") \
                        else np.array_equal(a, b)
                    assert same, f"{name} S={S} occupancy={occ} tables={tab}: per-run {k} differs from the oracle"
                assert res.hits == ref.hits and res.steps == ref.steps
    finally:
        sim.close()


def test_run_range_split_is_consistent(models_dir):
    """Runs are addressed by global index: two calls over halves equal one call."""
    net = _net(models_dir, "fischer:32", "P0.cs")
    q = cs.reach_query(300.0)
    sim = cs.GpuSimulator(net)
    try:
        whole = sim.run(q, seed=11, n_runs=3000, per_run=("digest",))
        a = sim.run(q, seed=11, n_runs=1000, run_begin=0, per_run=("digest",))
        b = sim.run(q, seed=11, n_runs=2000, run_begin=1000, per_run=("digest",))
        assert np.array_equal(whole.per_run["digest"], np.concatenate([a.per_run["digest"], b.per_run["digest"]]))
        assert whole.hits == a.hits + b.hits
    finally:
        sim.close()
