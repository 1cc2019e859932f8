"""The algorithmic work figures bench.py divides by kernel time are the
oracle cost model's per-race-step averages for each workload; a change of
either side must show here."""
from __future__ import annotations

import importlib.util
from pathlib import Path

import pytest

import oracle

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("workload,n_runs", [("fischer32", 1500), ("race8", 5000), ("sir350", 300)])
def test_cost_per_step_constants(workload, n_runs):
    b = _bench()
    net, q, _, _ = b._models(workload)
    steps, trans, ops = oracle.cost(net, q, seed=4242, n_runs=n_runs, threads=0)
    assert steps > 0
    per_step = ops / steps
    assert per_step == pytest.approx(b.COST_PER_STEP[workload], rel=5e-3), per_step


def test_cost_counts_every_race_step():
    """The work of a run grows with its race steps; a deterministic seed gives
    the same count twice."""
    b = _bench()
    net, q, _, _ = b._models("race8")
    a = oracle.cost(net, q, seed=1, n_runs=200, threads=1)
    c = oracle.cost(net, q, seed=1, n_runs=200, threads=3)
    assert a == c
    d = oracle.cost(net, q, seed=1, n_runs=400, threads=2)
    assert d[0] > a[0] and d[2] > a[2]
