"""smc_run_noalloc (the entry point that allocates nothing: its buffers live
in a fixed per-GPU scratch array) against smc_run on the same runs: two
calls in a row, the second smaller, must give identical totals."""
from __future__ import annotations

import ctypes as C

import pytest

import cusmc_amd as cs
from cusmc_amd import _capi as A
from cusmc_amd import generators as G

pytestmark = pytest.mark.gpu


def _noalloc(sim, q, seed, n, run_begin=0):
    res = A.Result()
    A.check(A.lib().smc_run_noalloc(sim._h, C.byref(q), C.c_uint64(seed), C.c_uint64(run_begin), C.c_uint64(n),
                                    None, C.byref(res)))
    return res


def _cases():
    fis = cs.ParsedNetwork.from_string(G.fischer_network(32)[0], "P0.cs")
    sir = cs.ParsedNetwork.from_string(G.sir_network(64, 0.3, 0.1)[0], "")
    return [("fischer32", fis, cs.reach_query(300.0), True),
            ("sir64", sir, cs.var_query(sir.variable_id("I"), 60.0, threshold=20), False)]


@pytest.mark.parametrize("k", [0, 1], ids=["fischer32-tuned", "sir64"])
def test_noalloc_matches_smc_run_on_two_calls(k):
    name, net, q, tune = _cases()[k]
    sim = cs.GpuSimulator(net)
    try:
        if tune:
            sim.autotune(q, seed=3, sample_runs=8192)
        for n, rb in ((300_000, 5), (3000, 11)):
            print(f"[noalloc] {name} runs={n} run_begin={rb}", flush=True)
            a = _noalloc(sim, q, 77, n, rb)
            b = sim.run(q, seed=77, n_runs=n, run_begin=rb)
            assert a.n_runs == n
            assert (a.hits, a.steps, a.transitions, a.truncated) == (b.hits, b.steps, b.transitions, b.truncated)
            assert (a.sum, a.sumsq, a.vmin, a.vmax, a.estimate) == (b.sum, b.sumsq, b.vmin, b.vmax, b.estimate)
            assert a.lanes_per_run == b.lanes_per_run
    finally:
        sim.close()


def test_noalloc_rejects_runs_beyond_its_scratch():
    _, net, q, _ = _cases()[0]
    sim = cs.GpuSimulator(net)
    try:
        res = A.Result()
        rc = A.lib().smc_run_noalloc(sim._h, C.byref(q), C.c_uint64(1), C.c_uint64(0), C.c_uint64((1 << 20) + 1),
                                     None, C.byref(res))
        assert rc == A.SMC_E_LIMIT
        assert b"1048576" in A.lib().smc_last_error()
        ok = _noalloc(sim, q, 1, 1 << 20)  # the maximum itself runs
        assert ok.n_runs == 1 << 20
    finally:
        sim.close()
