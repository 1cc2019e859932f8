"""Importance sampling on the GPU against the oracle.

* no tilt: the tilted advance kernel, run to the end or in lock-step
  macro-steps, reproduces the plain run kernel's digests bit for bit;
* fixed per-location tilts: digests, flags and the FP64 log2 likelihood
  ratios equal the oracle's bit for bit, with and without early stop;
* the recorded cross-entropy pairs (features, wins, ln2 dt r) equal the
  oracle's macro-step records;
* the MLP forward/backward (MFMA GEMMs) and the full cross-entropy gradient
  agree with the NumPy restatement (oracle/mlp.py) to bf16/fp32 tolerance;
* the trained estimator is unbiased against the exact CTMC, and on the
  rare-event SIR query the trained network gains over the untrained one.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cusmc_amd as cs
import oracle
from oracle import mlp
from cusmc_amd import generators as G

from _models import sir_exact

pytestmark = pytest.mark.gpu


def _sir(n, beta, gamma, i0=1):
    xml, _ = G.sir_network(n, beta, gamma, initial_infected=i0)
    return cs.ParsedNetwork.from_string(xml, "")


MODELS = {
    "fischer6": lambda: (cs.ParsedNetwork.from_string(G.fischer_network(6)[0], "P0.cs"), lambda n: cs.reach_query(40.0)),
    "race8": lambda: (cs.ParsedNetwork.from_string(G.race_network(8)[0], ""),
                      lambda n: cs.var_query(n.variable_id("count"), 40.0, threshold=20)),
    "sir30": lambda: (_sir(30, 0.3, 0.1, 2), lambda n: cs.var_query(n.variable_id("I"), 40.0, threshold=12)),
    "fischer33": lambda: (cs.ParsedNetwork.from_string(G.fischer_network(33)[0], "P0.cs"), lambda n: cs.reach_query(30.0)),
}


@pytest.fixture(scope="module")
def env():
    out = {}
    for name, mk in MODELS.items():
        net, qf = mk()
        q = qf(net)
        sim = cs.GpuSimulator(net)
        s = cs.ImportanceSampler(sim, q, runs_per_round=4096, rounds=1, epochs=2, seed=3)
        out[name] = (net, q, sim, s)
    yield out
    for net, q, sim, s in out.values():
        s.close()
        sim.close()


def _theta(net, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, net.net.n_locations).astype(np.float32)


def _var_scale(net, q):
    n = net.net
    kinds = [n.var_kind[v] for v in range(n.n_vars)]
    imax = float(n.n_components)
    for v in range(n.n_vars):
        if kinds[v] != 1:
            imax = max(imax, abs(n.var_init[v]))
    return np.array([np.float32(1.0 / q.time_bound) if kinds[v] == 1 else np.float32(1.0 / max(1.0, imax))
                     for v in range(n.n_vars)] or [1.0], dtype=np.float32)


@pytest.mark.parametrize("name", sorted(MODELS))
def test_untilted_matches_plain_kernel(env, name):
    net, q, sim, s = env[name]
    n = 2000
    plain = sim.run(q, seed=3, n_runs=n, per_run=("digest", "flag", "steps", "time"))
    for lock in (False, True):
        print(f"[is untilted] {name} lockstep={lock}", flush=True)
        r = s.sample(seed=3, n_runs=n, lockstep=lock, early_stop=False)
        assert np.array_equal(r.per_run["digest"], plain.per_run["digest"]), lock
        assert np.array_equal(r.per_run["flag"], plain.per_run["flag"])
        assert np.array_equal(r.per_run["steps"], plain.per_run["steps"])
        assert np.array_equal(r.per_run["time"], plain.per_run["time"])
        assert np.all(r.per_run["log2lr"] == 0.0)
        assert r.hits == plain.hits


@pytest.mark.parametrize("name", sorted(MODELS))
def test_fixed_tilt_matches_oracle(env, name):
    net, q, sim, s = env[name]
    theta = _theta(net, 5)
    n = 1500
    for lock, early in ((False, False), (True, False), (True, True)):
        print(f"[is fixed tilt] {name} lockstep={lock} early_stop={early}", flush=True)
        r = s.sample(seed=9, n_runs=n, theta=theta, lockstep=lock, early_stop=early)
        o = oracle.is_run(net, q, seed=9, n_runs=n, theta=theta, early_stop=early, threads=0)
        assert np.array_equal(r.per_run["digest"], o["digest"]), (lock, early)
        assert np.array_equal(r.per_run["flag"], o["flag"])
        assert np.array_equal(r.per_run["log2lr"], o["log2lr"])
        assert np.array_equal(r.per_run["steps"], o["steps"])
        assert np.array_equal(r.per_run["time"], o["time"])


@pytest.mark.parametrize("name", sorted(MODELS))
def test_staged_tables_match_oracle(env, name, monkeypatch):
    """The advance kernel with the model tables staged in LDS (persistent
    four-wave workgroups): fixed-tilt runs equal the oracle's bit for bit,
    recorded pairs too."""
    net, q, sim, _ = env[name]
    monkeypatch.setenv("SMC_IS_TABLES", "1")
    s = cs.ImportanceSampler(sim, q, runs_per_round=4096, rounds=1, epochs=2, seed=3)
    monkeypatch.delenv("SMC_IS_TABLES")
    try:
        theta = _theta(net, 5)
        n = 1500
        for lock, early in ((False, False), (True, True)):
            print(f"[is staged tables] {name} lockstep={lock} early_stop={early}", flush=True)
            r = s.sample(seed=9, n_runs=n, theta=theta, lockstep=lock, early_stop=early)
            o = oracle.is_run(net, q, seed=9, n_runs=n, theta=theta, early_stop=early, threads=0)
            for k in ("digest", "flag", "log2lr", "steps", "time"):
                assert np.array_equal(r.per_run[k], o[k]), (k, lock, early)
        runs = np.arange(100, 108, dtype=np.uint64)
        X, A, B = s.record(runs, seed=9, theta=theta, early_stop=True)
        X0, A0, B0 = env[name][3].record(runs, seed=9, theta=theta, early_stop=True)
        assert np.array_equal(X, X0) and np.array_equal(A, A0) and np.array_equal(B, B0)
    finally:
        s.close()


@pytest.mark.parametrize("name", ["sir30", "fischer6", "race8"])
def test_recorded_pairs_match_oracle(env, name):
    net, q, sim, s = env[name]
    theta = _theta(net, 6)
    runs = np.arange(100, 112, dtype=np.uint64)
    print(f"[is record] {name}", flush=True)
    X, A, B = s.record(runs, seed=9, theta=theta, early_stop=True)
    scale = _var_scale(net, q)
    per_run = [oracle.is_record(net, q, seed=9, run=int(r), theta=theta, var_scale=scale,
                                fpad=s.features_padded, lpad=s.locations_padded, early_stop=True) for r in runs]
    # slots are numbered macro-step-major, run-minor
    ox, oa, ob = [], [], []
    for k in range(max(len(p[0]) for p in per_run)):
        for p in per_run:
            if k < len(p[0]):
                ox.append(p[0][k]); oa.append(p[1][k]); ob.append(p[2][k])
    assert len(X) == len(ox)
    assert np.array_equal(X, np.array(ox))
    assert np.array_equal(A, np.array(oa))
    assert np.array_equal(B, np.array(ob))


def _random_params(s, seed, w3_scale=0.3):
    rng = np.random.default_rng(seed)
    fp, H, lp = s.features_padded, s.hidden, s.outputs_padded
    p = np.concatenate([rng.normal(0, math.sqrt(2.0 / fp), fp * H), rng.normal(0, 0.1, H),
                        rng.normal(0, math.sqrt(2.0 / H), H * H), rng.normal(0, 0.1, H),
                        rng.normal(0, w3_scale, H * lp), rng.normal(0, 0.3, lp)]).astype(np.float32)
    assert p.size == s.n_params
    return p


def _tiltable(net, s):
    n = net.net
    t = np.zeros(s.locations_padded, dtype=bool)
    for l in range(n.n_locations):
        t[l] = n.locations[l].type <= 1  # not branchpoint / urgent / committed
    return t


def _group_tiltable(net, s):
    t = _tiltable(net, s)
    gt = np.zeros(s.outputs_padded, dtype=bool)
    for l in range(net.net.n_locations):
        gt[s.group[l]] |= t[l]
    return gt


def _np_theta(s, net, p, x):
    """NumPy MLP: group outputs, then theta per location."""
    thg, cache = mlp.forward(p, x, s.features_padded, s.hidden, s.outputs_padded, s.config.tilt_max,
                             _group_tiltable(net, s))
    return mlp.expand(thg, s.group[:net.net.n_locations], s.locations_padded), cache


def test_tilt_groups_follow_templates(env):
    net, q, sim, s = env["sir30"]
    names = [net.location_name(l) for l in range(net.net.n_locations)]
    assert s.n_groups == 3
    for l, nm in enumerate(names):
        assert s.group[l] == {"S": 0, "I": 1, "R": 2}[nm]
    net, q, sim, s = env["fischer6"]
    # pid differs per process, the shape does not; P0 holds the goal location
    assert s.n_groups == 8
    nl = net.net.n_locations
    assert list(s.group[:4]) == [0, 1, 2, 3]
    assert all(list(s.group[4 * k:4 * k + 4]) == [4, 5, 6, 7] for k in range(1, nl // 4))
    p = cs.ImportanceSampler(sim, q, per_location=True, runs_per_round=1024, rounds=1)
    try:
        assert p.n_groups == net.net.n_locations
        assert np.array_equal(p.group[:net.net.n_locations], np.arange(net.net.n_locations))
    finally:
        p.close()


def test_mlp_forward_backward_matches_numpy(env):
    net, q, sim, s = env["sir30"]
    p = _random_params(s, 11)
    s.set_params(p)
    rng = np.random.default_rng(12)
    rows = 200  # not a multiple of 64: the library pads
    x = rng.normal(0.0, 1.0, (rows, s.features_padded)).astype(np.float32)
    g = rng.normal(0.0, 1.0, (rows, s.locations_padded)).astype(np.float32)
    tl = _tiltable(net, s)
    theta, grad = s.mlp_eval(x, g)
    th_np, cache = _np_theta(s, net, p, x)
    err = np.abs(theta - th_np)
    print(f"[mlp] theta max err {err.max():.3e} norm {np.linalg.norm(th_np):.3e}", flush=True)
    assert np.linalg.norm(theta - th_np) <= 2e-3 * np.linalg.norm(th_np)
    assert np.all(theta[:, ~tl] == 0.0)
    assert np.all(np.abs(theta) <= s.config.tilt_max)
    gm = np.where(tl[None, :] & (np.abs(theta) < s.config.tilt_max), g, 0.0)
    gg = mlp.group_sum(gm, s.group[:net.net.n_locations], s.outputs_padded)
    g_np = mlp.backward(p, cache, gg, s.features_padded, s.hidden, s.outputs_padded)
    rel = np.linalg.norm(grad - g_np) / np.linalg.norm(g_np)
    print(f"[mlp] grad rel err {rel:.3e}", flush=True)
    assert rel < 2e-2
    cos = float(grad @ g_np / (np.linalg.norm(grad) * np.linalg.norm(g_np)))
    assert cos > 0.999


def test_fused_forward_matches_gemm_path(env, monkeypatch):
    """The one-kernel forward pass and the three-GEMM path form the same
    products in the same order: tilts, gradients and a whole estimate agree
    bit for bit."""
    net, q, sim, _ = env["sir30"]
    a = cs.ImportanceSampler(sim, q, runs_per_round=4096, rounds=1, epochs=2, seed=3)
    monkeypatch.setenv("SMC_IS_FORWARD", "gemm")
    b = cs.ImportanceSampler(sim, q, runs_per_round=4096, rounds=1, epochs=2, seed=3)
    monkeypatch.delenv("SMC_IS_FORWARD")
    try:
        p = _random_params(a, 31, w3_scale=0.05)
        a.set_params(p)
        b.set_params(p)
        rng = np.random.default_rng(32)
        rows = 200
        x = rng.normal(0.0, 1.0, (rows, a.features_padded)).astype(np.float32)
        g = rng.normal(0.0, 1.0, (rows, a.locations_padded)).astype(np.float32)
        ta, ga = a.mlp_eval(x, g)
        tb, gb = b.mlp_eval(x, g)
        assert np.array_equal(ta.view(np.uint32), tb.view(np.uint32))
        assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
        ea = a.estimate(seed=9, batch_runs=4096, max_runs=8192, rel_target=0.0)
        eb = b.estimate(seed=9, batch_runs=4096, max_runs=8192, rel_target=0.0)
        print(f"[fused forward] estimate {ea.estimate!r} hits {ea.hits} runs {ea.n_runs}", flush=True)
        assert ea.n_runs == eb.n_runs and ea.hits == eb.hits and ea.estimate == eb.estimate
    finally:
        a.close()
        b.close()


def test_ce_gradient_matches_oracle(env):
    net, q, sim, s = env["sir30"]
    p = _random_params(s, 21, w3_scale=0.05)
    s.set_params(p)
    theta = _theta(net, 22)
    runs = np.arange(0, 48, dtype=np.uint64)
    w = np.random.default_rng(23).uniform(0.1, 1.0, len(runs)).astype(np.float32)
    X, A, B = s.record(runs, seed=4, theta=theta, weights=w, early_stop=True)
    grad = s.ce_gradient()
    # the same pairs from the oracle, with each pair's run weight
    scale = _var_scale(net, q)
    per_run = [oracle.is_record(net, q, seed=4, run=int(r), theta=theta, var_scale=scale,
                                fpad=s.features_padded, lpad=s.locations_padded, early_stop=True) for r in runs]
    ox, oa, ob, ow = [], [], [], []
    for k in range(max(len(pr[0]) for pr in per_run)):
        for i, pr in enumerate(per_run):
            if k < len(pr[0]):
                ox.append(pr[0][k]); oa.append(pr[1][k]); ob.append(pr[2][k]); ow.append(w[i])
    xo = mlp.bf16_bits_to_f32(np.array(ox))
    assert np.array_equal(X, np.array(ox))
    tl = _tiltable(net, s)
    th_np, cache = _np_theta(s, net, p, xo)
    G_np = mlp.ce_output_grad(th_np, np.array(oa), np.array(ob), np.array(ow), tl, s.config.tilt_max)
    gg = mlp.group_sum(G_np, s.group[:net.net.n_locations], s.outputs_padded)
    g_np = mlp.backward(p, cache, gg, s.features_padded, s.hidden, s.outputs_padded)
    rel = np.linalg.norm(grad - g_np) / np.linalg.norm(g_np)
    cos = float(grad @ g_np / (np.linalg.norm(grad) * np.linalg.norm(g_np)))
    print(f"[ce grad] pairs={len(ox)} rel err {rel:.3e} cos {cos:.6f}", flush=True)
    assert rel < 2e-2 and cos > 0.999


def test_is_estimate_unbiased_against_ctmc():
    n_ag, beta, gamma, k, T = 20, 0.2, 0.1, 19, 30.0
    exact = sir_exact(n_ag, beta, gamma, 1, k, T)
    net = _sir(n_ag, beta, gamma, 1)
    q = cs.var_query(net.variable_id("I"), T, threshold=k - 1)
    sim = cs.GpuSimulator(net)
    s = cs.ImportanceSampler(sim, q, rounds=3, runs_per_round=20000, epochs=10, seed=5)
    try:
        rounds = s.train()
        print("[is ctmc] rounds", rounds, flush=True)
        est = s.estimate(seed=77, batch_runs=32768, max_runs=1 << 19, rel_target=0.03)
        se = est.half_width / 1.959963984540054
        print(f"[is ctmc] est={est.estimate:.4e} exact={exact:.4e} rel_hw={est.rel_half_width:.3f} "
              f"runs={est.n_runs} hits={est.hits}", flush=True)
        assert abs(est.estimate - exact) < 4.0 * se
    finally:
        s.close()
        sim.close()


def test_trained_proposal_gains_over_untrained():
    """SIR, N=100, beta=0.15, gamma=0.1, one initial infected:
    Pr[<=50](<> I >= 90) = 6.638e-6 (exact CTMC)."""
    net = _sir(100, 0.15, 0.1, 1)
    q = cs.var_query(net.variable_id("I"), 50.0, threshold=89)
    exact = 6.638e-6
    sim = cs.GpuSimulator(net)
    s = cs.ImportanceSampler(sim, q, rounds=5, runs_per_round=20000, epochs=20, seed=7)
    try:
        n = 32768
        before = s.sample(seed=31, n_runs=n, use_mlp=True, lockstep=True, early_stop=True)
        rounds = s.train()
        for r in rounds:
            print(f"[is gain] round level={r['level']} elites={r['elites']} hits={r['hits']} "
                  f"est={r['estimate']:.3e} ms={r['ms']:.0f}", flush=True)
        after = s.estimate(seed=31, batch_runs=n, max_runs=n, rel_target=0.0001)
        se = after.half_width / 1.959963984540054
        print(f"[is gain] untrained hits={before.hits}/{n}; trained hits={after.hits}/{after.n_runs} "
              f"est={after.estimate:.4e} rel_hw={after.rel_half_width:.3f} ess={after.ess:.0f}", flush=True)
        assert before.hits <= 5
        assert after.hits >= 100 * max(before.hits, 1)
        assert after.rel_half_width < 0.25
        assert abs(after.estimate - exact) < 4.0 * se
    finally:
        s.close()
        sim.close()


def test_trained_window_and_rate_tilts_gain_on_race8():
    """race-8 mixes uniform windows (L0, L2: invariant x <= U, guard x >= L)
    and exponential locations (L1, L3).  Pr[<=20](<> count > 54) is about
    4e-5: the proposal trained with window and rate tilts against plain
    simulation of 4e7 runs."""
    net = cs.ParsedNetwork.from_string(G.race_network(8)[0], "")
    q = cs.var_query(net.variable_id("count"), 20.0, threshold=54)
    sim = cs.GpuSimulator(net)
    s = cs.ImportanceSampler(sim, q, rounds=5, runs_per_round=20000, epochs=20, seed=7)
    try:
        plain = sim.run(q, seed=5, n_runs=40_000_000)
        sp = math.sqrt(plain.estimate * (1 - plain.estimate) / plain.n_runs)
        n = 32768
        before = s.sample(seed=31, n_runs=n, use_mlp=True, lockstep=True, early_stop=True)
        rounds = s.train()
        for r in rounds:
            print(f"[is race8] round level={r['level']} elites={r['elites']} hits={r['hits']} "
                  f"est={r['estimate']:.3e} ms={r['ms']:.0f}", flush=True)
        after = s.estimate(seed=31, batch_runs=n, max_runs=n, rel_target=0.0001)
        se = after.half_width / 1.959963984540054
        print(f"[is race8] plain {plain.hits}/{plain.n_runs} = {plain.estimate:.3e}; untrained hits "
              f"{before.hits}/{n}; trained hits {after.hits}/{after.n_runs} est {after.estimate:.3e} "
              f"rel_hw {after.rel_half_width:.3f}", flush=True)
        assert plain.hits >= 400
        assert before.hits <= 10
        assert after.hits >= 50 * max(before.hits, 1)
        assert abs(after.estimate - plain.estimate) < 4.0 * math.sqrt(se ** 2 + sp ** 2)
        # relative variance per run: the trained proposal's is several times
        # smaller than plain simulation's (1 - p) / p
        p = plain.estimate
        rel_var_is = (after.rel_half_width / 1.959963984540054) ** 2 * after.n_runs
        assert rel_var_is < (1 - p) / p / 5, (rel_var_is, (1 - p) / p)
    finally:
        s.close()
        sim.close()
