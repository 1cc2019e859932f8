"""The oracle's importance-sampling mode, pinned on the CPU.

* theta = 0 (and no theta) reproduces the plain canonical race bit for bit;
* the log2 likelihood ratio of a one-component race has the closed form
  -theta + (2^theta - 1) r d, and mean(hit 2^log2lr) is the closed-form
  probability under any tilt;
* a uniform delay on [0, U] tilted by a = 2^theta (hazard times a) has the
  proposal law Pr[d <= T] = 1 - (1 - T/U)^a and the log2 likelihood ratio
  -theta + (1 - a) log2(1 - d/U), and mean(hit 2^log2lr) = T/U;
* on the agent SIR network, tilted estimates match the exact lumped CTMC;
* the recorded macro-step statistics A (wins) and B (ln2 dt r) of a run
  reproduce its log2 likelihood ratio, log2lr = sum_L -theta_L A_L +
  (2^theta_L - 1) B_L / ln2, which ties them to the race itself;
* the NumPy MLP's analytic gradient matches central finite differences, and
  the cross-entropy output gradient is the derivative of w ln2 log2lr.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cusmc_amd as cs
import oracle
from oracle import mlp
from cusmc_amd import generators as G

from _models import single_exponential, single_uniform, sir_exact


def _pads(net):
    nl = net.net.n_locations
    nv = net.net.n_vars
    return nl, nv, -(-(nl + nv + 1) // 64) * 64, -(-nl // 64) * 64


def _sir(n, beta, gamma, i0=1):
    xml, _ = G.sir_network(n, beta, gamma, initial_infected=i0)
    return cs.ParsedNetwork.from_string(xml, "")


def _loc_tilts(net, per_name):
    nl = net.net.n_locations
    return np.array([per_name.get(net.location_name(l), 0.0) for l in range(nl)], dtype=np.float32)


MODELS = {
    "fischer6": lambda: (cs.ParsedNetwork.from_string(G.fischer_network(6)[0], "P0.cs"), lambda n: cs.reach_query(40.0)),
    "race8": lambda: (cs.ParsedNetwork.from_string(G.race_network(8)[0], ""),
                      lambda n: cs.var_query(n.variable_id("count"), 40.0, threshold=20)),
    "sir30": lambda: (_sir(30, 0.3, 0.1, 2), lambda n: cs.var_query(n.variable_id("I"), 40.0, threshold=12)),
}


@pytest.mark.parametrize("name", sorted(MODELS))
def test_zero_tilt_is_the_plain_race(name):
    net, qf = MODELS[name]()
    q = qf(net)
    _, plain = oracle.run(net, q, seed=3, n_runs=300, per_run=("digest", "flag", "steps", "time"))
    for theta in (None, np.zeros(net.net.n_locations, np.float32)):
        r = oracle.is_run(net, q, seed=3, n_runs=300, theta=theta, threads=4)
        assert np.array_equal(r["digest"], plain["digest"])
        assert np.array_equal(r["flag"], plain["flag"])
        assert np.array_equal(r["steps"], plain["steps"])
        assert np.all(r["log2lr"] == 0.0)


def test_one_component_closed_form():
    rate, T, th = 0.7, 2.0, 1.3
    net = cs.ParsedNetwork.from_string(single_exponential(rate), "T.Goal")
    q = cs.reach_query(T)
    theta = _loc_tilts(net, {"l0": th, "Goal": -0.4})
    n = 200000
    r = oracle.is_run(net, q, seed=11, n_runs=n, theta=theta, threads=0)
    # every run's first (and, for hits, only) race is l0's exponential
    d = r["time"]
    fin = np.isfinite(d)
    th32 = float(np.float32(th))  # tilts are float32
    expect = -th32 + (2.0 ** th32 - 1.0) * rate * d[fin]
    hits = r["flag"].astype(bool)
    assert np.allclose(r["log2lr"][hits], expect[hits[fin]], rtol=1e-12, atol=1e-12)
    # unbiased under the tilt: E_q[hit 2^log2lr] = 1 - 2^(-rate T)
    w = np.where(hits, np.exp2(r["log2lr"]), 0.0)
    p = 1.0 - 2.0 ** (-rate * T)
    se = w.std() / math.sqrt(n)
    assert abs(w.mean() - p) < 4.0 * se, (w.mean(), p, se)
    # the tilt raised the hit rate above the base probability
    assert hits.mean() > p


def test_tilted_sir_matches_exact_ctmc():
    n_ag, beta, gamma, k, T = 20, 0.2, 0.1, 19, 30.0
    net = _sir(n_ag, beta, gamma, 1)
    q = cs.var_query(net.variable_id("I"), T, threshold=k - 1)
    exact = sir_exact(n_ag, beta, gamma, 1, k, T)
    assert 1e-4 < exact < 1e-2
    theta = _loc_tilts(net, {"S": 1.2, "I": -0.8})
    n = 40000
    for early in (False, True):
        r = oracle.is_run(net, q, seed=5, n_runs=n, theta=theta, early_stop=early, threads=0)
        w = np.where(r["flag"].astype(bool), np.exp2(r["log2lr"]), 0.0)
        se = w.std() / math.sqrt(n)
        assert abs(w.mean() - exact) < 4.0 * se, (early, w.mean(), exact, se)
        # and the tilt made the event common
        assert r["flag"].mean() > 20 * exact


@pytest.mark.parametrize("name", sorted(MODELS))
def test_macro_statistics_reproduce_log2lr(name):
    net, qf = MODELS[name]()
    q = qf(net)
    nl, nv, fpad, lpad = _pads(net)
    rng = np.random.default_rng(7)
    theta = rng.uniform(-1.0, 1.0, nl).astype(np.float32)
    runs = oracle.is_run(net, q, seed=9, n_runs=40, theta=theta, early_stop=True, threads=1)
    scale = np.full(max(nv, 1), 0.05, np.float32)
    th64 = theta.astype(np.float64)
    for run in range(40):
        X, Aa, Bb = oracle.is_record(net, q, seed=9, run=run, theta=theta, var_scale=scale, fpad=fpad, lpad=lpad,
                                     early_stop=True)
        assert len(X) >= 1
        Asum = Aa.sum(0, dtype=np.float64)[:nl]
        Bsum = Bb.sum(0, dtype=np.float64)[:nl]
        recon = float(np.sum(-th64 * Asum + (np.exp2(th64) - 1.0) * Bsum / math.log(2.0)))
        lr = runs["log2lr"][run]
        assert abs(recon - lr) <= 1e-4 + 1e-5 * abs(lr), (run, recon, lr)
        # the first macro-step starts in the initial state at time 0
        ones = np.flatnonzero(mlp.bf16_bits_to_f32(X[0][:nl]) == 1.0)
        assert len(ones) == net.n_components
        assert X[0][nl + nv] == 0


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    fpad, H, lpad, rows = 8, 6, 5, 7
    n = fpad * H + H + H * H + H + H * lpad + lpad
    p = rng.normal(0.0, 0.6, n)
    x = rng.normal(0.0, 1.0, (rows, fpad))
    g = rng.normal(0.0, 1.0, (rows, lpad))
    tiltable = np.array([1, 1, 0, 1, 1], dtype=bool)
    tmax = 4.0
    ana = mlp.backward_exact(p, x, g, fpad, H, lpad, tmax, tiltable)
    f = lambda pp: float(np.sum(g * mlp.forward_exact(pp, x, fpad, H, lpad, tmax, tiltable)))
    eps = 1e-6
    num = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = eps
        num[i] = (f(p + e) - f(p - e)) / (2 * eps)
    assert np.allclose(ana, num, rtol=1e-5, atol=1e-6)


def test_ce_output_gradient_is_derivative_of_log2lr():
    rng = np.random.default_rng(2)
    th = rng.uniform(-2, 2, (3, 4))
    A = rng.integers(0, 3, (3, 4)).astype(np.float64)
    B = rng.uniform(0, 2, (3, 4))
    w = rng.uniform(0.1, 1.0, 3)
    tiltable = np.array([1, 1, 1, 0], dtype=bool)
    # loss = sum_pairs w ln2 log2lr(theta), log2lr = sum_L -theta A + (2^theta - 1) B / ln2
    loss = lambda t: float(np.sum(w[:, None] * math.log(2) * (-t * A + (np.exp2(t) - 1) * B / math.log(2))))
    g = mlp.ce_output_grad(th, A, B, w, tiltable, 6.0)
    eps = 1e-6
    for i in range(3):
        for j in range(3):
            e = np.zeros_like(th)
            e[i, j] = eps
            assert abs((loss(th + e) - loss(th - e)) / (2 * eps) - g[i, j]) < 1e-6


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    r = mlp.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2 ** -6                 # tie to even (up)
    assert r[3] == -2.5
    assert np.isfinite(r[4])


@pytest.mark.parametrize("th", [1.5, -0.7])
def test_tilted_uniform_delay_closed_form(th):
    """Delay-window tilt: l0 with invariant x <= 8, one edge to Goal."""
    U, T = 8, 3.0
    net = cs.ParsedNetwork.from_string(single_uniform(U), "T.Goal")
    q = cs.reach_query(T)
    theta = _loc_tilts(net, {"l0": th})
    n = 100000
    r = oracle.is_run(net, q, seed=12, n_runs=n, theta=theta, threads=0)
    th32 = float(np.float32(th))
    a = 2.0 ** th32
    hits = r["flag"].astype(bool)
    p_prop = 1.0 - (1.0 - T / U) ** a
    assert abs(hits.mean() - p_prop) < 5 * math.sqrt(p_prop * (1 - p_prop) / n), (hits.mean(), p_prop)
    d = r["time"][hits]
    expect = -th32 + (1.0 - a) * np.log2(1.0 - d / U)
    assert np.allclose(r["log2lr"][hits], expect, rtol=1e-12, atol=1e-12)
    w = np.where(hits, np.exp2(r["log2lr"]), 0.0)
    se = w.std() / math.sqrt(n)
    assert abs(w.mean() - T / U) < 4 * se, (w.mean(), T / U, se)
    # the untilted window gives T/U; the tilt moved the proposal away from it
    assert abs(p_prop - T / U) > 0.05


def test_tilted_race8_unbiased_against_plain():
    """Race-8 mixes uniform (L0, L2) and exponential (L1, L3) racers: under
    tilts on all four locations the weighted estimate agrees with plain
    simulation (the cuSMC sampling, LITERAL mode) on a rarer threshold."""
    net = cs.ParsedNetwork.from_string(G.race_network(8)[0], "")
    q = cs.var_query(net.variable_id("count"), 20.0, threshold=48)
    n = 40000
    plain, _ = oracle.run(net, q, seed=13, n_runs=200000, mode=oracle.LITERAL, threads=0)
    theta = _loc_tilts(net, {"L0": 0.8, "L1": 0.6, "L2": 0.8, "L3": 0.6})
    r = oracle.is_run(net, q, seed=14, n_runs=n, theta=theta, early_stop=True, threads=0)
    w = np.where(r["flag"].astype(bool), np.exp2(r["log2lr"]), 0.0)
    se = w.std() / math.sqrt(n)
    p = plain.estimate
    sp = math.sqrt(max(p * (1 - p), 1e-12) / 200000)
    assert 1e-3 < p < 0.2, p
    assert abs(w.mean() - p) < 4 * math.sqrt(se ** 2 + sp ** 2), (w.mean(), p, se, sp)
    assert r["flag"].mean() > 3 * p  # the tilt made the event common
