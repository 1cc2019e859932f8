"""Pins of the CPU oracle against what the mathematics fixes.

Each test compares an oracle estimate with a closed form, an exact CTMC
probability or an exact frequency, within 5 standard errors (fixed seeds,
so the outcome is deterministic).  Both ways of sampling the race are
checked: LITERAL (cuSMC's per-component resampling, simulation.cu:409-541)
and CANONICAL (the GPU kernel's sampling of the same law).  A dropped ln 2,
a transposed weight, a wrong tie-break or a receiver order error moves one
of these answers by far more than the tolerance.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cusmc_amd as cs
from cusmc_amd import generators as G

import oracle

from _models import (clock_rate_race, first_winner, guarded_uniform, name_locations, receiver_goal, receiver_goal_p, single_exponential,
                     single_uniform, sir_exact, urgent_order, zeno)

MODES = [oracle.CANONICAL, oracle.LITERAL]
MODE_IDS = ["canonical", "literal"]


def _prob_ok(est, p, n, k=5.0):
    return abs(est - p) <= k * math.sqrt(max(p * (1 - p), 1e-12) / n) + 1e-12


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("rate,T", [(0.7, 2.0), (2.5, 0.4), (1.0, 1.0)])
def test_exponential_delay_closed_form(mode, rate, T):
    """Q1: delays are -log2(u)/rate, so Pr[<=T] = 1 - 2^(-rate T)."""
    net = cs.ParsedNetwork.from_string(single_exponential(rate), "")
    n = 60000
    res, _ = oracle.run(net, cs.reach_query(T), seed=3, n_runs=n, mode=mode, threads=8)
    p = 1 - 2 ** (-rate * T)
    assert _prob_ok(res.estimate, p, n), (res.estimate, p)
    # the natural-log reading would give 1 - e^(-rate T): must be excluded
    assert not _prob_ok(res.estimate, 1 - math.exp(-rate * T), n, k=3.0)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("U,T", [(8, 3.0), (5, 4.5)])
def test_uniform_delay_closed_form(mode, U, T):
    """Q2: a clock bound x <= U makes the delay uniform on (0, U - x]."""
    net = cs.ParsedNetwork.from_string(single_uniform(U), "")
    n = 60000
    res, _ = oracle.run(net, cs.reach_query(T), seed=4, n_runs=n, mode=mode, threads=8)
    assert _prob_ok(res.estimate, T / U, n), (res.estimate, T / U)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_race_winner_frequencies(mode):
    """Q4: the winner of an exponential race is component i w.p. r_i / sum r."""
    rates = [0.5, 1.0, 2.0, 0.25, 3.0]
    net = cs.ParsedNetwork.from_string(first_winner(rates), "")
    n = 60000
    _, ro = oracle.run(net, cs.var_query(0, 50.0, is_max=True), seed=5, n_runs=n, mode=mode,
                       threads=8, per_run=("value",))
    counts = np.bincount(ro["value"].astype(int), minlength=len(rates) + 1)[1:]
    p = np.array(rates) / sum(rates)
    assert counts.sum() == n
    for c, pi in zip(counts, p):
        assert _prob_ok(c / n, pi, n), (counts / n, p)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_zero_delay_tie_goes_to_lowest_index(mode):
    """Q3/Q4: urgent components race with delay 0; ties go to the lowest index."""
    net = cs.ParsedNetwork.from_string(urgent_order(), "")
    q = cs.var_query(0, 5.0)
    for run in range(200):
        ev, tm = oracle.trace(net, q, seed=6, run=run, mode=mode)
        assert [tuple(e[:2]) for e in ev[:2]] == [(0, 1), (1, 2)]
        assert tm[0] == 0.0 and tm[1] == 0.0


def test_zeno_runs_are_cut_off():
    """A cycle of zero delays never reaches the time bound (cuSMC would spin
    forever); the run ends after max_steps race rounds and is counted."""
    net = cs.ParsedNetwork.from_string(zeno(), "Z.never")
    for mode in MODES:
        res, ro = oracle.run(net, cs.reach_query(10.0, max_steps=5000), seed=1, n_runs=20, mode=mode,
                             per_run=("steps", "time", "flag"))
        assert (ro["steps"] == 5000).all() and (ro["time"] == 0.0).all() and not ro["flag"].any()


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("T", [3.0, 10.0])
def test_condtest_closed_form(mode, T, models_dir):
    """condTest.xml: `i := x<2 ? 5 : 0` and the null step that sticks when the
    first delay exceeds 2: Pr[<=T](<> l3) = 0.75 - 2 ln2 2^(-T) for T >= 2."""
    net = cs.ParsedNetwork.from_file(models_dir / "condTest.xml", "Template.l3")
    n = 80000
    res, _ = oracle.run(net, cs.reach_query(T), seed=7, n_runs=n, mode=mode, threads=8)
    p = 0.75 - 2 * math.log(2) * 2 ** (-T)
    assert _prob_ok(res.estimate, p, n), (res.estimate, p)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("T", [1.0, 2.0, 10.0])
def test_racetest_broadcast_order(mode, T, models_dir):
    """RaceTest.xml: Template4 broadcasts `test`; the eight receivers fire in
    component order (Template7..0 then Template8), so var ends at 8 once the
    sender has fired before T; a sender firing in the step that crosses T
    still counts (var = 4) but its receivers do not run.
    E[max var] = 8 (1 - 2^-T) + (4/9) 2^-T."""
    net = cs.ParsedNetwork.from_file(models_dir / "RaceTest.xml", "")
    n = 60000
    res, ro = oracle.run(net, cs.var_query(0, T, is_max=True, estimate=True, threshold=-1), seed=8,
                         n_runs=n, mode=mode, threads=8, per_run=("value",))
    mean = 8 * (1 - 2 ** -T) + (4 / 9) * 2 ** -T
    sd = ro["value"].std()
    assert abs(res.estimate - mean) <= 5 * sd / math.sqrt(n) + 1e-9, (res.estimate, mean)
    assert set(np.unique(ro["value"])) <= {0.0, 4.0, 8.0}


def test_weighted_branch_frequencies(models_dir):
    """UppaalWeightTest.xml: a branchpoint with weights 5:10:1 (Q10)."""
    net = cs.ParsedNetwork.from_file(models_dir / "UppaalWeightTest.xml", "")
    q = cs.reach_query(12.0)
    # the branch edges are the three edges leaving the branchpoint location
    n_ = net.net
    branch = [l for l in range(n_.n_locations) if n_.locations[l].type == 2][0]
    L = n_.locations[branch]
    bedges = {L.edge_first + k: n_.edges[L.edge_first + k] for k in range(L.edge_count)}
    weight_of = {e: n_.exprs[E.weight].value for e, E in bedges.items()}
    counts = {e: 0 for e in bedges}
    for run in range(4000):
        ev, _ = oracle.trace(net, q, seed=9, run=run)
        for step, comp, e, _ in ev:
            if e in counts:
                counts[e] += 1
    tot = sum(counts.values())
    W = sum(weight_of.values())
    assert sorted(weight_of.values()) == [1.0, 5.0, 10.0]
    for e, c in counts.items():
        assert _prob_ok(c / tot, weight_of[e] / W, tot), (counts, weight_of)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("n,beta,gamma,k,T", [(12, 0.9, 0.3, 5, 4.0), (20, 0.6, 0.2, 8, 6.0)])
def test_sir_exact_ctmc(mode, n, beta, gamma, k, T):
    """Agent SIR with broadcast infection waves vs the exact lumped CTMC."""
    xml, _ = G.sir_network(n, beta, gamma, initial_infected=1, seed=1)
    net = cs.ParsedNetwork.from_string(xml, "")
    q = cs.var_query(net.variable_id("I"), T, is_max=True, threshold=k - 1)
    runs = 40000
    res, _ = oracle.run(net, q, seed=10, n_runs=runs, mode=mode, threads=8)
    p = sir_exact(n, beta, gamma, 1, k, T)
    assert 0.02 < p < 0.98
    assert _prob_ok(res.estimate, p, runs), (res.estimate, p)


def test_literal_and_canonical_agree_in_law():
    """The two race samplings give the same estimates on larger networks."""
    cases = []
    xml, _ = G.race_network(8)
    net = cs.ParsedNetwork.from_string(xml, "")
    cases.append((net, cs.var_query(net.variable_id("count"), 20.0, threshold=9)))
    xml, _ = G.fischer_network(8)
    net2 = cs.ParsedNetwork.from_string(xml, "P0.cs")
    cases.append((net2, cs.reach_query(30.0)))
    xml, _ = G.sir_network(40, 0.5, 0.15, initial_infected=2, seed=3)
    net3 = cs.ParsedNetwork.from_string(xml, "")
    cases.append((net3, cs.var_query(net3.variable_id("I"), 15.0, threshold=11)))
    n = 20000
    for net_, q in cases:
        a, _ = oracle.run(net_, q, seed=11, n_runs=n, mode=oracle.CANONICAL, threads=8)
        b, _ = oracle.run(net_, q, seed=12, n_runs=n, mode=oracle.LITERAL, threads=8)
        p = (a.estimate + b.estimate) / 2
        se = math.sqrt(max(p * (1 - p), 1e-6) * 2 / n)
        assert abs(a.estimate - b.estimate) <= 5 * se, (a.estimate, b.estimate)


def test_tree_pick_proportional_and_positive():
    rng = np.random.default_rng(0)
    for size in [1, 2, 3, 7, 8, 33, 100]:
        rates = rng.uniform(0, 2, size) * (rng.uniform(size=size) < 0.7)
        if rates.sum() == 0:
            rates[0] = 1.0
        total, _ = oracle.tree_pick(rates, 0.5)
        assert total == pytest.approx(rates.sum(), rel=1e-12)
        hits = np.zeros(size)
        vs = (np.arange(20000) + 0.5) / 20000
        for v in vs:
            _, c = oracle.tree_pick(rates, float(v))
            assert rates[c] > 0
            hits[c] += 1
        assert np.allclose(hits / len(vs), rates / rates.sum(), atol=2e-3)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("live", [None, 0.2], ids=["absorbing", "live"])
@pytest.mark.parametrize("T", [0.7, 2.0, 5.0])
def test_receiver_goal_closed_form(mode, live, T):
    """Q16: a goal entered by a broadcast receiver counts only if the step
    after the receivers, raced by every component, keeps t < T
    (simulation.cu:409-415 fire receivers inside the iteration, :835-849 test
    the goal at the next loop top).  Without a live component the state is
    absorbing once the receiver has fired; with one, the components that can
    no longer fire still race in that step."""
    net = cs.ParsedNetwork.from_string(receiver_goal(live), "")
    n = 40000
    res, _ = oracle.run(net, cs.reach_query(T), seed=13, n_runs=n, mode=mode, threads=8)
    p = receiver_goal_p(T, live)
    assert _prob_ok(res.estimate, p, n), (res.estimate, p)
    # reading it as a winner-entered goal (hit iff the sender fires before T)
    # is excluded, and so is dropping the receiver's hit
    assert not _prob_ok(res.estimate, 1 - 2.0 ** -T, n, k=3.0)
    assert res.estimate > 0.05


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("T", [1.0, 3.0])
def test_racetest_location_queries(mode, T, models_dir):
    """RaceTest.xml with location queries: the sender's target (entered by
    the race winner) is reached w.p. 1 - 2^-T; a receiver's target (entered
    in the next iteration's broadcast) needs one more step of the nine
    components at the default rate 1: 1 - 2^-T - (2^-T - 2^-9T)/8."""
    xml = (models_dir / "RaceTest.xml").read_text()
    xml = name_locations(xml, {"id10": "sent", "id1": "got"})
    n = 40000
    for query, p in (("Process4.sent", 1 - 2.0 ** -T),
                     ("Process7.got", 1 - 2.0 ** -T - (2.0 ** -T - 2.0 ** (-9 * T)) / 8)):
        net = cs.ParsedNetwork.from_string(xml, query)
        res, _ = oracle.run(net, cs.reach_query(T), seed=14, n_runs=n, mode=mode, threads=8)
        assert _prob_ok(res.estimate, p, n), (query, res.estimate, p)


def test_clock_rate_races_every_component():
    """A rate that reads a clock changes at every null step of cuSMC's race,
    so the components that cannot fire still race: both samplings agree,
    and the rate is not frozen at its first value (that would give
    Pr = 1 - 2^(-0.1 T))."""
    net = cs.ParsedNetwork.from_string(clock_rate_race(), "")
    n = 30000
    T = 4.0
    a, _ = oracle.run(net, cs.reach_query(T), seed=15, n_runs=n, mode=oracle.CANONICAL, threads=8)
    b, _ = oracle.run(net, cs.reach_query(T), seed=16, n_runs=n, mode=oracle.LITERAL, threads=8)
    p = (a.estimate + b.estimate) / 2
    se = math.sqrt(max(p * (1 - p), 1e-6) * 2 / n)
    assert abs(a.estimate - b.estimate) <= 5 * se, (a.estimate, b.estimate)
    frozen = 1 - 2 ** (-0.1 * T)
    assert a.estimate > frozen + 0.1, (a.estimate, frozen)


def _agree(net, q, n, seed=21):
    a, _ = oracle.run(net, q, seed=seed, n_runs=n, mode=oracle.CANONICAL, threads=8)
    b, _ = oracle.run(net, q, seed=seed + 1, n_runs=n, mode=oracle.LITERAL, threads=8)
    p = (a.estimate + b.estimate) / 2
    se = math.sqrt(max(p * (1 - p), 1e-6) * 2 / n)
    assert abs(a.estimate - b.estimate) <= 5 * se, (a.estimate, b.estimate)
    return a.estimate


def test_location_queries_on_broadcast_models_agree_in_law(models_dir):
    """Canonical sampling (absorption, left-out null racers, the crossing
    correction) against cuSMC's per-component sampling, on location queries
    of models with broadcasts: SIR agents reaching I as winner or receiver,
    the receiver-goal model beside a clock-rate component, UppaalBehavior-
    Test2 and fischermulti."""
    n = 20000
    xml, _ = G.sir_network(24, 0.6, 0.2, initial_infected=2, seed=5)
    for query in ("A3.I", "A3.R"):
        net = cs.ParsedNetwork.from_string(xml, query)
        p = _agree(net, cs.reach_query(8.0), n)
        assert 0.02 < p < 0.98, (query, p)
    net = cs.ParsedNetwork.from_string(receiver_goal(0.3, clock_rate=True), "")
    _agree(net, cs.reach_query(2.0), n)
    # (UppaalBehaviorTest2's c1.f2 is left out: its runs are Zeno under
    # these semantics, a uniform delay bounded by a shrinking invariant gap)
    for fname, query, T in (("UppaalBehaviorTest2.xml", "c2.g2", 100.0),
                            ("fischermulti.xml", "P1.cs", 20.0),
                            ("UppaalBehaviorTest3.xml", "c2.g2", 20.0), ("UppaalBehaviorTest3.xml", "c1.f2", 5.0)):
        net = cs.ParsedNetwork.from_file(models_dir / fname, query)
        _agree(net, cs.reach_query(T), n)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("n_dead", [0, 3])
@pytest.mark.parametrize("T", [3.0, 4.5])
def test_redrawn_uniform_delay_closed_form(mode, n_dead, T):
    """A uniform delay redrawn at every step (simulation.cu:461-521) keeps
    the hazard 1/(W - tau) of one draw over the whole window: with the guard
    x >= 2 and the window x <= 5, Pr = (T - 2)/3, with or without null
    racers beside it.  The canonical sampling leaves the null racers out."""
    net = cs.ParsedNetwork.from_string(guarded_uniform(n_dead), "")
    n = 60000
    res, ro = oracle.run(net, cs.reach_query(T), seed=16, n_runs=n, mode=mode, threads=8, per_run=("steps",))
    p = (T - 2.0) / 3.0
    assert _prob_ok(res.estimate, p, n), (res.estimate, p)
    if mode == oracle.LITERAL and n_dead:
        assert ro["steps"].mean() > 3.0  # the null racers' steps are there in cuSMC's sampling
