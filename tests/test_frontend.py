"""UPPAAL XML front end: what a model file means (uppaal_xml_parser.cpp)."""
from __future__ import annotations

import math
import os
from pathlib import Path

import numpy as np
import pytest

import cusmc_amd as cs
from cusmc_amd import _capi as A
from cusmc_amd import generators as G

import oracle


# (file, components, variables, locations, edges)
CORPUS = [
    ("RaceTest.xml", 9, 1, 18, 9),
    ("UppaalWeightTest.xml", 1, 1, 5, 7),
    ("condTest.xml", 1, 2, 3, 2),
    ("fischermulti.xml", 6, 7, 24, 30),
    ("bluetoothSmacc.xml", 3, 18, 21, 44),
    ("agentBaseCovid_32_1.xml", 32, 2, 192, 192),
    ("UppaalBehaviorTest2.xml", 2, 6, 10, 10),
    ("fischer_10_29.xml", 10, 22, 40, 50),
    ("aloha_25.xml", 25, 3, 125, 200),
]


@pytest.mark.parametrize("f,nc,nv,nl,ne", CORPUS)
def test_corpus_models_parse(models_dir, f, nc, nv, nl, ne):
    net = cs.ParsedNetwork.from_file(models_dir / f, "")
    n = net.net
    assert (n.n_components, n.n_vars, n.n_locations, n.n_edges) == (nc, nv, nl, ne)


@pytest.mark.parametrize("f,nc,nv,nl,ne", [("CSMA_100.xml", 100, 900, 1800, 5500), ("aloha_50.xml", 50, 3, 250, 400),
                                            ("fischer_100_29.xml", 100, 202, 400, 500),
                                            ("agentBaseCovid_350.xml", 350, 2, 2100, 2100),
                                            ("agentBaseCovid_100_1.0.xml", 100, 2, 600, 600)])
def test_large_reference_models_parse(f, nc, nv, nl, ne):
    """The reference's large-model corpus (fixtures: xml_files/LargeModels and
    the agent-based covid models, gzipped)."""
    from _models import large_model
    net = cs.ParsedNetwork.from_string(large_model(f), "")
    n = net.net
    assert (n.n_components, n.n_vars, n.n_locations, n.n_edges) == (nc, nv, nl, ne)


def test_behavior_test3_parses(models_dir):
    net = cs.ParsedNetwork.from_file(models_dir / "UppaalBehaviorTest3.xml", "c2.g2")
    assert net.n_components == 2 and net.goal_count == 1


def test_variable_ids_follow_declaration_order(models_dir):
    net = cs.ParsedNetwork.from_file(models_dir / "fischer_10_29.xml", "")
    # globals first (id, in_critical), then per template: x, pid
    names = [net.variable_name(v) for v in range(net.n_vars)]
    assert names[:4] == ["id", "in_critical", "x", "pid"]
    kinds = net.var_kinds()
    assert kinds[2] == A.VAR_CLOCK and kinds[0] == A.VAR_INT and kinds[3] == A.VAR_INT
    assert net.net.var_init[3] == 1.0  # pid = 1 of P0


def test_racetest_channels_and_component_order(models_dir):
    net = cs.ParsedNetwork.from_file(models_dir / "RaceTest.xml", "")
    n = net.net
    names = [net.component_name(c) for c in range(n.n_components)]
    assert names == [f"Template{i}" for i in (7, 6, 5, 4, 3, 2, 1, 0, 8)]
    chans = [n.edges[e].channel for e in range(n.n_edges)]
    assert chans[3] > 0 and all(c == -chans[3] for i, c in enumerate(chans) if i != 3)


def test_conditional_update_and_clock_guard(models_dir):
    net = cs.ParsedNetwork.from_file(models_dir / "condTest.xml", "Template.l3")
    n = net.net
    # guard x < 2 on l1 -> l2 is a clock conjunct
    g = [n.constraints[k] for k in range(n.n_constraints)]
    assert any(c.uses_variable and c.var == 1 and c.op == 1 for c in g)
    # i := x<2 ? 5 : 0 becomes a COND node
    u = n.updates[0]
    assert n.exprs[u.expr].op == 20
    for x, want in [(0.5, 5.0), (3.0, 0.0), (1.99999, 5.0)]:
        assert oracle.eval_expr(net.net_ptr, u.expr, [0.0, x]) == want


def test_constants_go_through_float_and_six_decimals(models_dir):
    net = cs.ParsedNetwork.from_file(models_dir / "agentBaseCovid_32_1.xml", "")
    n = net.net
    # rate of E is `alpha` = 1.0/5.1 -> float -> "0.196078" -> 0.196078
    lits = {round(n.exprs[i].value, 9) for i in range(n.n_exprs) if n.exprs[i].op == 0}
    assert 0.196078 in lits
    assert 1.0 / 5.1 not in lits


def _doc(decl: str, body: str, tdecl: str = "") -> str:
    return (f'<nta><declaration>{decl}</declaration><template><name>T</name>'
            f'<declaration>{tdecl}</declaration>{body}</template><system>system T;</system></nta>')


def test_initial_values_truncate_to_int():
    xml = _doc("double d = 1.7; int k = -2; clock c;",
               '<location id="id0"><name>a</name></location><init ref="id0"/>')
    net = cs.ParsedNetwork.from_string(xml, "")
    assert [net.net.var_init[i] for i in range(3)] == [1.0, -2.0, 0.0]


def test_only_first_location_label_is_read():
    body = ('<location id="id0"><name>a</name><label kind="invariant">x &lt;= 3</label>'
            '<label kind="exponentialrate">5</label></location><init ref="id0"/>')
    net = cs.ParsedNetwork.from_string(_doc("", body, "clock x;"), "")
    L = net.net.locations[0]
    assert L.inv_count == 1 and L.rate == -1


def test_parentheses_are_stripped_from_guards():
    body = ('<location id="id0"><name>a</name></location><location id="id1"><name>b</name></location>'
            '<init ref="id0"/><transition><source ref="id0"/><target ref="id1"/>'
            '<label kind="guard">v &gt;= (1 + 2) * 3</label></transition>')
    net = cs.ParsedNetwork.from_string(_doc("int v = 0;", body), "")
    g = net.net.constraints[0]
    # reference reading: v >= 1 + 2 * 3 = 7, not 9
    assert oracle.eval_expr(net.net_ptr, g.rhs, [0.0]) == 7.0


def test_goal_marking_exact_and_named_goal():
    xml, _ = G.fischer_network(4)
    net = cs.ParsedNetwork.from_string(xml, "P2.cs")
    n = net.net
    goals = [l for l in range(n.n_locations) if n.locations[l].type == 1]
    assert len(goals) == 1 and net.location_name(goals[0]) == "cs"
    c = next(c for c in range(n.n_components) if n.comp_first_loc[c] <= goals[0] < n.comp_first_loc[c + 1])
    assert net.component_name(c) == "P2"
    # unmatched process name: first location with that name
    net2 = cs.ParsedNetwork.from_string(xml, "Nope.cs")
    assert net2.goal_count == 1


def test_parse_errors_are_reported():
    with pytest.raises(cs.SmcError) as e:
        cs.ParsedNetwork.from_file("/nonexistent/model.xml", "")
    assert e.value.code == A.SMC_E_PARSE
    body = ('<location id="id0"><name>a</name></location><location id="id1"><name>b</name></location>'
            '<init ref="id0"/><transition><source ref="id0"/><target ref="id1"/>'
            '<label kind="assignment">nope = 1</label></transition>')
    with pytest.raises(cs.SmcError) as e:
        cs.ParsedNetwork.from_string(_doc("int v;", body), "")
    assert "nope" in str(e.value)
    with pytest.raises(cs.SmcError):
        cs.ParsedNetwork.from_string("<nta><template>", "")


def _random_expr(rng, depth):
    if depth == 0 or rng.random() < 0.25:
        return rng.choice(["a", "b", "c", str(rng.integers(0, 9)), f"{rng.uniform(0, 5):.3f}"])
    op = rng.choice(["+", "-", "*", "/", "%", "^", "neg", "sqrt"])
    if op == "neg":
        return f"-({_random_expr(rng, depth - 1)})"
    if op == "sqrt":
        return f"sqrt({_random_expr(rng, depth - 1)})"
    return f"({_random_expr(rng, depth - 1)}) {op} ({_random_expr(rng, depth - 1)})"


def _py_eval(s, env):
    """Independent reading of the expression grammar: Python arithmetic with
    cuSMC's operator meanings (% on truncated ints with +1e-5 on the divisor,
    ^ as power, sqrt)."""
    import re
    def mod(x, y):
        ia, ib = int(x), int(y + 1e-5)
        if ib == 0:
            return 0.0
        r = abs(ia) % abs(ib)
        return float(r if ia >= 0 else -r)
    def pw(x, y):
        return float(x) ** float(y) if x > 0 or float(y).is_integer() else float("nan")
    def div(x, y):
        return x / y if y != 0 else (math.copysign(math.inf, x) * math.copysign(1, y) if x != 0 else math.nan)
    py = s.replace("^", "**")
    # precedence of the reference grammar: ^ binds tighter than * / %, parentheses are explicit here
    tree = compile(py, "<e>", "eval")
    return eval(tree, {"sqrt": lambda x: math.sqrt(x) if x >= 0 else math.nan}, dict(env))


def test_random_expressions_parse_and_evaluate():
    """Rates written as random arithmetic: the parsed tree evaluated by the
    oracle matches an independent Python evaluation."""
    rng = np.random.default_rng(3)
    env = {"a": 1.5, "b": 2.0, "c": 7.0}
    checked = 0
    for _ in range(300):
        e = _random_expr(rng, 3)
        if "%" in e or "^" in e:
            continue
        body = (f'<location id="id0"><name>a0</name><label kind="exponentialrate">{e}</label></location>'
                '<init ref="id0"/>')
        net = cs.ParsedNetwork.from_string(_doc("double a = 1.5; double b = 2; double c = 7;", body), "")
        # the front end truncates initial values to int (a -> 1.0); evaluate with explicit values
        r = net.net.locations[0].rate
        got = oracle.eval_expr(net.net_ptr, r, [env["a"], env["b"], env["c"]])
        try:
            want = _py_eval(e, env)
        except (ZeroDivisionError, ValueError, OverflowError):
            continue
        if isinstance(want, complex) or math.isnan(want) or math.isinf(want):
            continue
        assert got == pytest.approx(want, rel=1e-12, abs=1e-12), (e, got, want)
        checked += 1
    assert checked > 100
