"""The oracle's expression evaluator over cuSMC's whole operator set
(expressions.cu:10-95, Q21 tolerance semantics), pinned by brute force
against an independent Python restatement, and the simulator's own FP64 pow
against math.pow.

The trees are built directly as network arrays (tests/_models.RawNetwork):
the XML grammar cannot produce `||`, `==` or `!=` outside a `?:` condition,
but the evaluator (and the GPU bytecode that must match it bit for bit)
implements all of them.
"""
from __future__ import annotations

import math
import struct

import numpy as np

import oracle

from _models import BINARY, E_COND, E_POW, RawNetwork, UNARY, py_eval, random_tree


def _same(a: float, b: float) -> bool:
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return struct.pack("<d", a) == struct.pack("<d", b)


def _values(rng, n):
    pool = [0.0, 1.0, -1.0, 1e-6, -1e-6, 9e-6, 1.1e-5, 2.0, 3.0, 0.5, -2.5, 7.0]
    return [float(rng.choice(pool)) if rng.random() < 0.6 else float(rng.uniform(-5, 5)) for _ in range(n)]


def test_oracle_evaluator_matches_python_semantics_on_every_operator():
    rng = np.random.default_rng(2024)
    seen = set()
    checked = 0
    for k in range(400):
        net = RawNetwork()
        for _ in range(4):
            net.var(0.0)
        e = random_tree(rng, net, 4, 4)
        net.build()
        seen |= {x[0] for x in net.exprs}
        for _ in range(4):
            vals = _values(rng, 4)
            got = oracle.eval_expr(net.net_ptr, e, vals)
            want = py_eval(net.exprs, e, vals, pow_fn=oracle.pow)
            assert _same(got, want), (k, vals, got, want, net.exprs)
            checked += 1
    assert checked == 1600
    assert set(BINARY) | set(UNARY) | {E_COND} <= seen


def test_tolerance_semantics_of_logic_and_equality():
    """Q21: `==` holds within 1e-5, `!=` beyond it; `&&`, `||`, `!` read a
    value as true when |v| > 1e-5; `%` adds 1e-5 to the divisor before the
    int cast; `?:` tests |c| > 1e-5."""
    net = RawNetwork()
    x, y = net.var(0.0), net.var(0.0)
    X, Y = net.ref(x), net.ref(y)
    eq, ne = net.expr(17, X, Y), net.expr(18, X, Y)
    land, lor, lnot = net.expr(11, X, Y), net.expr(12, X, Y), net.expr(19, X)
    mod = net.expr(10, X, Y)
    cond = net.expr(20, X, net.lit(5.0), net.lit(0.0))
    net.build()
    ev = lambda e, a, b: oracle.eval_expr(net.net_ptr, e, [a, b])
    assert ev(eq, 1.0, 1.0 + 9e-6) == 1.0 and ev(eq, 1.0, 1.0 + 1.1e-5) == 0.0
    assert ev(ne, 1.0, 1.0 + 9e-6) == 0.0 and ev(ne, 1.0, 1.0 + 1.1e-5) == 1.0
    assert ev(land, 2e-5, -3.0) == 1.0 and ev(land, 9e-6, 3.0) == 0.0
    assert ev(lor, 9e-6, -9e-6) == 0.0 and ev(lor, 9e-6, -2e-5) == 1.0
    assert ev(lnot, 9e-6, 0) == 1.0 and ev(lnot, 2e-5, 0) == 0.0
    assert ev(mod, 7.9, 2.99999) == 1.0          # int(7.9) % int(2.99999 + 1e-5) = 7 % 3
    assert ev(mod, -7.0, 3.0) == -1.0 and ev(mod, 7.0, 0.0) == 0.0
    assert ev(cond, 2e-5, 0) == 5.0 and ev(cond, 9e-6, 0) == 0.0


def test_simulator_pow_against_math_pow():
    """smc_pow: exact square-and-multiply for integer exponents up to 1024,
    exp2(b log2 a) otherwise; within a few ulps of libm's pow."""
    rng = np.random.default_rng(5)
    for _ in range(4000):
        a = float(rng.uniform(0.01, 50.0))
        b = float(rng.choice([rng.uniform(-8, 8), float(rng.integers(-12, 13))]))
        got, want = oracle.pow(a, b), math.pow(a, b)
        assert abs(got - want) <= 4e-15 * abs(want) * max(1.0, abs(b)), (a, b, got, want)
    for a in (-2.0, -0.5, -3.0):
        for b in (2.0, 3.0, -1.0, -4.0):
            assert oracle.pow(a, b) == math.pow(a, b)
        assert math.isnan(oracle.pow(a, 0.5))
    assert oracle.pow(0.0, 2.0) == 0.0 and oracle.pow(0.0, -1.0) == math.inf
    assert oracle.pow(0.0, 0.5) == 0.0 and oracle.pow(5.0, 0.0) == 1.0
    assert E_POW == 7
