"""The shared random stream (include/smc_stream.h), compiled for the host.

Known-answer vectors pin Philox4x32-10; the conversions and the log2 used
for exponential delays are pinned against exact values and the math library.
The device build of the same header is pinned by the bit-exact GPU parity
tests (every delay and every edge choice enters the trajectory digests).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle

# Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key, output)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert oracle.philox(ctr, key) == out


def test_philox_counter_sensitivity():
    base = oracle.philox((1, 2, 3, 4), (5, 6))
    for i in range(4):
        c = [1, 2, 3, 4]
        c[i] ^= 1
        o = oracle.philox(tuple(c), (5, 6))
        flipped = sum(bin(a ^ b).count("1") for a, b in zip(base, o))
        assert 30 < flipped < 100  # avalanche: about half of 128 bits


def test_u01_range_and_extremes():
    L = oracle.lib()
    assert L.oracle_u01(0, 0) == 2.0 ** -53
    assert L.oracle_u01(0xFFFFFFFF, 0xFFFFFFFF) == 1.0
    assert L.oracle_u01(0x80000000, 0) == 0.5 + 2.0 ** -53
    rng = np.random.default_rng(1)
    for hi, lo in rng.integers(0, 2 ** 32, size=(2000, 2), dtype=np.uint64):
        u = L.oracle_u01(int(hi), int(lo))
        k = (int(hi) << 21) | (int(lo) >> 11)
        assert u == (k + 1) * 2.0 ** -53
        assert 0.0 < u <= 1.0


def test_shared_draw_words():
    """Uniform delays: components 2g, 2g+1 take words 0-1 and 2-3 of draw g;
    receiver choices: components 4g..4g+3 take word c mod 4 of draw g, each
    word a 32-bit uniform (k + 1) 2^-32 on (0, 1]."""
    import ctypes as C
    L = oracle.lib()
    rng = np.random.default_rng(7)
    for _ in range(200):
        w = [int(x) for x in rng.integers(0, 2 ** 32, size=4, dtype=np.uint64)]
        arr = (C.c_uint32 * 4)(*w)
        for c in range(12):
            assert L.oracle_shared_draw(c, 0) == c // 2
            assert L.oracle_shared_draw(c, 1) == c // 4
            hi, lo = w[2 * (c % 2)], w[2 * (c % 2) + 1]
            assert L.oracle_shared_u(arr, c, 0) == L.oracle_u01(hi, lo)
            assert L.oracle_shared_u(arr, c, 1) == (w[c % 4] + 1) * 2.0 ** -32
    ext = (C.c_uint32 * 4)(0, 0xFFFFFFFF, 5, 6)
    assert L.oracle_shared_u(ext, 0, 1) == 2.0 ** -32
    assert L.oracle_shared_u(ext, 1, 1) == 1.0


def test_log2_exact_at_powers_of_two():
    L = oracle.lib()
    for k in range(-1074 + 52, 1000, 37):
        assert L.oracle_log2(2.0 ** k) == float(k)


def test_log2_accuracy():
    L = oracle.lib()
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(2 ** -53, 1.0, 5000), rng.uniform(0.5, 2.0, 2000),
                         np.exp(rng.uniform(-700, 700, 2000))])
    for x in xs:
        got = L.oracle_log2(float(x))
        ref = math.log2(float(x))
        assert abs(got - ref) <= 4 * math.ulp(ref) + 1e-15, (x, got, ref)


def test_pow_integer_and_general():
    L = oracle.lib()
    for a in [0.5, 1.5, 2.0, -3.0, 7.0]:
        for b in range(-6, 9):
            assert L.oracle_pow(a, float(b)) == pytest.approx(a ** b, rel=1e-15)
    for a, b in [(2.0, 0.5), (10.0, 1.7), (0.3, 2.25), (7.5, -1.3)]:
        assert L.oracle_pow(a, b) == pytest.approx(a ** b, rel=1e-13)
    assert L.oracle_pow(0.0, 2.0) == 0.0
    assert math.isnan(L.oracle_pow(-2.0, 0.5))


def test_exponential_sample_law():
    """-log2(u) for u = u01(philox) is exponential with rate ln 2."""
    L = oracle.lib()
    n = 40000
    d = np.empty(n)
    for i in range(n):
        w = oracle.philox((i, 0, 7, 0), (123, 0))
        d[i] = -L.oracle_log2(L.oracle_u01(w[0], w[1]))
    # mean 1/ln2, and P(d > 1) = 1/2
    assert abs(d.mean() - 1 / math.log(2)) < 4 * (1 / math.log(2)) / math.sqrt(n)
    assert abs((d > 1.0).mean() - 0.5) < 4 * 0.5 / math.sqrt(n)
