"""CPU oracle of the simulation semantics — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, and only as the checker or the timed CPU baseline.
The product (cusmc_amd, libcusmc.so) never calls it.

Parity status of each function (see the header of oracle.cpp for the
reference file:line each step follows):
  expression evaluation   pinned: brute force vs numpy/float semantics and
                          the bytecode interpreter on random trees
  exponential delay (Q1)  pinned: closed form 1 - 2^(-lambda T)
  uniform delay (Q2)      pinned: closed form T/U
  race, tie-break (Q4)    pinned: literal vs canonical sampling agree in law;
                          winner frequencies lambda_i / sum lambda
  broadcast order (Q9)    pinned: RaceTest.xml closed form
  weighted pick (Q10)     pinned: UppaalWeightTest.xml 5:10:1
  ?: update, null step    pinned: condTest.xml closed form
  SIR race + broadcast    pinned: exact lumped CTMC (tests/golden)
  tilted race, log2 LR    pinned: theta = 0 reduces to the plain race bit for
                          bit; E[hit 2^log2lr] equals the exact CTMC value
                          under a tilt; log2 LR of a one-component race has
                          the closed form -theta + (2^theta - 1) r d
  macro-step statistics   pinned: their sum over a run reproduces the run's
                          log2 LR derivative by finite differences of theta
  MLP (mlp.py)            pinned: central finite differences of the fp64
                          forward pass
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"

CANONICAL, LITERAL = 0, 1


class OracleOutputs(C.Structure):
    _fields_ = [("flag", C.POINTER(C.c_uint8)), ("value", C.POINTER(C.c_double)),
                ("steps", C.POINTER(C.c_uint32)), ("transitions", C.POINTER(C.c_uint32)),
                ("digest", C.POINTER(C.c_uint64)), ("time", C.POINTER(C.c_double))]


class OracleResult(C.Structure):
    _fields_ = [("n_runs", C.c_uint64), ("hits", C.c_uint64), ("estimate", C.c_double),
                ("sum", C.c_double), ("sumsq", C.c_double), ("steps", C.c_uint64),
                ("transitions", C.c_uint64), ("seconds", C.c_double), ("threads", C.c_int32),
                ("reserved", C.c_int32)]


class OracleIsOutputs(C.Structure):
    _fields_ = [("flag", C.POINTER(C.c_uint8)), ("log2lr", C.POINTER(C.c_double)),
                ("digest", C.POINTER(C.c_uint64)), ("steps", C.POINTER(C.c_uint32)),
                ("transitions", C.POINTER(C.c_uint32)), ("time", C.POINTER(C.c_double))]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built ({LIB_PATH}); run `python cusmc_amd/build.py oracle`")
        from cusmc_amd import _capi as A
        l = C.CDLL(str(LIB_PATH))
        l.oracle_run.restype = C.c_int
        l.oracle_run.argtypes = [C.POINTER(A.Network), C.POINTER(A.Query), C.c_uint64, C.c_uint64, C.c_uint64,
                                 C.c_int, C.c_int, C.POINTER(OracleOutputs), C.POINTER(OracleResult)]
        l.oracle_trace.restype = C.c_long
        l.oracle_trace.argtypes = [C.POINTER(A.Network), C.POINTER(A.Query), C.c_uint64, C.c_uint64, C.c_int,
                                   C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_long]
        l.oracle_is_run.restype = C.c_int
        l.oracle_is_run.argtypes = [C.POINTER(A.Network), C.POINTER(A.Query), C.c_uint64, C.c_uint64, C.c_uint64,
                                    C.POINTER(C.c_float), C.c_int, C.c_int, C.POINTER(OracleIsOutputs)]
        l.oracle_is_record.restype = C.c_long
        l.oracle_is_record.argtypes = [C.POINTER(A.Network), C.POINTER(A.Query), C.c_uint64, C.c_uint64,
                                       C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float), C.c_int, C.c_int,
                                       C.POINTER(C.c_uint16), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_long]
        l.oracle_cost.restype = C.c_int
        l.oracle_cost.argtypes = [C.POINTER(A.Network), C.POINTER(A.Query), C.c_uint64, C.c_uint64, C.c_uint64,
                                  C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
        l.oracle_tree_pick.restype = C.c_double
        l.oracle_tree_pick.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_double, C.POINTER(C.c_int)]
        l.oracle_last_error.restype = C.c_char_p
        l.oracle_philox.restype = None
        l.oracle_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        l.oracle_u01.restype = C.c_double
        l.oracle_u01.argtypes = [C.c_uint32, C.c_uint32]
        l.oracle_shared_u.restype = C.c_double
        l.oracle_shared_u.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_int]
        l.oracle_shared_draw.restype = C.c_uint32
        l.oracle_shared_draw.argtypes = [C.c_uint32, C.c_int]
        l.oracle_log2.restype = C.c_double
        l.oracle_log2.argtypes = [C.c_double]
        l.oracle_pow.restype = C.c_double
        l.oracle_pow.argtypes = [C.c_double, C.c_double]
        l.oracle_eval_expr.restype = C.c_double
        l.oracle_eval_expr.argtypes = [C.POINTER(A.Network), C.c_int, C.POINTER(C.c_double)]
        _lib = l
    return _lib


def run(network, query, *, seed: int, n_runs: int, run_begin: int = 0, mode: int = CANONICAL,
        threads: int = 1, per_run: tuple[str, ...] = ()):
    """Simulate with the oracle; returns (OracleResult, {name: array})."""
    spec = {"flag": (np.uint8, C.c_uint8), "value": (np.float64, C.c_double),
            "steps": (np.uint32, C.c_uint32), "transitions": (np.uint32, C.c_uint32),
            "digest": (np.uint64, C.c_uint64), "time": (np.float64, C.c_double)}
    outs = OracleOutputs()
    arrays = {}
    for name in per_run:
        dt, ct = spec[name]
        a = np.zeros(n_runs, dtype=dt)
        arrays[name] = a
        setattr(outs, name, a.ctypes.data_as(C.POINTER(ct)))
    res = OracleResult()
    rc = lib().oracle_run(network.net_ptr, C.byref(query), C.c_uint64(seed), C.c_uint64(run_begin),
                          C.c_uint64(n_runs), int(mode), int(threads), C.byref(outs), C.byref(res))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return res, arrays


def trace(network, query, *, seed: int, run: int, mode: int = CANONICAL, cap: int = 100000):
    ev = np.zeros((cap, 4), dtype=np.int32)
    tm = np.zeros(cap, dtype=np.float64)
    n = lib().oracle_trace(network.net_ptr, C.byref(query), C.c_uint64(seed), C.c_uint64(run), int(mode),
                           ev.ctypes.data_as(C.POINTER(C.c_int32)), tm.ctypes.data_as(C.POINTER(C.c_double)),
                           cap)
    if n < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    n = min(n, cap)
    return ev[:n], tm[:n]


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return tuple(o)


def eval_expr(net_ptr, e: int, vars_) -> float:
    v = np.ascontiguousarray(vars_, dtype=np.float64)
    return lib().oracle_eval_expr(net_ptr, int(e), v.ctypes.data_as(C.POINTER(C.c_double)))


def pow(a: float, b: float) -> float:
    """The simulator's own FP64 pow (include/smc_stream.h smc_pow)."""
    return lib().oracle_pow(float(a), float(b))


def tree_pick(rates, v: float):
    r = np.ascontiguousarray(rates, dtype=np.float64)
    ch = C.c_int()
    total = lib().oracle_tree_pick(r.ctypes.data_as(C.POINTER(C.c_double)), len(r), float(v), C.byref(ch))
    return total, ch.value


def _theta_ptr(theta):
    if theta is None:
        return None, None
    th = np.ascontiguousarray(theta, dtype=np.float32)
    return th, th.ctypes.data_as(C.POINTER(C.c_float))


def is_run(network, query, *, seed: int, n_runs: int, run_begin: int = 0, theta=None, early_stop: bool = False,
           threads: int = 0):
    """Tilted runs (theta: log2 tilt per location, or None); returns a dict
    of per-run arrays flag, log2lr, digest, steps, transitions, time."""
    th, thp = _theta_ptr(theta)
    spec = {"flag": (np.uint8, C.c_uint8), "log2lr": (np.float64, C.c_double),
            "digest": (np.uint64, C.c_uint64), "steps": (np.uint32, C.c_uint32),
            "transitions": (np.uint32, C.c_uint32), "time": (np.float64, C.c_double)}
    outs = OracleIsOutputs()
    arrays = {}
    for name, (dt, ct) in spec.items():
        a = np.zeros(n_runs, dtype=dt)
        arrays[name] = a
        setattr(outs, name, a.ctypes.data_as(C.POINTER(ct)))
    rc = lib().oracle_is_run(network.net_ptr, C.byref(query), C.c_uint64(seed), C.c_uint64(run_begin),
                             C.c_uint64(n_runs), thp, int(bool(early_stop)), int(threads), C.byref(outs))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return arrays


def is_record(network, query, *, seed: int, run: int, theta, var_scale, fpad: int, lpad: int,
              early_stop: bool = False, cap: int = 4096):
    """Macro-steps of one tilted run: (X bf16 bits [k][fpad], A [k][lpad], B [k][lpad])."""
    th, thp = _theta_ptr(theta)
    vs = np.ascontiguousarray(var_scale, dtype=np.float32)
    X = np.zeros((cap, fpad), dtype=np.uint16)
    Aa = np.zeros((cap, lpad), dtype=np.float32)
    Bb = np.zeros((cap, lpad), dtype=np.float32)
    n = lib().oracle_is_record(network.net_ptr, C.byref(query), C.c_uint64(seed), C.c_uint64(run), thp,
                               int(bool(early_stop)), vs.ctypes.data_as(C.POINTER(C.c_float)), int(fpad), int(lpad),
                               X.ctypes.data_as(C.POINTER(C.c_uint16)), Aa.ctypes.data_as(C.POINTER(C.c_float)),
                               Bb.ctypes.data_as(C.POINTER(C.c_float)), cap)
    if n < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    if n > cap:
        raise RuntimeError(f"run has {n} macro-steps, more than cap={cap}")
    return X[:n], Aa[:n], Bb[:n]


def cost(network, query, *, seed: int, n_runs: int, run_begin: int = 0, threads: int = 0):
    """(steps, transitions, algorithmic lane-ops) of the runs, by the cost
    model of oracle.cpp."""
    st, tr, ops = C.c_uint64(), C.c_uint64(), C.c_double()
    rc = lib().oracle_cost(network.net_ptr, C.byref(query), C.c_uint64(seed), C.c_uint64(run_begin),
                           C.c_uint64(n_runs), int(threads), C.byref(st), C.byref(tr), C.byref(ops))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return st.value, tr.value, ops.value
