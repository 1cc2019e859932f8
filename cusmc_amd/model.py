"""Parsed networks and the GPU simulator, the host-side mirror of cuSMC's
smc() flow (smc.cu:283-491): parse a UPPAAL model, build the device model,
run a query, reduce.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _capi as A


class ParsedNetwork:
    """A UPPAAL model parsed by the native front end (uppaal_xml_parser)."""

    def __init__(self, handle: C.c_void_p, source: str):
        self._h = handle
        self.source = source

    @classmethod
    def from_file(cls, path: str | Path, loc_query: str = "") -> "ParsedNetwork":
        h = C.c_void_p()
        A.check(A.lib().smc_parse_uppaal(str(path).encode(), loc_query.encode(), C.byref(h)))
        return cls(h, str(path))

    @classmethod
    def from_string(cls, xml: str, loc_query: str = "") -> "ParsedNetwork":
        h = C.c_void_p()
        A.check(A.lib().smc_parse_uppaal_string(xml.encode(), loc_query.encode(), C.byref(h)))
        return cls(h, "<string>")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and A._lib is not None:
            A._lib.smc_parsed_destroy(h)
            self._h = None

    @property
    def net(self) -> A.Network:
        return A.lib().smc_parsed_network(self._h).contents

    @property
    def net_ptr(self):
        return A.lib().smc_parsed_network(self._h)

    @property
    def n_components(self) -> int:
        return self.net.n_components

    @property
    def n_vars(self) -> int:
        return self.net.n_vars

    def component_name(self, c: int) -> str:
        return A.lib().smc_parsed_component_name(self._h, c).decode()

    def location_name(self, l: int) -> str:
        return A.lib().smc_parsed_location_name(self._h, l).decode()

    def variable_name(self, v: int) -> str:
        return A.lib().smc_parsed_variable_name(self._h, v).decode()

    def variable_id(self, name: str) -> int:
        for v in range(self.n_vars):
            if self.variable_name(v) == name:
                return v
        raise KeyError(name)

    @property
    def goal_count(self) -> int:
        return A.lib().smc_parsed_goal_count(self._h)

    def var_kinds(self) -> list[int]:
        n = self.net
        return [n.var_kind[i] for i in range(n.n_vars)]


def reach_query(time_bound: float, *, max_steps: int = 0) -> A.Query:
    return A.Query(A.QUERY_REACH, -1, 0, 0, 0.0, float(time_bound), int(max_steps), 0)


def var_query(var: int, time_bound: float, *, is_max: bool = True, estimate: bool = False,
              threshold: float = 0.0, max_steps: int = 0) -> A.Query:
    return A.Query(A.QUERY_VAR, int(var), int(bool(is_max)), int(bool(estimate)), float(threshold),
                   float(time_bound), int(max_steps), 0)


@dataclass
class RunResult:
    n_runs: int
    hits: int
    estimate: float
    ch_eps: float
    clt_half: float
    steps: int
    transitions: int
    truncated: int
    kernel_ms: float
    reduce_ms: float
    wall_ms: float
    lanes_per_run: int
    waves: int
    sum: float = 0.0
    sumsq: float = 0.0
    vmin: float = 0.0
    vmax: float = 0.0
    per_run: dict = field(default_factory=dict)

    @property
    def runs_per_sec(self) -> float:
        return self.n_runs / (self.kernel_ms / 1e3) if self.kernel_ms > 0 else float("nan")

    @property
    def transitions_per_sec(self) -> float:
        return self.transitions / (self.kernel_ms / 1e3) if self.kernel_ms > 0 else float("nan")


class GpuSimulator:
    """The device model of one network on one GPU (smc_model)."""

    def __init__(self, network: ParsedNetwork, device: int = 0):
        self._h = None
        self.network = network
        h = C.c_void_p()
        A.check(A.lib().smc_model_create(network.net_ptr, int(device), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and A._lib is not None:
            A._lib.smc_model_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def autotune(self, query: A.Query, *, seed: int = 1, sample_runs: int = 65536) -> A.Launch:
        """Pick the fastest launch shape on a sample; later runs use it."""
        best = A.Launch()
        A.check(A.lib().smc_autotune(self._h, C.byref(query), C.c_uint64(seed), C.c_uint64(sample_runs),
                                     C.byref(best)))
        return best

    def run(self, query: A.Query, *, seed: int, n_runs: int, run_begin: int = 0,
            lanes_per_run: int = 0, waves: int = 0, waves_per_cu: int = 0, occupancy: int = 0,
            tables: int = 0, per_run: tuple[str, ...] = ()) -> RunResult:
        outs = A.RunOutputs()
        arrays = {}
        spec = {"flag": np.uint8, "value": np.float64, "steps": np.uint32, "transitions": np.uint32,
                "digest": np.uint64, "time": np.float64}
        ctype = {"flag": A.u8, "value": A.f64, "steps": A.u32, "transitions": A.u32, "digest": A.u64,
                 "time": A.f64}
        for name in per_run:
            arr = np.zeros(n_runs, dtype=spec[name])
            arrays[name] = arr
            setattr(outs, name, arr.ctypes.data_as(C.POINTER(ctype[name])))
        launch = A.Launch(int(lanes_per_run), int(waves), int(waves_per_cu), int(occupancy), int(tables))
        res = A.Result()
        A.check(A.lib().smc_run(self._h, C.byref(query), C.c_uint64(seed), C.c_uint64(run_begin),
                                C.c_uint64(n_runs), C.byref(launch), C.byref(outs), C.byref(res)))
        return _result(res, arrays)


def _result(res: A.Result, per_run=None) -> RunResult:
    return RunResult(res.n_runs, res.hits, res.estimate, res.ch_eps, res.clt_half, res.steps, res.transitions,
                     res.truncated, res.kernel_ms, res.reduce_ms, res.wall_ms, res.lanes_per_run, res.waves,
                     res.sum, res.sumsq, res.vmin, res.vmax, per_run or {})


class Batch:
    """Many (simulator, query, seed, runs) instances in one captured launch
    per round (smc_batch).  Each round simulates the next n_runs runs of
    every instance; run() returns per-instance totals over its rounds."""

    def __init__(self, items, *, tune: bool = False):
        self._h = None
        self._keep = [it[0] for it in items]
        arr = (A.BatchItem * len(items))()
        for k, (sim, query, seed, n_runs, *rest) in enumerate(items):
            arr[k] = A.BatchItem(sim._h, query, int(seed), int(rest[0]) if rest else 0, int(n_runs))
        h = C.c_void_p()
        A.check(A.lib().smc_batch_create(arr, len(items), int(bool(tune)), C.byref(h)))
        self._h = h
        self.n = len(items)

    def run(self, rounds: int = 1):
        res = (A.Result * self.n)()
        st = A.BatchStats()
        A.check(A.lib().smc_batch_run(self._h, int(rounds), res, C.byref(st)))
        return [_result(r) for r in res], st

    def close(self) -> None:
        if getattr(self, "_h", None) and A._lib is not None:
            A._lib.smc_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


@dataclass
class IsResult:
    n_runs: int
    hits: int
    estimate: float
    rel_half_width: float
    ms: float
    ess: float
    steps: int
    transitions: int
    per_run: dict = field(default_factory=dict)

    @property
    def half_width(self) -> float:
        return self.rel_half_width * self.estimate


def _is_result(r: A.IsResult, per_run=None) -> IsResult:
    return IsResult(r.n_runs, r.hits, r.estimate, r.rel_half_width, r.ms, r.ess, r.steps, r.transitions,
                    per_run or {})


class ImportanceSampler:
    """Learned importance sampling of one probability query (smc_is): an MLP
    over the runs' states tilts the exponential rates; every run carries its
    log2 likelihood ratio so that mean(hit * 2^log2lr) stays unbiased."""

    def __init__(self, sim: GpuSimulator, query: A.Query, *, hidden: int = 128, rounds: int = 5,
                 runs_per_round: int = 100000, elite_frac: float = 0.1, epochs: int = 20,
                 learning_rate: float = 1e-3, tilt_max: float = 4.0, per_location: bool = False, seed: int = 1):
        self._h = None
        self.sim = sim
        self.query = query
        self.config = A.IsConfig(int(hidden), int(rounds), int(runs_per_round), float(elite_frac), int(epochs),
                                 float(learning_rate), float(tilt_max), int(bool(per_location)), int(seed))
        h = C.c_void_p()
        A.check(A.lib().smc_is_create(sim._h, C.byref(query), C.byref(self.config), C.byref(h)))
        self._h = h
        fp, hd, op, lp, n = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_uint64()
        A.check(A.lib().smc_is_shape(self._h, C.byref(fp), C.byref(hd), C.byref(op), C.byref(lp), C.byref(n)))
        self.features_padded, self.hidden, self.outputs_padded = fp.value, hd.value, op.value
        self.locations_padded, self.n_params = lp.value, n.value
        self.group = np.zeros(max(sim.network.net.n_locations, 1), dtype=np.int32)
        self.n_groups = A.lib().smc_is_groups(self._h, self.group.ctypes.data_as(C.POINTER(C.c_int32)))

    def close(self) -> None:
        if getattr(self, "_h", None) and A._lib is not None:
            A._lib.smc_is_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def train(self):
        """Cross-entropy rounds; returns one dict per round."""
        rounds = (A.IsRound * self.config.rounds)()
        A.check(A.lib().smc_is_train(self._h, rounds))
        return [dict(level=r.level, elites=r.elites, hits=r.hits, estimate=r.estimate, ms=r.ms, ess=r.ess,
                     pairs=r.pairs) for r in rounds]

    def estimate(self, *, seed: int, run_begin: int = 0, batch_runs: int = 65536, max_runs: int = 0,
                 rel_target: float = 0.1) -> IsResult:
        r = A.IsResult()
        A.check(A.lib().smc_is_estimate(self._h, C.c_uint64(seed), C.c_uint64(run_begin), C.c_uint64(batch_runs),
                                        C.c_uint64(max_runs), float(rel_target), C.byref(r)))
        return _is_result(r)

    def sample(self, *, seed: int, n_runs: int, run_begin: int = 0, theta=None, use_mlp: bool = False,
               lockstep: bool = True, early_stop: bool = False) -> IsResult:
        """Runs under a fixed per-location tilt (theta), the MLP, or no tilt;
        per_run holds flag, log2lr, digest, steps, transitions, time."""
        spec = {"flag": (np.uint8, A.u8), "log2lr": (np.float64, A.f64), "digest": (np.uint64, A.u64),
                "steps": (np.uint32, A.u32), "transitions": (np.uint32, A.u32), "time": (np.float64, A.f64)}
        outs = A.IsOutputs()
        arrays = {}
        for name, (dt, ct) in spec.items():
            a = np.zeros(n_runs, dtype=dt)
            arrays[name] = a
            setattr(outs, name, a.ctypes.data_as(C.POINTER(ct)))
        th = None
        if theta is not None:
            th = np.ascontiguousarray(theta, dtype=np.float32)
        r = A.IsResult()
        A.check(A.lib().smc_is_sample(self._h, th.ctypes.data_as(C.POINTER(C.c_float)) if th is not None else None,
                                      int(bool(use_mlp)), C.c_uint64(seed), C.c_uint64(run_begin), C.c_uint64(n_runs),
                                      int(bool(lockstep)), int(bool(early_stop)), C.byref(outs), C.byref(r)))
        return _is_result(r, arrays)

    def record(self, runs, *, seed: int, theta=None, weights=None, early_stop: bool = False, cap: int = 1 << 16):
        """Replay runs and return their recorded macro-steps (X bf16 bits,
        A, B), numbered macro-step-major, run-minor."""
        rr = np.ascontiguousarray(runs, dtype=np.uint64)
        th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float32)
        ww = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        X = np.zeros((cap, self.features_padded), dtype=np.uint16)
        Aa = np.zeros((cap, self.locations_padded), dtype=np.float32)
        Bb = np.zeros((cap, self.locations_padded), dtype=np.float32)
        n = C.c_uint64()
        A.check(A.lib().smc_is_record(
            self._h, th.ctypes.data_as(C.POINTER(C.c_float)) if th is not None else None, C.c_uint64(seed),
            rr.ctypes.data_as(C.POINTER(A.u64)), len(rr), ww.ctypes.data_as(C.POINTER(C.c_float)) if ww is not None else None,
            int(bool(early_stop)), X.ctypes.data_as(C.POINTER(C.c_uint16)), Aa.ctypes.data_as(C.POINTER(C.c_float)),
            Bb.ctypes.data_as(C.POINTER(C.c_float)), C.c_uint64(cap), C.byref(n)))
        k = min(n.value, cap)
        return X[:k], Aa[:k], Bb[:k]

    def ce_gradient(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        A.check(A.lib().smc_is_ce_gradient(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        A.check(A.lib().smc_is_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.n_params
        A.check(A.lib().smc_is_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def mlp_eval(self, x, g=None):
        """theta = MLP(x) for x [rows, features_padded]; with g [rows,
        locations_padded] also the parameter gradient of sum(g * theta)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        rows = x.shape[0]
        theta = np.zeros((rows, self.locations_padded), dtype=np.float32)
        grad = None
        gp = None
        if g is not None:
            gg = np.ascontiguousarray(g, dtype=np.float32)
            grad = np.zeros(self.n_params, dtype=np.float32)
            gp = gg.ctypes.data_as(C.POINTER(C.c_float))
        A.check(A.lib().smc_is_mlp_eval(self._h, x.ctypes.data_as(C.POINTER(C.c_float)), rows, gp,
                                        theta.ctypes.data_as(C.POINTER(C.c_float)),
                                        grad.ctypes.data_as(C.POINTER(C.c_float)) if grad is not None else None))
        return theta, grad


def ch_epsilon(n: int, delta: float = 0.05) -> float:
    return A.lib().smc_ch_epsilon(int(n), float(delta))
