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
            lanes_per_run: int = 0, waves: int = 0, waves_per_cu: int = 0,
            per_run: tuple[str, ...] = ()) -> RunResult:
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
        launch = A.Launch(int(lanes_per_run), int(waves), int(waves_per_cu), 0)
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


def ch_epsilon(n: int, delta: float = 0.05) -> float:
    return A.lib().smc_ch_epsilon(int(n), float(delta))
