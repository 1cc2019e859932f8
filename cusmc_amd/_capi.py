"""ctypes view of the C ABI (include/smc_capi.h, include/smc_network.h).

Loading fails loudly: there is no Python or CPU fallback for the simulator.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("CUSMC_LIB", _HERE / "lib" / "libcusmc.so"))

i32, u32, u64, u8, f64 = C.c_int32, C.c_uint32, C.c_uint64, C.c_uint8, C.c_double
P = C.POINTER

# enums (smc_network.h)
QUERY_REACH, QUERY_VAR = 0, 1
VAR_INT, VAR_CLOCK = 0, 1
LOC_NORMAL, LOC_GOAL, LOC_BRANCH, LOC_URGENT, LOC_COMMITTED = range(5)
DEFAULT_MAX_STEPS = 1 << 24
SMC_OK, SMC_E_INVALID, SMC_E_PARSE, SMC_E_DEVICE, SMC_E_LIMIT, SMC_E_ALLOC = 0, -1, -2, -3, -4, -5


class ExprNode(C.Structure):
    _fields_ = [("op", i32), ("a", i32), ("b", i32), ("c", i32), ("var", i32), ("reserved", i32), ("value", f64)]


class Constraint(C.Structure):
    _fields_ = [("op", i32), ("uses_variable", i32), ("var", i32), ("lhs", i32), ("rhs", i32)]


class Location(C.Structure):
    _fields_ = [("type", i32), ("rate", i32), ("inv_first", i32), ("inv_count", i32),
                ("edge_first", i32), ("edge_count", i32)]


class Edge(C.Structure):
    _fields_ = [("src", i32), ("dst", i32), ("channel", i32), ("weight", i32), ("guard_first", i32),
                ("guard_count", i32), ("upd_first", i32), ("upd_count", i32)]


class Update(C.Structure):
    _fields_ = [("var", i32), ("expr", i32)]


class Network(C.Structure):
    _fields_ = [("n_components", i32), ("n_vars", i32), ("n_locations", i32), ("n_edges", i32),
                ("n_constraints", i32), ("n_updates", i32), ("n_exprs", i32), ("reserved", i32),
                ("var_kind", P(i32)), ("var_init", P(f64)), ("comp_first_loc", P(i32)),
                ("comp_init_loc", P(i32)), ("locations", P(Location)), ("edges", P(Edge)),
                ("constraints", P(Constraint)), ("updates", P(Update)), ("exprs", P(ExprNode))]


class Query(C.Structure):
    _fields_ = [("kind", i32), ("var", i32), ("is_max", i32), ("is_estimate", i32),
                ("threshold", f64), ("time_bound", f64), ("max_steps", u32), ("reserved", u32)]


class Launch(C.Structure):
    _fields_ = [("lanes_per_run", i32), ("waves", i32), ("waves_per_cu", i32), ("occupancy", i32),
                ("tables", i32)]


class RunOutputs(C.Structure):
    _fields_ = [("flag", P(u8)), ("value", P(f64)), ("steps", P(u32)), ("transitions", P(u32)),
                ("digest", P(u64)), ("time", P(f64))]


class Result(C.Structure):
    _fields_ = [("n_runs", u64), ("hits", u64), ("estimate", f64), ("sum", f64), ("sumsq", f64),
                ("vmin", f64), ("vmax", f64), ("ch_eps", f64), ("clt_half", f64), ("steps", u64),
                ("transitions", u64), ("truncated", u64), ("kernel_ms", f64), ("reduce_ms", f64), ("wall_ms", f64),
                ("lanes_per_run", i32), ("waves", i32)]


class BatchItem(C.Structure):
    _fields_ = [("model", C.c_void_p), ("query", Query), ("seed", u64), ("run_begin", u64), ("n_runs", u64)]


class BatchStats(C.Structure):
    _fields_ = [("graph_ms", f64), ("runs", u64), ("steps", u64), ("transitions", u64), ("groups", i32),
                ("rounds", i32)]


# Every symbol include/smc_capi.h declares, with its signature.
class IsConfig(C.Structure):
    _fields_ = [("hidden", C.c_int32), ("rounds", C.c_int32), ("runs_per_round", C.c_uint64),
                ("elite_frac", C.c_double), ("epochs", C.c_int32), ("learning_rate", C.c_float),
                ("tilt_max", C.c_float), ("per_location", C.c_int32), ("seed", C.c_uint64)]


class IsRound(C.Structure):
    _fields_ = [("level", C.c_double), ("elites", C.c_uint64), ("hits", C.c_uint64),
                ("estimate", C.c_double), ("ms", C.c_double), ("ess", C.c_double), ("pairs", C.c_uint64)]


class IsResult(C.Structure):
    _fields_ = [("n_runs", C.c_uint64), ("hits", C.c_uint64), ("estimate", C.c_double),
                ("rel_half_width", C.c_double), ("ms", C.c_double), ("ess", C.c_double),
                ("steps", C.c_uint64), ("transitions", C.c_uint64)]


class IsOutputs(C.Structure):
    _fields_ = [("flag", C.POINTER(C.c_uint8)), ("log2lr", C.POINTER(C.c_double)),
                ("digest", C.POINTER(C.c_uint64)), ("steps", C.POINTER(C.c_uint32)),
                ("transitions", C.POINTER(C.c_uint32)), ("time", C.POINTER(C.c_double))]


SIGNATURES = {
    "smc_parse_uppaal": (i32, [C.c_char_p, C.c_char_p, P(C.c_void_p)]),
    "smc_parse_uppaal_string": (i32, [C.c_char_p, C.c_char_p, P(C.c_void_p)]),
    "smc_parsed_network": (P(Network), [C.c_void_p]),
    "smc_parsed_component_name": (C.c_char_p, [C.c_void_p, C.c_int]),
    "smc_parsed_location_name": (C.c_char_p, [C.c_void_p, C.c_int]),
    "smc_parsed_variable_name": (C.c_char_p, [C.c_void_p, C.c_int]),
    "smc_parsed_goal_count": (C.c_int, [C.c_void_p]),
    "smc_parsed_destroy": (None, [C.c_void_p]),
    "smc_model_create": (i32, [P(Network), C.c_int, P(C.c_void_p)]),
    "smc_model_destroy": (None, [C.c_void_p]),
    "smc_run": (i32, [C.c_void_p, P(Query), u64, u64, u64, P(Launch), P(RunOutputs), P(Result)]),
    "smc_run_noalloc": (i32, [C.c_void_p, P(Query), u64, u64, u64, P(Launch), P(Result)]),
    "smc_autotune": (i32, [C.c_void_p, P(Query), u64, u64, P(Launch)]),
    "smc_batch_create": (i32, [P(BatchItem), C.c_int, C.c_int, P(C.c_void_p)]),
    "smc_batch_run": (i32, [C.c_void_p, C.c_int, P(Result), P(BatchStats)]),
    "smc_batch_destroy": (None, [C.c_void_p]),
    "smc_is_create": (i32, [C.c_void_p, P(Query), P(IsConfig), P(C.c_void_p)]),
    "smc_is_train": (i32, [C.c_void_p, P(IsRound)]),
    "smc_is_estimate": (i32, [C.c_void_p, u64, u64, u64, u64, f64, P(IsResult)]),
    "smc_is_sample": (i32, [C.c_void_p, P(C.c_float), C.c_int, u64, u64, u64, C.c_int, C.c_int, P(IsOutputs),
                            P(IsResult)]),
    "smc_is_shape": (i32, [C.c_void_p, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32), P(u64)]),
    "smc_is_groups": (C.c_int, [C.c_void_p, P(C.c_int32)]),
    "smc_is_get_params": (i32, [C.c_void_p, P(C.c_float)]),
    "smc_is_set_params": (i32, [C.c_void_p, P(C.c_float)]),
    "smc_is_mlp_eval": (i32, [C.c_void_p, P(C.c_float), C.c_int, P(C.c_float), P(C.c_float), P(C.c_float)]),
    "smc_is_record": (i32, [C.c_void_p, P(C.c_float), u64, P(u64), C.c_int, P(C.c_float), C.c_int, P(C.c_uint16),
                            P(C.c_float), P(C.c_float), u64, P(u64)]),
    "smc_is_ce_gradient": (i32, [C.c_void_p, P(C.c_float)]),
    "smc_is_destroy": (None, [C.c_void_p]),
    "smc_ch_epsilon": (f64, [u64, f64]),
    "smc_last_error": (C.c_char_p, []),
    "smc_version": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded libcusmc.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libcusmc.so not found at {LIB_PATH}: build it with "
                               f"`python -c 'import __graft_entry__ as g; g.build()'` or "
                               f"`python cusmc_amd/build.py`")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class SmcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(code: int) -> None:
    if code != SMC_OK:
        raise SmcError(code, lib().smc_last_error().decode(errors="replace"))
