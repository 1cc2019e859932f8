"""cusmc_amd — statistical model checking of networks of stochastic timed
automata on one AMD MI355X (gfx950).

The simulation path of cuSMC (UPPAAL XML in, `-m/-q` queries, statistics
out) with a from-scratch HIP run kernel behind a C ABI
(include/smc_capi.h).
"""
from ._capi import SmcError, lib  # noqa: F401
from .model import (Batch, GpuSimulator, ImportanceSampler, IsResult, ParsedNetwork, RunResult,  # noqa: F401
                    ch_epsilon, reach_query, var_query)

__all__ = ["Batch", "GpuSimulator", "ImportanceSampler", "IsResult", "ParsedNetwork", "RunResult", "SmcError",
           "ch_epsilon", "lib", "reach_query", "var_query"]
