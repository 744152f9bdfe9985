"""fastlbp_amd -- MI355X-native loopy belief propagation for causal factor graphs.

The engine is libfastlbp.so (C ABI in include/fastlbp.h, C++ drop-in classes
in include/ext and include/lbp).  This package is the Python verification
harness over it.
"""
from ._native import FlbpError, LIB_PATH, lib
from .engine import (FactorList, Generated, Graph, Plan, RunResult, Schedule, ScheduleInfo,
                     algorithmic_bytes_per_sweep, clamp, generate, rank_alarms, read_causal_fg, seed,
                     write_causal_fg)

__all__ = [
    "FlbpError", "LIB_PATH", "lib", "FactorList", "Generated", "Graph", "Plan", "RunResult",
    "Schedule", "ScheduleInfo", "algorithmic_bytes_per_sweep", "clamp", "generate", "rank_alarms",
    "read_causal_fg", "seed", "write_causal_fg",
]
