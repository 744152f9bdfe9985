"""ctypes binding of libfastlbp.so (the C ABI declared in include/fastlbp.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C fastlbp_amd/csrc``).  Loading fails loudly when it is missing:
there is no CPU fallback for the engine.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfastlbp.so")

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class FactorListC(C.Structure):
    _fields_ = [
        ("n_factors", C.c_uint64),
        ("type", i32p),
        ("head", u64p),
        ("body_off", u64p),
        ("body", u64p),
        ("p", f64p),
        ("q", f64p),
        ("head_clamped", u8p),
        ("mask0", f64p),
        ("mask1", f64p),
    ]


class CsrC(C.Structure):
    _fields_ = [
        ("n_vars", C.c_uint64),
        ("n_factors", C.c_uint64),
        ("n_edges", C.c_uint64),
        ("var_label", u64p),
        ("fac_off", u64p),
        ("edge_var", u64p),
        ("var_off", u64p),
        ("var_edge", u64p),
        ("fac_type", i32p),
        ("head_slot", u64p),
        ("head_clamped", u8p),
        ("p", f64p),
        ("q", f64p),
        ("mask0", f64p),
        ("mask1", f64p),
    ]


class GenParamsC(C.Structure):
    _fields_ = [
        ("n_vars", C.c_uint64),
        ("seed", C.c_uint64),
        ("n_layers", C.c_uint32),
        ("k_min", C.c_uint32),
        ("k_max", C.c_uint32),
        ("b_min", C.c_uint32),
        ("b_max", C.c_uint32),
        ("n_rules", C.c_uint32),
        ("leaf_frac", C.c_double),
        ("p_and_lo", C.c_double),
        ("p_and_hi", C.c_double),
        ("p_leaf_lo", C.c_double),
        ("p_leaf_hi", C.c_double),
        ("back_frac", C.c_double),
        ("n_alarms", C.c_uint64),
        ("extra_frac", C.c_double),
        ("extra_q", C.c_double),
        ("alarm_free", C.c_uint32),
    ]


class RunOptsC(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("maxiter", C.c_uint64),
        ("maxtime", C.c_double),
        ("damping", C.c_double),
        ("chunk", C.c_uint32),
    ]


class BatchOptsC(C.Structure):
    _fields_ = [
        ("tile_factors", C.c_uint32),
        ("tile_vars", C.c_uint32),
        ("grid_blocks", C.c_uint32),
        ("chunk", C.c_uint32),
        ("use_graph", C.c_uint32),
        ("layout", C.c_uint32),
    ]


class BatchGroupC(C.Structure):
    _fields_ = [
        ("graph", C.c_void_p),
        ("n_variants", C.c_uint32),
        ("clamp_off", C.c_void_p),
        ("clamp_var", C.c_void_p),
        ("clamp_val", C.c_void_p),
    ]


LAYOUT_AUTO, LAYOUT_PULL, LAYOUT_SCATTER = 0, 1, 2   # FLBP_LAYOUT_* of fastlbp.h


class ScheduleInfoC(C.Structure):
    _fields_ = [
        ("n_order", C.c_uint64),
        ("n_waves", C.c_uint64),
        ("n_fv_items", C.c_uint64),
        ("n_vf_items", C.c_uint64),
        ("n_levels", C.c_uint64),
        ("updates", C.c_int32),
        ("respect_preorder", C.c_int32),
        ("n_exec_items", C.c_uint64),
    ]


UPDATES_SEQFIX, UPDATES_PARALL, UPDATES_TOPO = 0, 1, 2   # FLBP_UPDATES_* of fastlbp.h


class RunResultC(C.Structure):
    _fields_ = [
        ("iters", C.c_uint64),
        ("sweeps", C.c_uint64),
        ("maxdiff", C.c_double),
        ("seconds", C.c_double),
    ]


class HistResultC(C.Structure):
    _fields_ = [
        ("yet_to_converge", C.c_double),
        ("maxdiff", C.c_double),
        ("iters", C.c_uint64),
        ("reason", C.c_int),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "flbp_graph_build": (C.c_int, [C.POINTER(FactorListC), C.POINTER(C.c_void_p)]),
    "flbp_graph_csr": (C.c_int, [C.c_void_p, C.POINTER(CsrC)]),
    "flbp_graph_free": (None, [C.c_void_p]),
    "flbp_write_graph": (C.c_int, [C.POINTER(CsrC), C.c_char_p]),
    "flbp_read_graph": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "flbp_read_causal_fg": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "flbp_write_causal_fg": (C.c_int, [C.POINTER(FactorListC), C.c_char_p]),
    "flbp_clamp": (C.c_int, [C.POINTER(FactorListC), C.c_uint64, u64p, u8p, C.POINTER(C.c_void_p)]),
    "flbp_flist_view": (C.c_int, [C.c_void_p, C.POINTER(FactorListC)]),
    "flbp_flist_free": (None, [C.c_void_p]),
    "flbp_gen_defaults": (None, [C.POINTER(GenParamsC)]),
    "flbp_generate": (C.c_int, [C.POINTER(GenParamsC), C.POINTER(C.c_void_p)]),
    "flbp_generated_view": (C.c_int, [C.c_void_p, C.POINTER(FactorListC)]),
    "flbp_generated_extra": (C.c_int, [C.c_void_p, u64p, C.POINTER(u64p), C.POINTER(i32p)]),
    "flbp_generated_free": (None, [C.c_void_p]),
    "flbp_rank_alarms": (C.c_int, [C.c_uint64, u64p, f64p, u64p]),
    "flbp_plan_create": (C.c_int, [C.POINTER(CsrC), C.c_int, C.POINTER(C.c_void_p)]),
    "flbp_plan_destroy": (None, [C.c_void_p]),
    "flbp_plan_init": (C.c_int, [C.c_void_p, C.c_int]),
    "flbp_plan_run": (C.c_int, [C.c_void_p, C.POINTER(RunOptsC), C.POINTER(RunResultC)]),
    "flbp_plan_run_fixed": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(RunResultC)]),
    "flbp_plan_run_hist": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double,
                                     C.c_double, C.POINTER(HistResultC), C.POINTER(C.c_double)]),
    "flbp_plan_run_hist_noalloc": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double,
                                     C.c_double, C.POINTER(HistResultC), C.POINTER(C.c_double)]),
    "flbp_plan_beliefs": (C.c_int, [C.c_void_p, f64p, C.c_uint64]),
    "flbp_plan_reset": (C.c_int, [C.c_void_p]),
    "flbp_plan_get_messages": (C.c_int, [C.c_void_p, f64p, f64p]),
    "flbp_plan_set_messages": (C.c_int, [C.c_void_p, f64p, f64p]),
    "flbp_plan_set_p": (C.c_int, [C.c_void_p, f64p, C.c_uint64]),
    "flbp_plan_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "flbp_plan_last_timing": (C.c_int, [C.c_void_p, f64p, f64p, u64p, u64p]),
    "flbp_plan_device_bytes": (C.c_uint64, [C.c_void_p]),
    "flbp_plan_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "flbp_plan_get_variant": (C.c_int, [C.c_void_p]),
    "flbp_plan_use_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "flbp_plan_beliefs_at": (C.c_int, [C.c_void_p, C.c_uint64, u64p, f64p]),
    "flbp_plan_set_head_clamps": (C.c_int, [C.c_void_p, C.c_uint64, u64p, u8p]),
    "flbp_plan_reset_head_clamps": (C.c_int, [C.c_void_p]),
    "flbp_plan_clamp_vars": (C.c_int, [C.c_void_p, C.c_uint64, u64p, u8p]),
    "flbp_batch_default_opts": (None, [C.POINTER(BatchOptsC)]),
    "flbp_batch_create": (C.c_int, [C.c_uint64, C.POINTER(C.POINTER(CsrC)), C.c_int, C.POINTER(BatchOptsC),
                                    C.POINTER(C.c_void_p)]),
    "flbp_batch_create_groups": (C.c_int, [C.c_uint64, C.POINTER(BatchGroupC), C.c_int, C.POINTER(BatchOptsC),
                                             C.POINTER(C.c_void_p)]),
    "flbp_batch_destroy": (None, [C.c_void_p]),
    "flbp_batch_n_variants": (C.c_uint64, [C.c_void_p]),
    "flbp_batch_beliefs_at": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, u64p, f64p]),
    "flbp_batch_set_p": (C.c_int, [C.c_void_p, C.c_uint64, f64p, C.c_uint64]),
    "flbp_batch_init": (C.c_int, [C.c_void_p, C.c_int]),
    "flbp_batch_run": (C.c_int, [C.c_void_p, C.POINTER(RunOptsC), u64p, f64p, C.POINTER(RunResultC)]),
    "flbp_batch_beliefs": (C.c_int, [C.c_void_p, C.c_uint64, f64p, C.c_uint64]),
    "flbp_batch_set_opts": (C.c_int, [C.c_void_p, C.POINTER(BatchOptsC)]),
    "flbp_batch_get_opts": (C.c_int, [C.c_void_p, C.POINTER(BatchOptsC)]),
    "flbp_batch_reset_counters": (C.c_int, [C.c_void_p]),
    "flbp_batch_autotune": (C.c_int, [C.c_void_p, C.c_double, C.c_uint32, C.POINTER(BatchOptsC)]),
    "flbp_batch_device_bytes": (C.c_uint64, [C.c_void_p]),
    "flbp_adjoint_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_double, C.POINTER(C.c_void_p)]),
    "flbp_adjoint_destroy": (None, [C.c_void_p]),
    "flbp_adjoint_forward": (C.c_int, [C.c_void_p, f64p]),
    "flbp_adjoint_backward": (C.c_int, [C.c_void_p, f64p, f64p]),
    "flbp_adjoint_last_timing": (C.c_int, [C.c_void_p, f64p, f64p]),
    "flbp_adjoint_backward_sparse": (C.c_int, [C.c_void_p, C.c_uint64, u64p, f64p]),
    "flbp_adjoint_clear_grad": (C.c_int, [C.c_void_p]),
    "flbp_adjoint_get_dl_dp": (C.c_int, [C.c_void_p, f64p]),
    "flbp_adjoint_set_rules": (C.c_int, [C.c_void_p, i32p, C.c_uint64]),
    "flbp_adjoint_rule_grad": (C.c_int, [C.c_void_p, f64p]),
    "flbp_adjoint_device_bytes": (C.c_uint64, [C.c_void_p]),
    "flbp_tied_p": (C.c_int, [C.c_uint64, f64p, i32p, C.c_uint64, f64p, f64p]),
    "flbp_rule_grad": (C.c_int, [C.c_uint64, f64p, f64p, i32p, C.c_uint64, f64p]),
    "flbp_schedule_create": (C.c_int, [C.POINTER(CsrC), C.c_int, C.POINTER(C.c_void_p)]),
    "flbp_schedule_from_order": (C.c_int, [C.POINTER(CsrC), u64p, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    "flbp_schedule_destroy": (None, [C.c_void_p]),
    "flbp_schedule_get_info": (C.c_int, [C.c_void_p, C.POINTER(ScheduleInfoC)]),
    "flbp_schedule_order": (C.c_int, [C.c_void_p, u64p]),
    "flbp_schedule_waves": (C.c_int, [C.c_void_p, u64p, u64p, u64p, u64p]),
    "flbp_schedule_levels": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "flbp_schedule_exec": (C.c_int, [C.c_void_p, u8p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "flbp_seed": (None, [C.c_uint32]),
    "flbp_plan_set_schedule": (C.c_int, [C.c_void_p, C.c_void_p]),
    "flbp_plan_get_updates": (C.c_int, [C.c_void_p]),
    "flbp_plan_set_seq_tail": (C.c_int, [C.c_void_p, C.c_uint32]),
    "flbp_last_error": (C.c_char_p, []),
    "flbp_version": (C.c_char_p, []),
}

_lib = None


class FlbpError(RuntimeError):
    """A libfastlbp.so entry point returned a negative status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "or make -C fastlbp_amd/csrc (there is no CPU fallback)")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(fn: str, rc: int) -> None:
    if rc != 0:
        msg = lib().flbp_last_error().decode(errors="replace")
        raise FlbpError(fn, rc, msg)


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    check(name, rc)
    return rc
