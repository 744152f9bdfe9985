"""flbp_plan_run_hist_noalloc (the belief-history run with its history in a
fixed per-device array) against flbp_plan_run_hist (history allocated per
call), both through the C ABI: the same stop reason, sweep count, moved
fraction, max change, low-pass beliefs and plan beliefs, bit for bit, on
calls in a row whose sizes shrink, so that a stale history of the previous
call would show; and the size limit refused before any sweep."""
import ctypes as C
import math

import numpy as np
import pytest

import fastlbp_amd as fl
from fastlbp_amd._native import HistResultC, lib

pytestmark = pytest.mark.gpu

FLBP_ERANGE = -5


def run(entry, graph, tol, min_iters, max_iters, hist_len, damping):
    plan = fl.Plan(graph, device=0)
    plan.init()
    res = HistResultC()
    low = np.full(max(graph.n_vars, 1), np.nan)
    rc = getattr(lib(), entry)(plan._h, tol, min_iters, max_iters, hist_len, math.inf, damping, C.byref(res),
                               low.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, (entry, rc, lib().flbp_last_error())
    summary = (res.yet_to_converge, res.maxdiff, int(res.iters), int(res.reason))
    return summary, low[:graph.n_vars], plan.beliefs()


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def check(graph, tol, min_iters, max_iters, hist_len, damping, want_reason):
    print(f"[no-alloc] V={graph.n_vars} E={graph.n_edges} hist_len={hist_len} tol={tol} damping={damping}",
          flush=True)
    s_ref, low_ref, bel_ref = run("flbp_plan_run_hist", graph, tol, min_iters, max_iters, hist_len, damping)
    s_new, low_new, bel_new = run("flbp_plan_run_hist_noalloc", graph, tol, min_iters, max_iters, hist_len, damping)
    assert s_ref[3] in want_reason, s_ref
    assert s_new[2:] == s_ref[2:]
    assert same_bits(s_new[:2], s_ref[:2])
    assert same_bits(low_new, low_ref)
    assert same_bits(bel_new, bel_ref)
    if s_ref[3] != 0:   # the low-pass beliefs are ring means, not the last beliefs
        assert not same_bits(low_ref, bel_ref)


def test_noalloc_history_matches_the_allocating_run_on_shrinking_calls():
    big = fl.Graph(fl.generate(20000, 7).factors)
    small = fl.Graph(fl.generate(2000, 8).factors)
    # the ring is filled past its length: low-pass = mean over the ring
    check(big, 1e-300, 0, 30, 16, 0.3, (1, 2))
    # a smaller graph and a shorter ring next: stale counters or ring slots
    # of the first call would change its means
    check(small, 1e-300, 5, 12, 4, 0.0, (1, 2))
    # converged: the low-pass beliefs are the beliefs
    check(small, 1e-6, 0, 1000, 8, 0.0, (0,))


def test_noalloc_history_beyond_the_fixed_array_is_refused():
    g = fl.Graph(fl.generate(20000, 7).factors)
    plan = fl.Plan(g, device=0)
    plan.init()
    res = HistResultC()
    low = np.zeros(g.n_vars)
    for hist_len, ok in ((417, False), (416, True)):
        rc = lib().flbp_plan_run_hist_noalloc(plan._h, 1e-6, 0, 3, hist_len, math.inf, 0.0, C.byref(res),
                                               low.ctypes.data_as(C.POINTER(C.c_double)))
        if ok:
            assert rc == 0, lib().flbp_last_error()
        else:
            assert rc == FLBP_ERANGE
            assert b"fixed" in lib().flbp_last_error()
