"""Batched solve of many graphs and clamped variants on one GPU (flbp_batch_*).

A batch holds groups: one graph structure solved for K = 1, 2, 4, ..., 64
variants, each variant a set of head clamps of variables without body uses
(the clamp of such a variable keeps the structure, so the variant is the
graph rebuilt with those clamps).  Independent graphs are groups of one
variant.  Variants are numbered across groups in order.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from typing import Optional, Sequence

import numpy as np

from ._native import BatchGroupC, BatchOptsC, CsrC, RunOptsC, RunResultC, call, lib
from .engine import Graph, RunResult, _ptr

MAX_VARIANTS_PER_GROUP = 64


class Group:
    """One structure and its variants: clamps[k] = (variable indices, 0/1
    values) of variant k (empty for the unclamped graph)."""

    def __init__(self, graph: Graph, clamps: Optional[Sequence] = None):
        self.graph = graph
        self.clamps = list(clamps) if clamps is not None else [((), ())]
        K = len(self.clamps)
        if K < 1 or K > MAX_VARIANTS_PER_GROUP or K & (K - 1):
            raise ValueError("variants per group must be 1, 2, 4, ..., 64")

    @property
    def n_variants(self) -> int:
        return len(self.clamps)


def split_variants(graph: Graph, clamps: Sequence, per_group: int) -> list:
    """Variants of one graph in groups of per_group (a power of two), the
    last groups smaller powers of two as the count requires."""
    out, i = [], 0
    while i < len(clamps):
        k = per_group
        while k > len(clamps) - i:
            k //= 2
        out.append(Group(graph, clamps[i:i + k]))
        i += k
    return out


class Batch:
    def __init__(self, graphs_or_groups: Sequence, device: int = 0, **opts):
        groups = [g if isinstance(g, Group) else Group(g) for g in graphs_or_groups]
        self.groups = groups
        self.variant_graph = [g.graph for g in groups for _ in range(g.n_variants)]
        self._csrs = [g.graph.csr_c() for g in groups]
        self._keep = []
        arr = (BatchGroupC * len(groups))()
        for i, (g, c) in enumerate(zip(groups, self._csrs)):
            arr[i].graph = C.cast(C.pointer(c), C.c_void_p)
            arr[i].n_variants = g.n_variants
            if any(len(v) for v, _ in g.clamps):
                off = np.zeros(g.n_variants + 1, np.uint64)
                for k, (v, _) in enumerate(g.clamps):
                    off[k + 1] = off[k] + len(v)
                var = np.ascontiguousarray(np.concatenate([np.asarray(v, np.uint64) for v, _ in g.clamps]), np.uint64)
                val = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint8) for _, x in g.clamps]), np.uint8)
                self._keep += [off, var, val]
                arr[i].clamp_off = off.ctypes.data
                arr[i].clamp_var = var.ctypes.data
                arr[i].clamp_val = val.ctypes.data
        o = BatchOptsC()
        lib().flbp_batch_default_opts(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        self._h = C.c_void_p()
        call("flbp_batch_create_groups", len(groups), arr, device, C.byref(o), C.byref(self._h))

    @property
    def n_variants(self) -> int:
        return int(lib().flbp_batch_n_variants(self._h))

    def close(self):
        if self._h:
            lib().flbp_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, logdomain: bool = False):
        call("flbp_batch_init", self._h, 1 if logdomain else 0)

    def run(self, tol: float, maxiter: int, damping: float = 0.0, maxtime: float = math.inf, chunk: int = 0):
        NV = len(self.variant_graph)
        iters = np.zeros(NV, np.uint64)
        md = np.zeros(NV, np.float64)
        tot = RunResultC()
        o = RunOptsC(tol, maxiter, maxtime, damping, chunk)
        call("flbp_batch_run", self._h, C.byref(o), _ptr(iters, C.c_uint64), _ptr(md, C.c_double), C.byref(tot))
        return iters.astype(np.int64), md, RunResult(int(tot.iters), int(tot.sweeps), float(tot.maxdiff),
                                                      float(tot.seconds))

    def beliefs(self, v: int) -> np.ndarray:
        out = np.empty(self.variant_graph[v].n_vars, np.float64)
        call("flbp_batch_beliefs", self._h, v, _ptr(out, C.c_double), out.shape[0])
        return out

    def beliefs_at(self, v: int, var_indices) -> np.ndarray:
        idx = np.ascontiguousarray(var_indices, np.uint64)
        out = np.empty(idx.shape[0], np.float64)
        call("flbp_batch_beliefs_at", self._h, v, idx.shape[0], _ptr(idx, C.c_uint64), _ptr(out, C.c_double))
        return out

    def set_p(self, group: int, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float64)
        call("flbp_batch_set_p", self._h, group, _ptr(p, C.c_double), p.shape[0])

    def opts(self) -> dict:
        o = BatchOptsC()
        call("flbp_batch_get_opts", self._h, C.byref(o))
        return {f: getattr(o, f) for f, _ in BatchOptsC._fields_}

    def set_opts(self, **kw):
        o = BatchOptsC()
        call("flbp_batch_get_opts", self._h, C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        call("flbp_batch_set_opts", self._h, C.byref(o))

    def size_class(self) -> str:
        """Autotune cache key: powers of two of the K-interleaved message
        slots, of the groups and of the largest variant count."""
        slots = sum(g.graph.n_edges * g.n_variants for g in self.groups)
        kmax = max(g.n_variants for g in self.groups)
        return (f"slots2^{max(slots, 1).bit_length() - 1}-groups2^{len(self.groups).bit_length() - 1}"
                f"-K{kmax}")

    def autotune(self, damping: float, trial_sweeps: int = 8, cache: Optional[str] = None) -> dict:
        """Times the candidate tile bounds and keeps the fastest.  With a cache file, a batch of a size class tuned before
        takes the stored choice without timing, and a new choice is stored."""
        key = self.size_class()
        if cache and os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
            if key in known:
                self.set_opts(**known[key])
                self.reset_counters()
                self.init()
                return self.opts()
        o = BatchOptsC()
        call("flbp_batch_autotune", self._h, damping, trial_sweeps, C.byref(o))
        best = {f: getattr(o, f) for f, _ in BatchOptsC._fields_}
        if cache:
            known = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    known = json.load(f)
            known[key] = {k: best[k] for k in ("tile_factors", "tile_vars")}
            tmp = cache + ".tmp"
            with open(tmp, "w") as f:
                json.dump(known, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return best

    def reset_counters(self):
        call("flbp_batch_reset_counters", self._h)

    @property
    def device_bytes(self) -> int:
        return int(lib().flbp_batch_device_bytes(self._h))
