"""Batched solve of many graphs on one GPU (flbp_batch_*)."""
from __future__ import annotations

import ctypes as C
import math
from typing import Sequence

import numpy as np

from ._native import BatchOptsC, CsrC, RunOptsC, RunResultC, call, lib
from .engine import Graph, RunResult, _ptr


class Batch:
    def __init__(self, graphs: Sequence[Graph], device: int = 0, **opts):
        self.graphs = list(graphs)
        self._csrs = [g.csr_c() for g in self.graphs]
        arr = (C.POINTER(CsrC) * len(self._csrs))(*[C.pointer(c) for c in self._csrs])
        o = BatchOptsC()
        lib().flbp_batch_default_opts(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        self._h = C.c_void_p()
        call("flbp_batch_create", len(self._csrs), arr, device, C.byref(o), C.byref(self._h))

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
        G = len(self.graphs)
        iters = np.zeros(G, np.uint64)
        md = np.zeros(G, np.float64)
        tot = RunResultC()
        o = RunOptsC(tol, maxiter, maxtime, damping, chunk)
        call("flbp_batch_run", self._h, C.byref(o), _ptr(iters, C.c_uint64), _ptr(md, C.c_double), C.byref(tot))
        return iters.astype(np.int64), md, RunResult(int(tot.iters), int(tot.sweeps), float(tot.maxdiff),
                                                      float(tot.seconds))

    def beliefs(self, g: int) -> np.ndarray:
        out = np.empty(self.graphs[g].n_vars, np.float64)
        call("flbp_batch_beliefs", self._h, g, _ptr(out, C.c_double), out.shape[0])
        return out

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

    def autotune(self, damping: float, trial_sweeps: int = 8) -> dict:
        o = BatchOptsC()
        call("flbp_batch_autotune", self._h, damping, trial_sweeps, C.byref(o))
        return {f: getattr(o, f) for f, _ in BatchOptsC._fields_}

    def reset_counters(self):
        call("flbp_batch_reset_counters", self._h)

    @property
    def device_bytes(self) -> int:
        return int(lib().flbp_batch_device_bytes(self._h))
