"""Python verification harness over libfastlbp.so.

Thin numpy wrappers around the C ABI (include/fastlbp.h): factor lists,
graphs in the reference's CSR orders, the synthetic generator, clamping,
and the GPU plan.  Every call goes through the native library; nothing here
computes BP itself.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._native import (UPDATES_PARALL, UPDATES_SEQFIX, UPDATES_TOPO, CsrC, FactorListC, GenParamsC, HistResultC,
                      RunOptsC, RunResultC, ScheduleInfoC, call, check, f64p, i32p, lib, u8p, u64p)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _arr(ptr, n: int, dtype) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


@dataclass
class FactorList:
    """Flat std::vector<CausalFactor> (include/ext/causal_factor.h)."""

    type: np.ndarray          # int32 [F]
    head: np.ndarray          # uint64 [F]
    body_off: np.ndarray      # uint64 [F+1]
    body: np.ndarray          # uint64 [body_off[-1]]
    p: np.ndarray             # float64 [F]
    q: np.ndarray             # float64 [F]
    head_clamped: np.ndarray  # uint8 [F]
    mask0: np.ndarray         # float64 [F]
    mask1: np.ndarray         # float64 [F]

    def __post_init__(self):
        self.type = np.ascontiguousarray(self.type, dtype=np.int32)
        self.head = np.ascontiguousarray(self.head, dtype=np.uint64)
        self.body_off = np.ascontiguousarray(self.body_off, dtype=np.uint64)
        self.body = np.ascontiguousarray(self.body, dtype=np.uint64)
        self.p = np.ascontiguousarray(self.p, dtype=np.float64)
        self.q = np.ascontiguousarray(self.q, dtype=np.float64)
        self.head_clamped = np.ascontiguousarray(self.head_clamped, dtype=np.uint8)
        self.mask0 = np.ascontiguousarray(self.mask0, dtype=np.float64)
        self.mask1 = np.ascontiguousarray(self.mask1, dtype=np.float64)
        if self.body.size == 0:
            self.body = np.zeros(1, dtype=np.uint64)[:0]

    @property
    def n_factors(self) -> int:
        return int(self.type.shape[0])

    @staticmethod
    def build(factors: Sequence[tuple]) -> "FactorList":
        """From (type, head, body, p[, q[, clamped, mask0, mask1]]) tuples."""
        t, h, off, body, p, q, cl, m0, m1 = [], [], [0], [], [], [], [], [], []
        for f in factors:
            ty, hd, bd, pp = f[0], f[1], list(f[2]), f[3]
            qq = f[4] if len(f) > 4 else 0.0
            c, a, b = (f[5], f[6], f[7]) if len(f) > 5 else (0, 1.0, 1.0)
            t.append(ty); h.append(hd); body.extend(bd); off.append(len(body))
            p.append(pp); q.append(qq); cl.append(c); m0.append(a); m1.append(b)
        return FactorList(np.array(t), np.array(h, dtype=np.uint64), np.array(off, dtype=np.uint64),
                          np.array(body, dtype=np.uint64), np.array(p), np.array(q),
                          np.array(cl, dtype=np.uint8), np.array(m0), np.array(m1))

    def as_c(self) -> FactorListC:
        return FactorListC(self.n_factors, _ptr(self.type, C.c_int32), _ptr(self.head, C.c_uint64),
                           _ptr(self.body_off, C.c_uint64), _ptr(self.body, C.c_uint64),
                           _ptr(self.p, C.c_double), _ptr(self.q, C.c_double),
                           _ptr(self.head_clamped, C.c_uint8), _ptr(self.mask0, C.c_double),
                           _ptr(self.mask1, C.c_double))

    @staticmethod
    def from_c(v: FactorListC) -> "FactorList":
        F = int(v.n_factors)
        off = _arr(v.body_off, F + 1, np.uint64)
        nb = int(off[-1]) if F else 0
        return FactorList(_arr(v.type, F, np.int32), _arr(v.head, F, np.uint64), off,
                          _arr(v.body, nb, np.uint64), _arr(v.p, F, np.float64),
                          _arr(v.q, F, np.float64), _arr(v.head_clamped, F, np.uint8),
                          _arr(v.mask0, F, np.float64), _arr(v.mask1, F, np.float64))

    def with_p(self, p: np.ndarray) -> "FactorList":
        return FactorList(self.type, self.head, self.body_off, self.body, np.asarray(p, np.float64),
                          self.q, self.head_clamped, self.mask0, self.mask1)


def _take_flist(handle: C.c_void_p) -> FactorList:
    v = FactorListC()
    call("flbp_flist_view", handle, C.byref(v))
    try:
        return FactorList.from_c(v)
    finally:
        lib().flbp_flist_free(handle)


def read_causal_fg(path: str) -> FactorList:
    h = C.c_void_p()
    call("flbp_read_causal_fg", path.encode(), C.byref(h))
    return _take_flist(h)


def write_causal_fg(fl: FactorList, path: str) -> None:
    c = fl.as_c()
    call("flbp_write_causal_fg", C.byref(c), path.encode())


def clamp(fl: FactorList, labels: Sequence[int], values: Sequence[int]) -> FactorList:
    """CausalFactorGraph::clamp per (label, value), in ascending label order."""
    lab = np.ascontiguousarray(labels, dtype=np.uint64)
    val = np.ascontiguousarray(values, dtype=np.uint8)
    c = fl.as_c()
    h = C.c_void_p()
    call("flbp_clamp", C.byref(c), len(lab), _ptr(lab, C.c_uint64), _ptr(val, C.c_uint8), C.byref(h))
    return _take_flist(h)


@dataclass
class Generated:
    factors: FactorList
    alarms: np.ndarray   # uint64 labels
    rule: np.ndarray     # int32 per factor, -1 = untied


def gen_params(**kw) -> GenParamsC:
    prm = GenParamsC()
    lib().flbp_gen_defaults(C.byref(prm))
    for k, v in kw.items():
        if not hasattr(prm, k):
            raise TypeError(f"unknown generator parameter {k}")
        setattr(prm, k, v)
    return prm


def generate(n_vars: int, seed: int, **kw) -> Generated:
    prm = gen_params(n_vars=n_vars, seed=seed, **kw)
    h = C.c_void_p()
    call("flbp_generate", C.byref(prm), C.byref(h))
    try:
        v = FactorListC()
        call("flbp_generated_view", h, C.byref(v))
        fl = FactorList.from_c(v)
        na = C.c_uint64()
        ap = u64p()
        rp = i32p()
        call("flbp_generated_extra", h, C.byref(na), C.byref(ap), C.byref(rp))
        alarms = _arr(ap, int(na.value), np.uint64)
        rule = _arr(rp, fl.n_factors, np.int32)
    finally:
        lib().flbp_generated_free(h)
    return Generated(fl, alarms, rule)


class Graph:
    """A factor graph in the reference's CSR orders (flbp_graph_build)."""

    def __init__(self, fl: Optional[FactorList]):
        if fl is None:
            return
        h = C.c_void_p()
        c = fl.as_c()
        call("flbp_graph_build", C.byref(c), C.byref(h))
        self._take(h)

    @classmethod
    def read(cls, path: str) -> "Graph":
        """A graph from a binary graph file (flbp_read_graph)."""
        h = C.c_void_p()
        call("flbp_read_graph", path.encode(), C.byref(h))
        g = cls(None)
        g._take(h)
        return g

    def write(self, path: str) -> None:
        """Writes the binary graph file (flbp_write_graph)."""
        c = self.csr_c()
        call("flbp_write_graph", C.byref(c), path.encode())

    def _take(self, h) -> None:
        """Copies the CSR of a native graph handle and frees the handle."""
        try:
            v = CsrC()
            call("flbp_graph_csr", h, C.byref(v))
            V, F, E = int(v.n_vars), int(v.n_factors), int(v.n_edges)
            self.n_vars, self.n_factors, self.n_edges = V, F, E
            self.var_label = _arr(v.var_label, V, np.uint64)
            self.fac_off = _arr(v.fac_off, F + 1, np.uint64)
            self.edge_var = _arr(v.edge_var, E, np.uint64)
            self.var_off = _arr(v.var_off, V + 1, np.uint64)
            self.var_edge = _arr(v.var_edge, E, np.uint64)
            self.fac_type = _arr(v.fac_type, F, np.int32)
            self.head_slot = _arr(v.head_slot, F, np.uint64)
            self.head_clamped = _arr(v.head_clamped, F, np.uint8)
            self.p = _arr(v.p, F, np.float64)
            self.q = _arr(v.q, F, np.float64)
            self.mask0 = _arr(v.mask0, F, np.float64)
            self.mask1 = _arr(v.mask1, F, np.float64)
        finally:
            lib().flbp_graph_free(h)

    def csr_c(self) -> CsrC:
        return CsrC(self.n_vars, self.n_factors, self.n_edges,
                    _ptr(self.var_label, C.c_uint64), _ptr(self.fac_off, C.c_uint64),
                    _ptr(self.edge_var, C.c_uint64), _ptr(self.var_off, C.c_uint64),
                    _ptr(self.var_edge, C.c_uint64), _ptr(self.fac_type, C.c_int32),
                    _ptr(self.head_slot, C.c_uint64), _ptr(self.head_clamped, C.c_uint8),
                    _ptr(self.p, C.c_double), _ptr(self.q, C.c_double),
                    _ptr(self.mask0, C.c_double), _ptr(self.mask1, C.c_double))

    def head_factors(self, var_indices) -> np.ndarray:
        """The first factor (lowest index) whose head each variable is, -1 for
        a variable that heads none.  A variable may head several factors;
        Plan.clamp_vars clamps all of them."""
        head_var = self.edge_var[self.head_slot].astype(np.int64)
        out = np.full(self.n_vars, -1, dtype=np.int64)
        out[head_var[::-1]] = np.arange(self.n_factors)[::-1]
        return out[np.asarray(var_indices, dtype=np.int64)]

    def n_head_factors(self, var_indices) -> np.ndarray:
        """Number of factors whose head each variable is."""
        head_var = self.edge_var[self.head_slot].astype(np.int64)
        cnt = np.bincount(head_var, minlength=self.n_vars)
        return cnt[np.asarray(var_indices, dtype=np.int64)]

    def body_uses(self, var_indices) -> np.ndarray:
        """Factors in which each variable appears other than as the head."""
        v = np.asarray(var_indices, dtype=np.int64)
        deg = (self.var_off[v + 1] - self.var_off[v]).astype(np.int64)
        return deg - self.n_head_factors(v)

    def index_of(self, labels) -> np.ndarray:
        labels = np.asarray(labels, dtype=np.uint64)
        idx = np.searchsorted(self.var_label, labels)
        if np.any(idx >= self.n_vars) or np.any(self.var_label[np.minimum(idx, self.n_vars - 1)] != labels):
            raise KeyError("label not in graph")
        return idx


@dataclass
class ScheduleInfo:
    n_order: int
    n_waves: int
    n_fv_items: int
    n_vf_items: int
    n_levels: int
    updates: int
    respect_preorder: bool
    n_exec_items: int


class Schedule:
    """A sequential update order split into the reference's waves
    (flbp_schedule_*); built on the host, no GPU needed.

    ``Schedule(graph, "SEQFIX")`` shuffles the factors with the calling
    thread's std::mt19937 (seeded 42 per thread, advanced by every SEQFIX
    schedule; see :func:`seed`), ``Schedule(graph, "TOPO")`` takes the
    breadth-first preorder, and ``Schedule(graph, order=...)`` splits an
    explicit order of var-major edge positions."""

    def __init__(self, graph: Graph, updates: str = "SEQFIX", order=None, respect_preorder: bool = False):
        self.graph = graph
        self._h = C.c_void_p()
        c = graph.csr_c()
        if order is not None:
            o = np.ascontiguousarray(order, dtype=np.uint64)
            buf = o if o.size else np.zeros(1, np.uint64)
            call("flbp_schedule_from_order", C.byref(c), _ptr(buf, C.c_uint64), o.shape[0],
                 1 if respect_preorder else 0, C.byref(self._h))
        else:
            code = {"SEQFIX": UPDATES_SEQFIX, "TOPO": UPDATES_TOPO}.get(updates)
            if code is None:
                raise ValueError("updates must be 'SEQFIX' or 'TOPO' (a plan without schedule is PARALL)")
            call("flbp_schedule_create", C.byref(c), code, C.byref(self._h))

    def close(self):
        if self._h:
            lib().flbp_schedule_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> ScheduleInfo:
        i = ScheduleInfoC()
        call("flbp_schedule_get_info", self._h, C.byref(i))
        return ScheduleInfo(int(i.n_order), int(i.n_waves), int(i.n_fv_items), int(i.n_vf_items),
                            int(i.n_levels), int(i.updates), bool(i.respect_preorder), int(i.n_exec_items))

    def order(self) -> np.ndarray:
        n = self.info().n_order
        out = np.zeros(max(n, 1), np.uint64)
        call("flbp_schedule_order", self._h, _ptr(out, C.c_uint64))
        return out[:n]

    def waves(self):
        """(fv_off, fv_edge, vf_off, vf_slot): wave w updates the var-major
        edges fv_edge[fv_off[w]:fv_off[w+1]] after recomputing the
        factor-major v->f slots vf_slot[vf_off[w]:vf_off[w+1]]."""
        i = self.info()
        fo = np.zeros(i.n_waves + 1, np.uint64)
        vo = np.zeros(i.n_waves + 1, np.uint64)
        fe = np.zeros(max(i.n_fv_items, 1), np.uint64)
        vs = np.zeros(max(i.n_vf_items, 1), np.uint64)
        call("flbp_schedule_waves", self._h, _ptr(fo, C.c_uint64), _ptr(fe, C.c_uint64), _ptr(vo, C.c_uint64),
             _ptr(vs, C.c_uint64))
        return fo, fe[:i.n_fv_items], vo, vs[:i.n_vf_items]


    def levels(self):
        """(fv_level, vf_level): the dependency level of each item of waves()."""
        i = self.info()
        a = np.zeros(max(i.n_fv_items, 1), np.uint32)
        b = np.zeros(max(i.n_vf_items, 1), np.uint32)
        call("flbp_schedule_levels", self._h, _ptr(a, C.c_uint32), _ptr(b, C.c_uint32))
        return a[:i.n_fv_items], b[:i.n_vf_items]


    def exec_items(self):
        """(kind, a, b, c, level) of what runs per sweep on the device."""
        n = self.info().n_exec_items
        k = np.zeros(max(n, 1), np.uint8)
        arrs = [np.zeros(max(n, 1), np.uint32) for _ in range(4)]
        call("flbp_schedule_exec", self._h, _ptr(k, C.c_uint8), *[_ptr(x, C.c_uint32) for x in arrs])
        return (k[:n],) + tuple(x[:n] for x in arrs)


def seed(value: int) -> None:
    """Reseed the calling thread's SEQFIX generator (rnd_gen.seed)."""
    lib().flbp_seed(int(value) & 0xffffffff)


@dataclass
class RunResult:
    iters: int
    sweeps: int
    maxdiff: float
    seconds: float


class Plan:
    """A graph resident on one GPU (flbp_plan_*)."""

    def __init__(self, graph: Graph, device: int = 0):
        self.graph = graph
        self._h = C.c_void_p()
        c = graph.csr_c()
        call("flbp_plan_create", C.byref(c), device, C.byref(self._h))

    def close(self):
        if self._h:
            lib().flbp_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, logdomain: bool = False) -> None:
        call("flbp_plan_init", self._h, 1 if logdomain else 0)

    def reset(self) -> None:
        """Iteration counter and old beliefs back to a fresh plan's."""
        call("flbp_plan_reset", self._h)

    def run(self, tol: float, maxiter: int, damping: float = 0.0, maxtime: float = math.inf,
            chunk: int = 0) -> RunResult:
        o = RunOptsC(tol, maxiter, maxtime, damping, chunk)
        r = RunResultC()
        call("flbp_plan_run", self._h, C.byref(o), C.byref(r))
        return RunResult(int(r.iters), int(r.sweeps), float(r.maxdiff), float(r.seconds))

    def run_fixed(self, n: int, damping: float = 0.0) -> RunResult:
        r = RunResultC()
        call("flbp_plan_run_fixed", self._h, n, damping, C.byref(r))
        return RunResult(int(r.iters), int(r.sweeps), float(r.maxdiff), float(r.seconds))

    def run_hist(self, tol: float, min_iters: int, max_iters: int, hist_len: int, maxtime: float = math.inf,
                 damping: float = 0.0):
        """CausalBP::run(tol, minIters, maxIters, histLength): (yet-to-converge
        fraction, low-pass beliefs, reason 0/1/2, last max diff, sweeps)."""
        r = HistResultC()
        low = np.empty(max(self.graph.n_vars, 1), dtype=np.float64)
        call("flbp_plan_run_hist", self._h, tol, min_iters, max_iters, hist_len, maxtime, damping, C.byref(r),
             _ptr(low, C.c_double))
        return float(r.yet_to_converge), low[:self.graph.n_vars], int(r.reason), float(r.maxdiff), int(r.iters)

    def beliefs(self) -> np.ndarray:
        out = np.empty(self.graph.n_vars, dtype=np.float64)
        call("flbp_plan_beliefs", self._h, _ptr(out, C.c_double), out.shape[0])
        return out

    def messages(self):
        E = self.graph.n_edges
        fv = np.empty(2 * max(E, 1), dtype=np.float64)
        vf = np.empty(2 * max(E, 1), dtype=np.float64)
        call("flbp_plan_get_messages", self._h, _ptr(fv, C.c_double), _ptr(vf, C.c_double))
        return fv[:2 * E].reshape(E, 2), vf[:2 * E].reshape(E, 2)

    def set_messages(self, fv_varmajor: Optional[np.ndarray], vf_facmajor: Optional[np.ndarray]):
        a = None if fv_varmajor is None else np.ascontiguousarray(fv_varmajor, dtype=np.float64).ravel()
        b = None if vf_facmajor is None else np.ascontiguousarray(vf_facmajor, dtype=np.float64).ravel()
        call("flbp_plan_set_messages", self._h, None if a is None else _ptr(a, C.c_double),
             None if b is None else _ptr(b, C.c_double))

    def set_p(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float64)
        call("flbp_plan_set_p", self._h, _ptr(p, C.c_double), p.shape[0])

    def beliefs_at(self, var_indices) -> np.ndarray:
        """Beliefs of the given variables, gathered on the device."""
        v = np.ascontiguousarray(var_indices, dtype=np.uint64)
        out = np.empty(v.shape[0], np.float64)
        call("flbp_plan_beliefs_at", self._h, v.shape[0], _ptr(v, C.c_uint64), _ptr(out, C.c_double))
        return out

    def set_head_clamps(self, factors, values) -> None:
        """Head clamps in place (gen_clamped's head rule) on AND/OR factors;
        equals rebuilding the graph clamped at variables with no body uses."""
        f = np.ascontiguousarray(factors, dtype=np.uint64)
        x = np.ascontiguousarray(values, dtype=np.uint8)
        if f.shape != x.shape:
            raise ValueError("factors and values differ in length")
        call("flbp_plan_set_head_clamps", self._h, f.shape[0], _ptr(f, C.c_uint64), _ptr(x, C.c_uint8))

    def clamp_vars(self, var_indices, values) -> None:
        """CausalFactorGraph::clamp in place for variables without body uses:
        every AND/OR factor they head gets the head clamp (equals the rebuilt
        clamped graph)."""
        v = np.ascontiguousarray(var_indices, dtype=np.uint64)
        x = np.ascontiguousarray(values, dtype=np.uint8)
        if v.shape != x.shape:
            raise ValueError("variables and values differ in length")
        call("flbp_plan_clamp_vars", self._h, v.shape[0], _ptr(v, C.c_uint64), _ptr(x, C.c_uint8))

    def reset_head_clamps(self) -> None:
        call("flbp_plan_reset_head_clamps", self._h)

    @property
    def variant(self) -> int:
        """The kernel variant in use (a new plan starts with 4)."""
        v = int(lib().flbp_plan_get_variant(self._h))
        if v < 0:
            check("flbp_plan_get_variant", v)
        return v

    def set_variant(self, variant: int) -> None:
        """Kernel shape and message layout: 4 the state layout (the default:
        f->v factor-major plus per-variable products, v->f derived in the
        factor pass); 0 LDS-staged, v->f var-major and f->v factor-major; 1
        one lane per factor/variable, same layout; 2 LDS-staged, both
        factor-major (push); 3 LDS-staged, v->f factor-major and f->v
        var-major (both passes scatter).  Results are identical; switching
        layouts moves the stored messages."""
        call("flbp_plan_set_variant", self._h, variant)

    def set_schedule(self, schedule: Optional[Schedule]) -> None:
        """Run sweeps by a sequential schedule (None: PARALL).  The plan keeps
        a copy; products and zero counts start at what init() leaves."""
        call("flbp_plan_set_schedule", self._h, None if schedule is None else schedule._h)

    @property
    def updates(self) -> int:
        """FLBP_UPDATES_PARALL without a schedule, else the schedule's."""
        return int(lib().flbp_plan_get_updates(self._h))

    def set_seq_tail(self, max_items: int) -> None:
        """Under a schedule, the trailing levels of at most max_items items
        each (0 .. 1024) run in one single-workgroup launch; 0 launches
        every level.  Results are the same either way."""
        call("flbp_plan_set_seq_tail", self._h, C.c_uint32(max_items))

    def use_graphs(self, enable: bool = True) -> None:
        """Replay chunks of sweeps from captured hipGraphs (off by default:
        no measured gain over direct launches)."""
        call("flbp_plan_use_graphs", self._h, 1 if enable else 0)

    def profile(self, enable: bool = True) -> None:
        call("flbp_plan_profile", self._h, 1 if enable else 0)

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        na, nb = C.c_uint64(), C.c_uint64()
        call("flbp_plan_last_timing", self._h, C.byref(a), C.byref(b), C.byref(na), C.byref(nb))
        return float(a.value), float(b.value), int(na.value), int(nb.value)

    @property
    def device_bytes(self) -> int:
        return int(lib().flbp_plan_device_bytes(self._h))


def rank_alarms(labels, marginals) -> np.ndarray:
    """Alarm indices by marginal descending, ties by label ascending."""
    labels = np.ascontiguousarray(labels, dtype=np.uint64)
    marg = np.ascontiguousarray(marginals, dtype=np.float64)
    out = np.empty(labels.shape[0], dtype=np.uint64)
    call("flbp_rank_alarms", labels.shape[0], _ptr(labels, C.c_uint64), _ptr(marg, C.c_double),
         _ptr(out, C.c_uint64))
    return out.astype(np.int64)


def algorithmic_bytes_per_sweep(n_vars: int, n_factors: int, n_edges: int, damping: float) -> int:
    """HBM bytes one sweep must move at minimum: 68 E + 24 F + 20 V
    (+16 E when damped).

    Per edge: read v->f pair (16), write f->v pair (16), gather f->v pair
    (16), scatter v->f pair (16), 4-byte slot map.  Per factor: offset (4),
    meta word (4), p and q (16).  Per variable: offset (4), old belief read
    and new belief write (16).  Damping reads the previous f->v pair (16).
    """
    b = 68 * n_edges + 24 * n_factors + 20 * n_vars
    if damping != 0.0:
        b += 16 * n_edges
    return b
