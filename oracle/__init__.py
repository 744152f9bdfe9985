"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU FP64 restatement of the reference's CausalBP path -- PARALL, SEQFIX
and TOPO updates (see oracle.cpp for
the per-function citations).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module; the engine never does.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DUMP = os.path.join(HERE, "_ref", "ref_dump")

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class OrcFactorList(C.Structure):
    _fields_ = [("n_factors", C.c_uint64), ("type", i32p), ("head", u64p), ("body_off", u64p),
                ("body", u64p), ("p", f64p), ("q", f64p), ("head_clamped", u8p),
                ("mask0", f64p), ("mask1", f64p)]


_SIG = {
    "orc_build": (C.c_int, [C.POINTER(OrcFactorList), C.POINTER(C.c_void_p)]),
    "orc_free": (None, [C.c_void_p]),
    "orc_nvars": (C.c_uint64, [C.c_void_p]),
    "orc_nfactors": (C.c_uint64, [C.c_void_p]),
    "orc_nedges": (C.c_uint64, [C.c_void_p]),
    "orc_dump": (C.c_int, [C.c_void_p, u64p, u64p, u64p, u64p, u64p, u64p, i32p]),
    "orc_init": (C.c_int, [C.c_void_p, C.c_int]),
    "orc_run": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.c_double, C.c_int, u64p, f64p]),
    "orc_run_fixed": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.c_int]),
    "orc_sweep": (C.c_int, [C.c_void_p, C.c_double, C.c_int, f64p]),
    "orc_run_hist": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, f64p, f64p,
                               C.POINTER(C.c_int), f64p]),
    "orc_beliefs": (C.c_int, [C.c_void_p, f64p]),
    "orc_iters": (C.c_uint64, [C.c_void_p]),
    "orc_messages": (C.c_int, [C.c_void_p, f64p, f64p, f64p, f64p]),
    "orc_set_messages": (C.c_int, [C.c_void_p, f64p, f64p, f64p, f64p]),
    "orc_fv_message": (None, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                              C.c_double, C.c_double, f64p, f64p, f64p, f64p]),
    "orc_vf_message": (None, [C.c_uint64, C.c_uint64, f64p, f64p, f64p, f64p]),
    "orc_clamp": (C.c_int, [C.POINTER(OrcFactorList), C.c_uint64, u64p, u8p, C.POINTER(C.c_void_p)]),
    "orc_flist_view": (None, [C.c_void_p, C.POINTER(OrcFactorList)]),
    "orc_flist_free": (None, [C.c_void_p]),
    "orc_grad": (C.c_int, [C.POINTER(OrcFactorList), i32p, C.c_uint64, f64p, C.c_uint64, u64p, f64p,
                           C.c_uint64, C.c_double, f64p, f64p]),
    "orc_loss": (C.c_int, [C.POINTER(OrcFactorList), i32p, C.c_uint64, f64p, C.c_uint64, u64p, f64p,
                           C.c_uint64, C.c_double, f64p]),
    "orc_rng_new": (C.c_void_p, [C.c_uint32]),
    "orc_rng_free": (None, [C.c_void_p]),
    "orc_rng_next": (C.c_uint32, [C.c_void_p]),
    "orc_shuffle": (C.c_int, [C.c_void_p, C.c_uint64, u64p]),
    "orc_std_shuffle": (C.c_int, [C.c_uint32, C.c_uint64, u64p]),
    "orc_order_seqfix": (C.c_int, [C.c_void_p, C.c_void_p, u64p]),
    "orc_order_topo": (C.c_int, [C.c_void_p, u64p]),
    "orc_seq_set": (C.c_int, [C.c_void_p, u64p, C.c_uint64, C.c_int]),
    "orc_seq_clear": (C.c_int, [C.c_void_p]),
    "orc_seq_info": (C.c_int, [C.c_void_p, u64p, u64p, u64p]),
    "orc_seq_waves": (C.c_int, [C.c_void_p, u64p, u64p, u64p, u64p]),
    "orc_products": (C.c_int, [C.c_void_p, f64p, f64p, u64p, u64p]),
    "orc_seq_sweep_items": (C.c_int, [C.c_void_p, C.c_uint64, u8p, u64p, f64p]),
    "orc_seq_sweep_exec": (C.c_int, [C.c_void_p, C.c_uint64, u8p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32), f64p]),
    "orc_fv_message_seq": (None, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                                  C.c_double, C.c_double, f64p, f64p, f64p, f64p]),
    "orc_last_error": (C.c_char_p, []),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
        l = C.CDLL(LIB_PATH)
        for n, (r, a) in _SIG.items():
            f = getattr(l, n)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def _chk(name, rc):
    if rc != 0:
        raise RuntimeError(f"{name}: {lib().orc_last_error().decode()}")


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class _ListHold:
    """Keeps the numpy arrays of a factor list alive while C reads them."""

    def __init__(self, fl):
        self.a = [np.ascontiguousarray(fl.type, np.int32), np.ascontiguousarray(fl.head, np.uint64),
                  np.ascontiguousarray(fl.body_off, np.uint64), np.ascontiguousarray(fl.body, np.uint64),
                  np.ascontiguousarray(fl.p, np.float64), np.ascontiguousarray(fl.q, np.float64),
                  np.ascontiguousarray(fl.head_clamped, np.uint8), np.ascontiguousarray(fl.mask0, np.float64),
                  np.ascontiguousarray(fl.mask1, np.float64)]
        if self.a[3].size == 0:
            self.a[3] = np.zeros(1, np.uint64)
        t, h, o, b, p, q, c, m0, m1 = self.a
        self.c = OrcFactorList(t.shape[0], _p(t, C.c_int32), _p(h, C.c_uint64), _p(o, C.c_uint64),
                               _p(b, C.c_uint64), _p(p, C.c_double), _p(q, C.c_double),
                               _p(c, C.c_uint8), _p(m0, C.c_double), _p(m1, C.c_double))


class OracleGraph:
    def __init__(self, fl):
        self._hold = _ListHold(fl)
        self._h = C.c_void_p()
        _chk("orc_build", lib().orc_build(C.byref(self._hold.c), C.byref(self._h)))
        L = lib()
        self.n_vars = int(L.orc_nvars(self._h))
        self.n_factors = int(L.orc_nfactors(self._h))
        self.n_edges = int(L.orc_nedges(self._h))

    def __del__(self):
        try:
            if self._h:
                lib().orc_free(self._h)
        except Exception:
            pass

    def dump(self):
        V, F, E = self.n_vars, self.n_factors, self.n_edges
        d = dict(var_label=np.zeros(V, np.uint64), fac_off=np.zeros(F + 1, np.uint64),
                 edge_var=np.zeros(E, np.uint64), var_off=np.zeros(V + 1, np.uint64),
                 var_src=np.zeros(E, np.uint64), head_slot=np.zeros(F, np.uint64),
                 code=np.zeros(E, np.int32))
        _chk("orc_dump", lib().orc_dump(
            self._h, _p(d["var_label"], C.c_uint64), _p(d["fac_off"], C.c_uint64),
            _p(d["edge_var"], C.c_uint64), _p(d["var_off"], C.c_uint64), _p(d["var_src"], C.c_uint64),
            _p(d["head_slot"], C.c_uint64), _p(d["code"], C.c_int32)))
        return d

    def init(self, logdomain=False):
        _chk("orc_init", lib().orc_init(self._h, 1 if logdomain else 0))

    def run(self, tol, maxiter, damping=0.0, threads=1):
        it = C.c_uint64()
        md = C.c_double()
        _chk("orc_run", lib().orc_run(self._h, tol, maxiter, damping, threads, C.byref(it), C.byref(md)))
        return int(it.value), float(md.value)

    def run_fixed(self, n, damping=0.0, threads=1):
        _chk("orc_run_fixed", lib().orc_run_fixed(self._h, n, damping, threads))

    def run_hist(self, tol, min_iters, max_iters, hist_len, damping=0.0):
        """The 4-argument run: (yet-to-converge fraction, low-pass beliefs,
        exit reason 0/1/2, last max diff)."""
        yet, md, reason = C.c_double(), C.c_double(), C.c_int()
        low = np.zeros(max(self.n_vars, 1), np.float64)
        _chk("orc_run_hist", lib().orc_run_hist(self._h, tol, min_iters, max_iters, hist_len, damping,
                                               C.byref(yet), _p(low, C.c_double), C.byref(reason), C.byref(md)))
        return float(yet.value), low[:self.n_vars], int(reason.value), float(md.value)

    def sweep(self, damping=0.0, threads=1):
        md = C.c_double()
        _chk("orc_sweep", lib().orc_sweep(self._h, damping, threads, C.byref(md)))
        return float(md.value)

    def beliefs(self):
        out = np.zeros(self.n_vars, np.float64)
        _chk("orc_beliefs", lib().orc_beliefs(self._h, _p(out, C.c_double)))
        return out

    @property
    def iters(self):
        return int(lib().orc_iters(self._h))

    def messages(self):
        E = self.n_edges
        a = [np.zeros(max(E, 1), np.float64) for _ in range(4)]
        _chk("orc_messages", lib().orc_messages(self._h, *[_p(x, C.c_double) for x in a]))
        fv = np.stack([a[0][:E], a[1][:E]], axis=1)
        vf = np.stack([a[2][:E], a[3][:E]], axis=1)
        return fv, vf

    def set_messages(self, fv, vf):
        a = [np.ascontiguousarray(fv[:, 0]), np.ascontiguousarray(fv[:, 1]),
             np.ascontiguousarray(vf[:, 0]), np.ascontiguousarray(vf[:, 1])]
        _chk("orc_set_messages", lib().orc_set_messages(self._h, *[_p(x, C.c_double) for x in a]))


    # ---- sequential schedules (SEQFIX / TOPO) ----
    def order_seqfix(self, rng: "Rng") -> np.ndarray:
        out = np.zeros(max(self.n_edges, 1), np.uint64)
        _chk("orc_order_seqfix", lib().orc_order_seqfix(self._h, rng._h, _p(out, C.c_uint64)))
        return out[:self.n_edges]

    def order_topo(self) -> np.ndarray:
        out = np.zeros(max(self.n_edges, 1), np.uint64)
        _chk("orc_order_topo", lib().orc_order_topo(self._h, _p(out, C.c_uint64)))
        return out[:self.n_edges]

    def set_sequence(self, order, respect_preorder=False):
        o = np.ascontiguousarray(order, np.uint64)
        if o.size == 0:
            o = np.zeros(1, np.uint64)
            n = 0
        else:
            n = o.shape[0]
        _chk("orc_seq_set", lib().orc_seq_set(self._h, _p(o, C.c_uint64), n, 1 if respect_preorder else 0))

    def clear_sequence(self):
        _chk("orc_seq_clear", lib().orc_seq_clear(self._h))

    def waves(self):
        """(fv_off, fv_edge, vf_off, vf_slot) of the wave split."""
        w, a, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _chk("orc_seq_info", lib().orc_seq_info(self._h, C.byref(w), C.byref(a), C.byref(b)))
        W, nf, nv = int(w.value), int(a.value), int(b.value)
        fo = np.zeros(W + 1, np.uint64)
        vo = np.zeros(W + 1, np.uint64)
        fe = np.zeros(max(nf, 1), np.uint64)
        vs = np.zeros(max(nv, 1), np.uint64)
        _chk("orc_seq_waves", lib().orc_seq_waves(self._h, _p(fo, C.c_uint64), _p(fe, C.c_uint64),
                                                 _p(vo, C.c_uint64), _p(vs, C.c_uint64)))
        return fo, fe[:nf], vo, vs[:nv]

    def sweep_items(self, kind, item):
        """One sweep of explicitly ordered items (checker of a level split)."""
        k = np.ascontiguousarray(kind, np.uint8)
        it = np.ascontiguousarray(item, np.uint64)
        md = C.c_double()
        _chk("orc_seq_sweep_items", lib().orc_seq_sweep_items(self._h, k.shape[0], _p(k, C.c_uint8),
                                                             _p(it, C.c_uint64), C.byref(md)))
        return float(md.value)

    def sweep_exec(self, kind, a, b, c):
        """One sweep of engine execution items (checker of the level split)."""
        k = np.ascontiguousarray(kind, np.uint8)
        arrs = [np.ascontiguousarray(x, np.uint32) for x in (a, b, c)]
        md = C.c_double()
        _chk("orc_seq_sweep_exec", lib().orc_seq_sweep_exec(self._h, k.shape[0], _p(k, C.c_uint8),
                                                           *[_p(x, C.c_uint32) for x in arrs], C.byref(md)))
        return float(md.value)

    def products(self):
        V = self.n_vars
        a = [np.zeros(max(V, 1), np.float64) for _ in range(2)] + [np.zeros(max(V, 1), np.uint64) for _ in range(2)]
        _chk("orc_products", lib().orc_products(self._h, _p(a[0], C.c_double), _p(a[1], C.c_double),
                                               _p(a[2], C.c_uint64), _p(a[3], C.c_uint64)))
        return [x[:V] for x in a]


class Rng:
    """A std::mt19937 owned by the oracle (the reference's rnd_gen)."""

    def __init__(self, seed: int = 42):
        self._h = C.c_void_p(lib().orc_rng_new(seed))

    def __del__(self):
        try:
            if self._h:
                lib().orc_rng_free(self._h)
        except Exception:
            pass

    def next(self) -> int:
        return int(lib().orc_rng_next(self._h))

    def shuffle(self, n: int) -> np.ndarray:
        out = np.zeros(max(n, 1), np.uint64)
        _chk("orc_shuffle", lib().orc_shuffle(self._h, n, _p(out, C.c_uint64)))
        return out[:n]


def std_shuffle(seed: int, n: int) -> np.ndarray:
    out = np.zeros(max(n, 1), np.uint64)
    _chk("orc_std_shuffle", lib().orc_std_shuffle(seed, n, _p(out, C.c_uint64)))
    return out[:n]


def fv_message(code, m0, m1, src, head, p, q=0.0, mask0=1.0, mask1=1.0, seq=False):
    m0 = np.ascontiguousarray(m0, np.float64)
    m1 = np.ascontiguousarray(m1, np.float64)
    o0, o1 = C.c_double(), C.c_double()
    f = lib().orc_fv_message_seq if seq else lib().orc_fv_message
    f(code, m0.shape[0], src, head, p, q, mask0, mask1, _p(m0, C.c_double),
      _p(m1, C.c_double), C.byref(o0), C.byref(o1))
    return o0.value, o1.value


def vf_message(m0, m1, src):
    m0 = np.ascontiguousarray(m0, np.float64)
    m1 = np.ascontiguousarray(m1, np.float64)
    o0, o1 = C.c_double(), C.c_double()
    lib().orc_vf_message(m0.shape[0], src, _p(m0, C.c_double), _p(m1, C.c_double), C.byref(o0),
                         C.byref(o1))
    return o0.value, o1.value


def clamp(fl, labels, values, factory):
    """Returns factory(type, head, body_off, body, p, q, head_clamped, mask0, mask1)."""
    hold = _ListHold(fl)
    lab = np.ascontiguousarray(labels, np.uint64)
    val = np.ascontiguousarray(values, np.uint8)
    h = C.c_void_p()
    _chk("orc_clamp", lib().orc_clamp(C.byref(hold.c), lab.shape[0], _p(lab, C.c_uint64),
                                       _p(val, C.c_uint8), C.byref(h)))
    v = OrcFactorList()
    lib().orc_flist_view(h, C.byref(v))
    F = int(v.n_factors)

    def arr(ptr, n, dt):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)

    off = arr(v.body_off, F + 1, np.uint64)
    out = factory(arr(v.type, F, np.int32), arr(v.head, F, np.uint64), off,
                  arr(v.body, int(off[-1]), np.uint64), arr(v.p, F, np.float64), arr(v.q, F, np.float64),
                  arr(v.head_clamped, F, np.uint8), arr(v.mask0, F, np.float64),
                  arr(v.mask1, F, np.float64))
    lib().orc_flist_free(h)
    return out


def loss(fl, rule, n_rules, theta, alarm_labels, targets, T, damping):
    hold = _ListHold(fl)
    rule = np.ascontiguousarray(rule, np.int32)
    theta = np.ascontiguousarray(theta, np.float64)
    al = np.ascontiguousarray(alarm_labels, np.uint64)
    tg = np.ascontiguousarray(targets, np.float64)
    out = C.c_double()
    _chk("orc_loss", lib().orc_loss(C.byref(hold.c), _p(rule, C.c_int32), n_rules, _p(theta, C.c_double),
                                     al.shape[0], _p(al, C.c_uint64), _p(tg, C.c_double), T, damping,
                                     C.byref(out)))
    return out.value


def grad(fl, rule, n_rules, theta, alarm_labels, targets, T, damping):
    hold = _ListHold(fl)
    rule = np.ascontiguousarray(rule, np.int32)
    theta = np.ascontiguousarray(theta, np.float64)
    al = np.ascontiguousarray(alarm_labels, np.uint64)
    tg = np.ascontiguousarray(targets, np.float64)
    out = C.c_double()
    g = np.zeros(n_rules, np.float64)
    _chk("orc_grad", lib().orc_grad(C.byref(hold.c), _p(rule, C.c_int32), n_rules, _p(theta, C.c_double),
                                     al.shape[0], _p(al, C.c_uint64), _p(tg, C.c_double), T, damping,
                                     C.byref(out), _p(g, C.c_double)))
    return out.value, g


def ref_dump(path, observations=None):
    """Runs oracle/_ref/ref_dump (the reference's own graph model) and parses
    its output.  Returns None when it has not been built."""
    if not os.path.exists(REF_DUMP):
        return None
    args = [REF_DUMP, path] + ([observations] if observations else [])
    txt = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    return parse_dump(txt)


def parse_dump(txt):
    lines = txt.strip("\n").split("\n")
    V, F, E = map(int, lines[0].split())
    labels = [int(x) for x in lines[1:1 + V]]
    facs = []
    for ln in lines[1 + V:1 + V + F]:
        t = ln.split()
        nb = [tuple(map(int, x.split(":"))) for x in t[8:]]
        facs.append(dict(type=int(t[0]), p=float(t[1]), q=float(t[2]), clamped=int(t[3]),
                         mask0=float(t[4]), mask1=float(t[5]), head=int(t[6]), nbF=nb))
    nbv = []
    for ln in lines[1 + V + F:1 + V + F + V]:
        t = ln.split()
        nbv.append([tuple(map(int, x.split(":"))) for x in t[1:]])
    return dict(V=V, F=F, E=E, labels=labels, factors=facs, nbV=nbv)
