"""Clamped feedback variants of one graph, solved and differentiated together.

A variant of a graph clamps some alarm variables without body uses to 0 or
1 (examples/fl_wrapper.cu:26-45: clamp, then rebuild; for such variables
the rebuild keeps the structure and only the factors they head change,
causal_factor.cpp:75-83).  K variants become one graph of K disjoint
copies, interleaved:

  - factor f of copy k sits at position f*K + k, with its slots in the
    base factor's order;
  - the variable of label l in copy k has label l*K + k, so the copies of a
    variable are adjacent in the label-ordered variable list, and each
    copy's edge list keeps the base variable's order.

Each copy's arithmetic is then that of its variant solved alone, to the
bit.  One plan sweeps all K copies per launch (one launch per pass for the
K variants instead of K), the K copies of a factor are neighbours in the
factor-major message array and the K copies of a variable neighbours in
the variable arrays, so the gathers of K variants share lines.  Variant
k's clamps are applied in place on copy k (flbp_plan_clamp_vars), so a
batch of variants is re-clamped without a rebuild.  One adjoint
(flbp_adjoint_*) differentiates the K copies' T-sweep solves together,
and its tied-rule reduction sums dL/dtheta over the copies.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from .adjoint import Adjoint, tied_p
from .engine import FactorList, Graph, Plan


def interleave(flist: FactorList, K: int) -> FactorList:
    """K interleaved copies of a factor list (factor f of copy k at f*K + k,
    label l of copy k as l*K + k)."""
    if K < 1:
        raise ValueError("K must be positive")
    F = flist.n_factors
    k = np.arange(K, dtype=np.uint64)
    deg = np.diff(flist.body_off.astype(np.int64))
    udeg = np.repeat(deg, K)
    uoff = np.zeros(F * K + 1, dtype=np.uint64)
    np.cumsum(udeg, out=uoff[1:])
    # element j of union factor u = f*K + c: base element body_off[f] + j, label * K + c
    u = np.repeat(np.arange(F * K, dtype=np.int64), udeg)
    j = np.arange(int(uoff[-1]), dtype=np.int64) - uoff[u].astype(np.int64)
    f, c = u // K, (u % K).astype(np.uint64)
    body = flist.body[flist.body_off[f].astype(np.int64) + j] * np.uint64(K) + c
    head = (flist.head[:, None] * np.uint64(K) + k[None, :]).ravel()
    rep = lambda a: np.repeat(a, K)
    return FactorList(rep(flist.type), head, uoff, body, rep(flist.p), rep(flist.q), rep(flist.head_clamped),
                      rep(flist.mask0), rep(flist.mask1))


class VariantBatch:
    """K clamped variants of one graph with tied rule logits, resident on one
    GPU: the T-sweep damped solve from init of all K (one plan over the
    interleaved copies) and its adjoint."""

    def __init__(self, flist: FactorList, rule: np.ndarray, n_rules: int, alarm_labels, K: int,
                 n_sweeps: int, checkpoint_every: int, damping: float, device: int = 0):
        self.K = K
        self.n_rules = n_rules
        self.graph = Graph(interleave(flist, K))
        self.base_p = self.graph.p.copy()
        self.rule = np.ascontiguousarray(np.repeat(np.asarray(rule, np.int32), K), np.int32)
        alarm_labels = np.asarray(alarm_labels, np.uint64)
        # alarm_idx[c] = union variable indices of the alarms in copy c
        self.alarm_idx = np.stack([self.graph.index_of(alarm_labels * np.uint64(K) + np.uint64(c))
                                   for c in range(K)])
        self.plan = Plan(self.graph, device)
        self.adj = Adjoint(self.plan, n_sweeps, checkpoint_every, damping)
        self.adj.set_rules(self.rule, n_rules)
        self.p = self.base_p

    def close(self):
        self.adj.close()
        self.plan.close()

    def set_clamps(self, clamps: Sequence) -> None:
        """clamps[c] = (alarm labels, 0/1 values) of the variant in copy c
        (K entries; labels of the base graph, variables without body uses)."""
        if len(clamps) != self.K:
            raise ValueError(f"{len(clamps)} variants for {self.K} copies")
        self.plan.reset_head_clamps()
        for c, (labels, values) in enumerate(clamps):
            labels = np.asarray(labels, np.uint64)
            if labels.size:
                idx = self.graph.index_of(labels * np.uint64(self.K) + np.uint64(c))
                self.plan.clamp_vars(idx, np.asarray(values, np.uint8))

    def set_theta(self, theta: np.ndarray) -> None:
        self.p = tied_p(self.base_p, self.rule, np.asarray(theta, np.float64))
        self.plan.set_p(self.p)

    def forward(self) -> np.ndarray:
        """T sweeps from init of every copy; the alarm marginals [K, n_alarms]."""
        self.adj.forward(download=False)
        flat = self.plan.beliefs_at(self.alarm_idx.ravel())
        return flat.reshape(self.alarm_idx.shape)

    def clear_grad(self) -> None:
        self.adj.clear_grad()

    def backward(self, dl_db: np.ndarray) -> None:
        """Adds dL/dp for dL/db[c, a] at alarm a of copy c to the device
        accumulator (the adjoint of the last forward)."""
        dl_db = np.asarray(dl_db, np.float64)
        if dl_db.shape != self.alarm_idx.shape:
            raise ValueError("dl_db must be [K, n_alarms]")
        self.adj.backward_sparse(self.alarm_idx.ravel(), dl_db.ravel())

    def rule_grad(self) -> np.ndarray:
        """dL/dtheta of everything accumulated since clear_grad, summed over
        the copies (device reduction)."""
        return self.adj.rule_grad()

    def last_timing(self):
        return self.adj.last_timing()

    @property
    def device_bytes(self) -> int:
        return self.plan.device_bytes + self.adj.device_bytes
