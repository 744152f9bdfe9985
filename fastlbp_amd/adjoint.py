"""Gradients of alarm marginals with respect to tied rule logits.

Adjoint wraps the C-ABI reverse-mode entry points (flbp_adjoint_*): a
forward of T damped sweeps from init with checkpoints, and a backward that
recomputes each segment and runs the hand-written reverse kernels.
TiedMarginals is the torch.autograd.Function of the verification harness:
theta (rule logits) -> alarm marginals after T sweeps, with
p_I = sigmoid(theta[rule(I)]) for every tied AND factor.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import call, lib
from .engine import Graph, Plan, _ptr


def tied_p(base_p: np.ndarray, rule: np.ndarray, theta: np.ndarray) -> np.ndarray:
    base_p = np.ascontiguousarray(base_p, np.float64)
    rule = np.ascontiguousarray(rule, np.int32)
    theta = np.ascontiguousarray(theta, np.float64)
    out = np.empty_like(base_p)
    call("flbp_tied_p", base_p.shape[0], _ptr(base_p, C.c_double), _ptr(rule, C.c_int32),
         theta.shape[0], _ptr(theta, C.c_double), _ptr(out, C.c_double))
    return out


def rule_grad(dl_dp: np.ndarray, p: np.ndarray, rule: np.ndarray, n_rules: int) -> np.ndarray:
    dl_dp = np.ascontiguousarray(dl_dp, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    rule = np.ascontiguousarray(rule, np.int32)
    out = np.zeros(n_rules, np.float64)
    call("flbp_rule_grad", dl_dp.shape[0], _ptr(dl_dp, C.c_double), _ptr(p, C.c_double),
         _ptr(rule, C.c_int32), n_rules, _ptr(out, C.c_double))
    return out


class Adjoint:
    """T damped sweeps from init on a plan, with their reverse pass."""

    def __init__(self, plan: Plan, n_sweeps: int, checkpoint_every: int, damping: float):
        self.plan = plan
        self._h = C.c_void_p()
        call("flbp_adjoint_create", plan._h, n_sweeps, checkpoint_every, damping, C.byref(self._h))

    def close(self):
        if self._h:
            lib().flbp_adjoint_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, download: bool = True):
        """T sweeps from init; all beliefs, or None (read selected ones with
        plan.beliefs_at) when download is False."""
        if not download:
            call("flbp_adjoint_forward", self._h, None)
            return None
        out = np.empty(self.plan.graph.n_vars, np.float64)
        call("flbp_adjoint_forward", self._h, _ptr(out, C.c_double))
        return out

    def backward(self, dl_db: np.ndarray) -> np.ndarray:
        dl_db = np.ascontiguousarray(dl_db, np.float64)
        out = np.empty(self.plan.graph.n_factors, np.float64)
        call("flbp_adjoint_backward", self._h, _ptr(dl_db, C.c_double), _ptr(out, C.c_double))
        return out

    def backward_sparse(self, var_indices, dl_db) -> None:
        """Adds dL/dp for dL/db = dl_db at var_indices (0 elsewhere) to the
        device accumulator."""
        v = np.ascontiguousarray(var_indices, np.uint64)
        g = np.ascontiguousarray(dl_db, np.float64)
        if v.shape != g.shape:
            raise ValueError("var_indices and dl_db differ in length")
        call("flbp_adjoint_backward_sparse", self._h, v.shape[0], _ptr(v, C.c_uint64), _ptr(g, C.c_double))

    def clear_grad(self) -> None:
        call("flbp_adjoint_clear_grad", self._h)

    def dl_dp(self) -> np.ndarray:
        out = np.empty(self.plan.graph.n_factors, np.float64)
        call("flbp_adjoint_get_dl_dp", self._h, _ptr(out, C.c_double))
        return out

    def set_rules(self, rule: np.ndarray, n_rules: int) -> None:
        r = np.ascontiguousarray(rule, np.int32)
        self._n_rules = n_rules
        call("flbp_adjoint_set_rules", self._h, _ptr(r, C.c_int32), n_rules)

    def rule_grad(self) -> np.ndarray:
        """dL/dtheta from the accumulated dL/dp, reduced on the device."""
        out = np.empty(self._n_rules, np.float64)
        call("flbp_adjoint_rule_grad", self._h, _ptr(out, C.c_double))
        return out

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        call("flbp_adjoint_last_timing", self._h, C.byref(a), C.byref(b))
        return float(a.value), float(b.value)

    @property
    def device_bytes(self) -> int:
        return int(lib().flbp_adjoint_device_bytes(self._h))


class TiedProblem:
    """A graph whose AND probabilities are tied to rule logits, resident on
    one GPU, with the alarm variables whose marginals are differentiated."""

    def __init__(self, graph: Graph, rule: np.ndarray, n_rules: int, alarm_labels,
                 n_sweeps: int, checkpoint_every: int, damping: float, device: int = 0):
        self.graph = graph
        self.base_p = graph.p.copy()
        self.rule = np.ascontiguousarray(rule, np.int32)
        self.n_rules = n_rules
        self.alarm_idx = graph.index_of(alarm_labels)
        self.plan = Plan(graph, device)
        self.adj = Adjoint(self.plan, n_sweeps, checkpoint_every, damping)
        self.p = self.base_p
        self.last_theta = None   # theta of the forward the adjoint holds

    def marginals(self, theta: np.ndarray) -> np.ndarray:
        theta = np.array(theta, np.float64, copy=True)
        self.p = tied_p(self.base_p, self.rule, theta)
        self.plan.set_p(self.p)
        out = self.adj.forward()[self.alarm_idx]
        self.last_theta = theta
        return out

    def grad_at(self, theta: np.ndarray, dl_dalarm: np.ndarray) -> np.ndarray:
        """dL/dtheta at theta: the adjoint differentiates its last forward, so
        the forward is run again first when that was at another theta."""
        theta = np.asarray(theta, np.float64)
        if self.last_theta is None or not np.array_equal(self.last_theta.view(np.uint64), theta.view(np.uint64)):
            self.marginals(theta)
        return self.grad(dl_dalarm)

    def grad(self, dl_dalarm: np.ndarray) -> np.ndarray:
        """dL/dtheta at the theta of the last forward."""
        dl_db = np.zeros(self.graph.n_vars, np.float64)
        np.add.at(dl_db, self.alarm_idx, dl_dalarm)
        dl_dp = self.adj.backward(dl_db)
        return rule_grad(dl_dp, self.p, self.rule, self.n_rules)


def tied_marginals_function():
    """torch.autograd.Function: theta (float64 CPU tensor) -> alarm marginals."""
    import torch

    class TiedMarginals(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, problem: TiedProblem):
            ctx.problem = problem
            # the theta of this node: several applications may run before one
            # backward, and each must be differentiated at its own theta
            ctx.theta = theta.detach().cpu().numpy().astype(np.float64).copy()
            b = problem.marginals(ctx.theta)
            return torch.from_numpy(b)

        @staticmethod
        def backward(ctx, grad_out):
            g = ctx.problem.grad_at(ctx.theta, grad_out.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g), None

    return TiedMarginals
