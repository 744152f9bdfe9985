"""The hand-written adjoint on the GPU.

- forward beliefs bit-identical to the oracle's T fixed sweeps;
- gradient vs the oracle's CPU unrolled FP64 reverse-mode gradient
  (orc_grad, itself pinned by central differences in test_oracle_grad.py);
- gradient vs central finite differences of the GPU forward (h = 1e-6) on
  the 100k-variable, 200-rule, T = 50, damping 0.5 configuration;
- the checkpoint interval does not change the result;
- the torch.autograd.Function wrapper.
Tolerance: |g - ref| <= 1e-6 * max|ref| + 1e-6 * |ref| (relative with an
absolute floor at the gradient's scale).
"""
import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from fastlbp_amd.adjoint import TiedProblem, tied_marginals_function, tied_p
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu


def make(n_vars, seed, n_rules, n_alarms, T, k, damping):
    gen = fl.generate(n_vars, seed, n_rules=n_rules, n_alarms=n_alarms)
    rng = np.random.default_rng(seed)
    theta = rng.normal(2.5, 0.3, n_rules)
    y = rng.integers(0, 2, len(gen.alarms)).astype(np.float64)
    prob = TiedProblem(fl.Graph(gen.factors), gen.rule, n_rules, gen.alarms, T, k, damping)
    return gen, theta, y, prob


def seq_loss(b, y):
    """sum of squares in alarm order, as the oracle adds them"""
    L = 0.0
    for bi, yi in zip(b.tolist(), y.tolist()):
        L += (bi - yi) * (bi - yi)
    return L


def loss_and_grad(prob, theta, y):
    b = prob.marginals(theta)
    r = b - y
    return seq_loss(b, y), prob.grad(2.0 * r)


def oracle_forward(gen, theta, T, damping):
    o = oracle.OracleGraph(gen.factors.with_p(tied_p(gen.factors.p, gen.rule, theta)))
    o.init()
    o.run_fixed(T, damping)
    return o


def close(g, ref, rel=1e-6):
    scale = np.max(np.abs(ref))
    return np.all(np.abs(g - ref) <= rel * scale + rel * np.abs(ref))


@pytest.mark.parametrize("seed,damping,T,k", [(1, 0.5, 20, 5), (2, 0.0, 15, 4), (3, 0.3, 25, 25)])
def test_gradient_matches_oracle_unrolled(seed, damping, T, k):
    n_rules = 20
    gen, theta, y, prob = make(3000, seed, n_rules, 60, T, k, damping)
    L, g = loss_and_grad(prob, theta, y)
    o = oracle_forward(gen, theta, T, damping)
    assert same_bits(prob.adj.forward(), o.beliefs())   # the forward is the reference arithmetic
    Lo, go = oracle.grad(gen.factors, gen.rule, n_rules, theta, gen.alarms, y, T, damping)
    assert L == Lo
    assert close(g, go, rel=1e-9), np.max(np.abs(g - go))


def test_checkpoint_interval_does_not_change_gradient():
    gen, theta, y, prob = make(2000, 4, 10, 40, 24, 1, 0.5)
    _, g1 = loss_and_grad(prob, theta, y)
    for k in (5, 7, 24):
        p2 = TiedProblem(prob.graph, gen.rule, 10, gen.alarms, 24, k, 0.5)
        _, g2 = loss_and_grad(p2, theta, y)
        assert np.array_equal(g1, g2)


def test_c4_gradient_vs_central_differences_and_oracle():
    """V = 100k, 200 tied rules, 1000 alarms, T = 50, damping 0.5, checkpoint every 10."""
    n_rules = 200
    gen, theta, y, prob = make(100_000, 4, n_rules, 1000, 50, 10, 0.5)
    L, g = loss_and_grad(prob, theta, y)
    h = 1e-6
    fd = np.zeros(n_rules)
    for r in range(n_rules):
        tp, tm = theta.copy(), theta.copy()
        tp[r] += h
        tm[r] -= h
        bp = prob.marginals(tp)
        bm = prob.marginals(tm)
        fd[r] = (np.sum((bp - y) ** 2) - np.sum((bm - y) ** 2)) / (2 * h)
    assert close(g, fd, rel=1e-6), np.max(np.abs(g - fd) / np.max(np.abs(fd)))
    Lo, go = oracle.grad(gen.factors, gen.rule, n_rules, theta, gen.alarms, y, 50, 0.5)
    assert L == Lo and close(g, go, rel=1e-9)


def test_torch_autograd_function():
    import torch
    gen, theta, y, prob = make(2000, 6, 8, 30, 20, 5, 0.5)
    fn = tied_marginals_function()
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    b = fn.apply(th, prob)
    loss = ((b - torch.from_numpy(y)) ** 2).sum()
    loss.backward()
    _, g = loss_and_grad(prob, theta, y)
    assert np.array_equal(th.grad.numpy(), g)


def test_gradient_with_in_place_head_clamps():
    """The adjoint of a plan with head clamps applied in place equals the
    adjoint of the graph rebuilt by clamp() (alarms without body uses)."""
    gen = fl.generate(20000, 9, n_rules=50, n_alarms=300)
    g = fl.Graph(gen.factors)
    aidx = g.index_of(gen.alarms)
    rng = np.random.default_rng(9)
    free = gen.alarms[g.body_uses(aidx) == 0]
    labels = rng.choice(free, size=25, replace=False)
    values = rng.integers(0, 2, 25).astype(np.uint8)
    theta = rng.normal(2.5, 0.3, 50)
    y = rng.integers(0, 2, len(gen.alarms)).astype(np.float64)
    inplace = TiedProblem(g, gen.rule, 50, gen.alarms, 20, 5, 0.5)
    inplace.plan.clamp_vars(g.index_of(labels), values)
    rebuilt = TiedProblem(fl.Graph(fl.clamp(gen.factors, labels, values)), gen.rule, 50, gen.alarms, 20, 5, 0.5)
    la, ga = loss_and_grad(inplace, theta, y)
    lb, gb = loss_and_grad(rebuilt, theta, y)
    assert la == lb
    assert same_bits(ga, gb)


def test_sparse_backward_and_device_rule_grad():
    """backward_sparse gives the dense backward's dL/dp bit for bit (repeated
    variables add), successive calls accumulate, and the device rule
    reduction matches rule_grad (and repeats its own bits)."""
    from fastlbp_amd.adjoint import rule_grad
    gen = fl.generate(30000, 12, n_rules=40, n_alarms=200)
    g = fl.Graph(gen.factors)
    prob = TiedProblem(g, gen.rule, 40, gen.alarms, 15, 4, 0.5)
    rng = np.random.default_rng(12)
    theta = rng.normal(2.5, 0.3, 40)
    p = tied_p(prob.base_p, prob.rule, theta)
    prob.plan.set_p(p)
    aidx = prob.alarm_idx
    free = np.flatnonzero(g.body_uses(aidx) == 0)
    adj = prob.adj
    adj.set_rules(prob.rule, 40)
    for trial in range(2):
        pick = rng.choice(free, size=10, replace=False)
        prob.plan.clamp_vars(aidx[pick], rng.integers(0, 2, 10).astype(np.uint8))
        full = adj.forward()
        assert same_bits(prob.plan.beliefs_at(aidx), full[aidx])
        gb = rng.normal(0, 1, len(aidx))
        dl_db = np.zeros(g.n_vars)
        np.add.at(dl_db, aidx, gb)
        np.add.at(dl_db, aidx[:5], gb[:5])
        dense = adj.backward(dl_db)
        adj.clear_grad()
        adj.backward_sparse(np.concatenate([aidx, aidx[:5]]), np.concatenate([gb, gb[:5]]))
        assert same_bits(adj.dl_dp(), dense)
        adj.clear_grad()
        adj.backward_sparse(aidx, gb)
        adj.backward_sparse(aidx[:5], gb[:5])
        assert close(adj.dl_dp(), dense)
        host = rule_grad(adj.dl_dp(), p, prob.rule, 40)
        dev = adj.rule_grad()
        assert close(dev, host)
        assert same_bits(dev, adj.rule_grad())
        prob.plan.reset_head_clamps()


@pytest.mark.parametrize("seed,damping,T,k", [(21, 0.4, 15, 5), (22, 0.0, 12, 12)])
def test_gradient_of_high_degree_factors_matches_oracle(seed, damping, T, k):
    """OR heads over up to 14 clauses and AND bodies of up to 10 variables:
    factors above the exact-degree code (degree > 6) take the O(degree)
    reverse path; the gradient equals the oracle's unrolled reverse mode
    (itself pinned by central differences on such factors)."""
    n_rules = 12
    gen = fl.generate(2500, seed, n_rules=n_rules, n_alarms=60, k_max=14, b_max=10)
    g_ = fl.Graph(gen.factors)
    assert np.diff(g_.fac_off).max() > 6
    rng = np.random.default_rng(seed)
    theta = rng.normal(-1.0, 0.3, n_rules)
    y = rng.integers(0, 2, len(gen.alarms)).astype(np.float64)
    prob = TiedProblem(g_, gen.rule, n_rules, gen.alarms, T, k, damping)
    L, g = loss_and_grad(prob, theta, y)
    o = oracle_forward(gen, theta, T, damping)
    assert same_bits(prob.adj.forward(), o.beliefs())
    Lo, go = oracle.grad(gen.factors, gen.rule, n_rules, theta, gen.alarms, y, T, damping)
    assert L == Lo
    assert np.max(np.abs(go)) > 1e-3
    assert close(g, go, rel=1e-9), np.max(np.abs(g - go))


def test_gradient_of_a_wide_factor_and_variable_vs_central_differences():
    """An OR over 700 clauses (wider than a staged block of the reverse
    kernels) and a leaf used by 1500 ANDs; the 1500 ANDs tied to 6 rules.
    The oracle's unrolled reverse mode is cubic in the degree, so the pin is
    central differences of the GPU forward (h = 1e-6)."""
    facs = [(0, 0, [], 0.7), (0, 1, [], 0.6)]
    nxt, clauses, rule = 2, [], [-1, -1]
    for i in range(1500):
        facs.append((1, nxt, [0, 1] if i % 3 else [0], 0.5, 0.01))
        rule.append(i % 6)
        if i < 700:
            clauses.append(nxt)
        nxt += 1
    facs.append((2, nxt, clauses, 1.0, 0.0))
    rule.append(-1)
    flist = fl.FactorList.build(facs)
    rule = np.array(rule, np.int32)
    g_ = fl.Graph(flist)
    assert np.diff(g_.fac_off).max() == 701 and np.diff(g_.var_off).max() == 1501
    labels = np.array([nxt, 3, 800, 1499], np.uint64)
    rng = np.random.default_rng(5)
    theta = rng.normal(-3.0, 0.3, 6)
    y = np.array([1.0, 0.0, 1.0, 0.0])
    prob = TiedProblem(g_, rule, 6, labels, 12, 4, 0.3)
    L, g = loss_and_grad(prob, theta, y)
    h = 1e-6
    fd = np.zeros(6)
    for r in range(6):
        tp, tm = theta.copy(), theta.copy()
        tp[r] += h
        tm[r] -= h
        fd[r] = (seq_loss(prob.marginals(tp), y) - seq_loss(prob.marginals(tm), y)) / (2 * h)
    assert np.max(np.abs(fd)) > 1e-4
    assert close(g, fd, rel=1e-6), (g, fd)
