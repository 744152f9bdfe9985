"""The oracle's unrolled reverse-mode gradient (orc_grad) against central
finite differences of its own forward (orc_loss), h = 1e-6 on the rule
logits -- the pin for the CPU FP64 reference gradient of the adjoint."""
import numpy as np
import pytest

import fastlbp_amd as fl
import oracle


def problem(n_vars, seed, n_rules, n_alarms):
    gen = fl.generate(n_vars, seed, n_rules=n_rules, n_alarms=n_alarms)
    rng = np.random.default_rng(seed)
    theta = rng.normal(2.5, 0.3, n_rules)
    targets = rng.integers(0, 2, len(gen.alarms)).astype(np.float64)
    return gen, theta, targets


@pytest.mark.parametrize("seed,damping,T", [(1, 0.5, 20), (2, 0.0, 12), (3, 0.3, 30)])
def test_reverse_mode_matches_central_differences(seed, damping, T):
    n_rules = 12
    gen, theta, targets = problem(400, seed, n_rules, 40)
    args = (gen.factors, gen.rule, n_rules)
    L, g = oracle.grad(*args, theta, gen.alarms, targets, T, damping)
    assert L == oracle.loss(*args, theta, gen.alarms, targets, T, damping)
    h = 1e-6
    fd = np.zeros(n_rules)
    for r in range(n_rules):
        tp, tm = theta.copy(), theta.copy()
        tp[r] += h
        tm[r] -= h
        fd[r] = (oracle.loss(*args, tp, gen.alarms, targets, T, damping) -
                 oracle.loss(*args, tm, gen.alarms, targets, T, damping)) / (2 * h)
    scale = np.max(np.abs(fd))
    assert scale > 1e-6
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6 * scale)


def test_gradient_is_zero_for_unused_rules():
    gen, theta, targets = problem(300, 7, 500, 20)   # more rules than AND factors use
    L, g = oracle.grad(gen.factors, gen.rule, 500, theta, gen.alarms, targets, 10, 0.5)
    unused = np.setdiff1d(np.arange(500), gen.rule[gen.rule >= 0])
    assert unused.size > 0 and np.all(g[unused] == 0.0)


def test_reverse_mode_on_high_degree_factors():
    """OR heads over up to 14 clauses and AND bodies of up to 10 variables
    (factor degrees up to 15): the reverse mode still matches central
    differences -- the reference the GPU adjoint's general-degree factor
    code is checked against.  Rule probabilities near sigmoid(-1) keep the
    wide noisy-ORs away from saturation, so the gradient is not tiny."""
    n_rules = 8
    gen = fl.generate(300, 21, n_rules=n_rules, n_alarms=30, k_max=14, b_max=10)
    assert np.max(np.diff(gen.factors.body_off)) >= 10
    rng = np.random.default_rng(21)
    theta = rng.normal(-1.0, 0.3, n_rules)
    targets = rng.integers(0, 2, len(gen.alarms)).astype(np.float64)
    args = (gen.factors, gen.rule, n_rules)
    L, g = oracle.grad(*args, theta, gen.alarms, targets, 15, 0.4)
    h = 1e-6
    fd = np.zeros(n_rules)
    for r in range(n_rules):
        tp, tm = theta.copy(), theta.copy()
        tp[r] += h
        tm[r] -= h
        fd[r] = (oracle.loss(*args, tp, gen.alarms, targets, 15, 0.4) -
                 oracle.loss(*args, tm, gen.alarms, targets, 15, 0.4)) / (2 * h)
    scale = np.max(np.abs(fd))
    assert scale > 1e-6
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6 * scale)
