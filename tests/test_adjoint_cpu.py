"""The autograd wrapper of the adjoint (CPU only, the device part mocked).

The adjoint differentiates the last forward its plan ran.  TiedMarginals
must differentiate each application at its own theta, also when several
applications run before one backward (f(theta1) + f(theta2), or an LBFGS
closure that evaluates again)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from fastlbp_amd.adjoint import TiedProblem, tied_marginals_function, tied_p


class _Graph:
    def __init__(self, p):
        self.p = p
        self.n_vars = self.n_factors = p.shape[0]


class _Plan:
    """Keeps the p of the last set_p, as the device plan does."""
    def set_p(self, p):
        self.p = np.array(p)


class _Adjoint:
    """b_i = p_i^2 for variable i = factor i; backward differentiates the
    plan's p of the last forward, as the device adjoint does."""
    def __init__(self, plan):
        self.plan = plan

    def forward(self):
        self.fwd_p = self.plan.p.copy()
        return self.fwd_p ** 2

    def backward(self, dl_db):
        return dl_db * 2.0 * self.fwd_p


def mock_problem(n_rules=3, n_f=7, seed=0):
    rng = np.random.default_rng(seed)
    prob = TiedProblem.__new__(TiedProblem)
    prob.graph = _Graph(rng.uniform(0.2, 0.8, n_f))
    prob.base_p = prob.graph.p.copy()
    prob.rule = rng.integers(-1, n_rules, n_f).astype(np.int32)
    prob.rule[:n_rules] = np.arange(n_rules)
    prob.n_rules = n_rules
    prob.alarm_idx = np.arange(n_f)
    prob.plan = _Plan()
    prob.adj = _Adjoint(prob.plan)
    prob.p = prob.base_p
    prob.last_theta = None
    return prob


def exact_grad(prob, theta, w):
    """d/dtheta of sum_i w_i * p_i(theta)^2, p = sigmoid(theta[rule])."""
    p = tied_p(prob.base_p, prob.rule, theta)
    g = np.zeros(prob.n_rules)
    for i, r in enumerate(prob.rule):
        if r >= 0:
            g[r] += w[i] * 2 * p[i] * p[i] * (1 - p[i])
    return g


def test_two_forwards_before_one_backward():
    prob = mock_problem()
    f = tied_marginals_function()
    t1 = np.array([0.3, -1.2, 2.0])
    t2 = np.array([1.1, 0.4, -0.7])
    th1 = torch.tensor(t1, dtype=torch.float64, requires_grad=True)
    th2 = torch.tensor(t2, dtype=torch.float64, requires_grad=True)
    w1 = np.linspace(0.5, 1.5, 7)
    w2 = np.linspace(-1.0, 2.0, 7)
    y = (f.apply(th1, prob) * torch.from_numpy(w1)).sum() + (f.apply(th2, prob) * torch.from_numpy(w2)).sum()
    y.backward()
    np.testing.assert_allclose(th1.grad.numpy(), exact_grad(prob, t1, w1), rtol=1e-13)
    np.testing.assert_allclose(th2.grad.numpy(), exact_grad(prob, t2, w2), rtol=1e-13)


def test_same_theta_reuses_the_forward():
    prob = mock_problem(seed=1)
    f = tied_marginals_function()
    t = np.array([0.2, 0.1, -0.3])
    th = torch.tensor(t, dtype=torch.float64, requires_grad=True)
    (f.apply(th, prob) * 2.0).sum().backward()
    runs = []
    orig = prob.adj.forward
    prob.adj.forward = lambda: runs.append(1) or orig()
    th.grad = None
    out = f.apply(th, prob)
    n = len(runs)
    (out * 2.0).sum().backward()
    assert len(runs) == n          # no second forward when theta is unchanged
    np.testing.assert_allclose(th.grad.numpy(), exact_grad(prob, t, np.full(7, 2.0)), rtol=1e-13)
