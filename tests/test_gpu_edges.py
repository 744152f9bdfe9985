"""Edge cases on the GPU against the oracle: the empty graph (no variables,
no factors), a graph of prior factors only, batches and variant batches
that mix them with ordinary graphs, and NaN probabilities spreading
through messages, products and beliefs."""
import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from fastlbp_amd.adjoint import TiedProblem
from fastlbp_amd.batch import Batch, Group
from fastlbp_amd.variants import VariantBatch
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu


def empty():
    return fl.FactorList.build([])


def priors_only():
    return fl.FactorList.build([(0, 3, [], 0.25), (0, 8, [], 0.5), (0, 1, [], 0.999)])


@pytest.mark.parametrize("variant", [0, 3, 4])
def test_empty_graph(variant):
    """The reference's loop runs one sweep (max over no beliefs = 0 <= tol)."""
    print(f"[edges] empty graph, variant {variant}", flush=True)
    g = fl.Graph(empty())
    assert (g.n_vars, g.n_factors, g.n_edges) == (0, 0, 0)
    plan = fl.Plan(g)
    plan.set_variant(variant)
    plan.init()
    r = plan.run(1e-9, 100, 0.0)
    o = oracle.OracleGraph(empty())
    o.init()
    it, md = o.run(1e-9, 100, 0.0)
    assert (r.iters, r.maxdiff) == (it, md)
    assert plan.beliefs().shape == (0,)
    plan.run_fixed(3, 0.3)
    plan.init()
    yet, low, reason, md, iters = plan.run_hist(1e-6, 0, 10, 4)
    o2 = oracle.OracleGraph(empty())
    o2.init()
    oyet, olow, oreason, omd = o2.run_hist(1e-6, 0, 10, 4)
    assert reason == oreason and md == omd and low.shape == (0,)
    assert (np.isnan(yet) and np.isnan(oyet)) or yet == oyet


def test_priors_only_graph():
    print("[edges] priors only", flush=True)
    fl_ = priors_only()
    plan = fl.Plan(fl.Graph(fl_))
    plan.init()
    r = plan.run(1e-9, 100, 0.2)
    o = oracle.OracleGraph(fl_)
    o.init()
    it, md = o.run(1e-9, 100, 0.2)
    assert r.iters == it and same_bits(plan.beliefs(), o.beliefs())


def test_batch_mixing_empty_tiny_and_generated_graphs():
    lists = [empty(), priors_only(), fl.generate(3000, 31).factors, empty(), fl.generate(200, 32).factors]
    print("[edges] batch of empty, tiny and generated graphs", flush=True)
    b = Batch([Group(fl.Graph(x)) for x in lists])
    b.init()
    iters, md, _ = b.run(1e-9, 300, 0.3)
    for v, x in enumerate(lists):
        o = oracle.OracleGraph(x)
        o.init()
        it, omd = o.run(1e-9, 300, 0.3)
        assert int(iters[v]) == it, v
        assert same_bits(b.beliefs(v), o.beliefs()), v


def test_variant_batch_of_one_is_the_single_problem():
    """K = 1: the interleaved graph is the graph itself."""
    n_rules, T, k, d = 12, 10, 5, 0.3
    gen = fl.generate(2000, 33, n_rules=n_rules, n_alarms=40)
    rng = np.random.default_rng(33)
    theta = rng.normal(2.5, 0.3, n_rules)
    y = rng.integers(0, 2, len(gen.alarms)).astype(np.float64)
    print("[edges] variant batch K = 1", flush=True)
    vb = VariantBatch(gen.factors, gen.rule, n_rules, gen.alarms, 1, T, k, d)
    vb.set_clamps([((), ())])
    vb.set_theta(theta)
    b = vb.forward()
    vb.clear_grad()
    vb.backward(2.0 * (b - y[None, :]))
    g = vb.rule_grad()
    vb.close()
    prob = TiedProblem(fl.Graph(gen.factors), gen.rule, n_rules, gen.alarms, T, k, d)
    b1 = prob.marginals(theta)
    g1 = prob.grad(2.0 * (b1 - y))
    assert same_bits(b[0], b1)
    assert np.allclose(g, g1, rtol=1e-12, atol=1e-14 * np.max(np.abs(g1)))


@pytest.mark.parametrize("variant", [0, 3, 4])
def test_nan_probabilities_spread_as_in_the_oracle(variant):
    """Factors with p = NaN make NaN messages, which spread into products,
    beliefs and the rescaled product pairs (a NaN non-unit value, or a NaN
    unit once both values are NaN: the 8-byte product records of the state
    layout keep both).  Every message and belief matches the oracle's
    sweeps, NaN for NaN, and the finite ones bit for bit."""
    print(f"[edges] NaN probabilities, variant {variant}", flush=True)
    gen = fl.generate(6000, 71)
    f = gen.factors
    rng = np.random.default_rng(71)
    p = f.p.copy()
    p[rng.choice(f.n_factors, size=5, replace=False)] = np.nan
    flist = fl.FactorList(f.type, f.head, f.body_off, f.body, p, f.q, f.head_clamped, f.mask0, f.mask1)
    g = fl.Graph(flist)
    plan = fl.Plan(g)
    plan.set_variant(variant)
    plan.init()
    o = oracle.OracleGraph(flist)
    o.init()
    for n in (1, 1, 6):   # 14 NaN beliefs after one sweep, 1,799 after two, all after eight
        plan.run_fixed(n, 0.3)
        o.run_fixed(n, 0.3)
        fv, vf = plan.messages()
        ofv, ovf = o.messages()
        assert same_bits(fv, ofv) and same_bits(vf, ovf), n
        assert same_bits(plan.beliefs(), o.beliefs()), n
        if n == 1:
            b = plan.beliefs()
            assert np.isnan(b).any() and np.isfinite(b).any()
    r = plan.run(1e-9, 200, 0.3)   # the max change is NaN: the reference's loop stops
    it, md = o.run(1e-9, 200, 0.3)
    assert r.iters == it and same_bits([r.maxdiff], [md])
    assert same_bits(plan.beliefs(), o.beliefs())
