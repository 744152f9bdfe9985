"""Clamped variants solved and differentiated together on the GPU
(fastlbp_amd.variants.VariantBatch: K interleaved copies on one plan, one
adjoint):

- each copy's T-sweep alarm marginals bit-identical to the oracle's solve
  of its variant alone (the base factor list clamped and rebuilt);
- dL/dtheta of the summed loss against the sum over variants of the
  oracle's unrolled reverse-mode gradient (orc_grad);
- re-clamping the same plan for the next micro-batch of variants leaves no
  trace of the previous one;
- at 100k variables x 4 copies, dL/dtheta against central finite
  differences of the GPU forward (h = 1e-6) on the rules with the largest
  gradients.
"""
import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from fastlbp_amd.adjoint import tied_p
from fastlbp_amd.variants import VariantBatch
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu


def close(g, ref, rel):
    scale = np.max(np.abs(ref))
    return np.all(np.abs(g - ref) <= rel * scale + rel * np.abs(ref))


def make_variants(gen, n, rng, max_clamps=8):
    g = fl.Graph(gen.factors)
    aidx = g.index_of(gen.alarms)
    free = gen.alarms[g.body_uses(aidx) == 0]
    out = []
    for _ in range(n):
        m = int(rng.integers(1, max_clamps + 1))
        lab = np.sort(rng.choice(free, size=m, replace=False)).astype(np.uint64)
        out.append((lab, rng.integers(0, 2, m).astype(np.uint8)))
    return out


def seq_loss(b, y):
    L = 0.0
    for bi, yi in zip(b.tolist(), y.tolist()):
        L += (bi - yi) * (bi - yi)
    return L


@pytest.mark.parametrize("K,T,k,damping", [(4, 16, 4, 0.3), (2, 12, 12, 0.5)])
def test_variant_batch_matches_oracle_per_variant(K, T, k, damping):
    n_rules = 16
    gen = fl.generate(3000, 21, n_rules=n_rules, n_alarms=60)
    rng = np.random.default_rng(K)
    theta = rng.normal(2.5, 0.3, n_rules)
    vb = VariantBatch(gen.factors, gen.rule, n_rules, gen.alarms, K, T, k, damping)
    vb.set_theta(theta)
    for batch in range(2):   # two micro-batches on the same plan
        vs = make_variants(gen, K, rng)
        ys = rng.integers(0, 2, (K, len(gen.alarms))).astype(np.float64)
        print(f"[variants] K={K} T={T} k={k} damping={damping} micro-batch {batch}", flush=True)
        vb.set_clamps(vs)
        b = vb.forward()
        vb.clear_grad()
        vb.backward(2.0 * (b - ys))
        g = vb.rule_grad()
        g_ref = np.zeros(n_rules)
        for c, ((lab, val), y) in enumerate(zip(vs, ys)):
            fc = fl.clamp(gen.factors, lab, val)
            assert np.array_equal(fc.head, gen.factors.head)   # head clamps keep the factor order
            o = oracle.OracleGraph(fc.with_p(tied_p(fc.p, gen.rule, theta)))
            o.init()
            o.run_fixed(T, damping)
            gc = fl.Graph(fc)
            assert same_bits(b[c], o.beliefs()[gc.index_of(gen.alarms)]), c
            Lo, go = oracle.grad(fc, gen.rule, n_rules, theta, gen.alarms, y, T, damping)
            assert Lo == seq_loss(b[c], y)
            g_ref += go
        assert np.max(np.abs(g_ref)) > 1e-4
        assert close(g, g_ref, 1e-9), np.max(np.abs(g - g_ref))
    vb.close()


def test_variant_batch_gradient_by_finite_differences():
    n_rules, K, T, k, damping = 200, 4, 20, 5, 0.3
    gen = fl.generate(100_000, 23, n_rules=n_rules, n_alarms=1000)
    rng = np.random.default_rng(23)
    theta = rng.normal(2.5, 0.3, n_rules)
    vs = make_variants(gen, K, rng, max_clamps=50)
    ys = rng.integers(0, 2, (K, len(gen.alarms))).astype(np.float64)
    vb = VariantBatch(gen.factors, gen.rule, n_rules, gen.alarms, K, T, k, damping)
    vb.set_clamps(vs)

    def loss(th):
        vb.set_theta(th)
        b = vb.forward()
        return float(np.sum((b - ys) ** 2)), b

    L0, b = loss(theta)
    vb.clear_grad()
    vb.backward(2.0 * (b - ys))
    g = vb.rule_grad()
    h = 1e-6
    top = np.argsort(-np.abs(g))[:4]
    for r in top:
        e = np.zeros(n_rules)
        e[r] = h
        fd = (loss(theta + e)[0] - loss(theta - e)[0]) / (2 * h)
        print(f"[variants] rule {r}: adjoint {g[r]:.9e} central difference {fd:.9e}", flush=True)
        assert abs(g[r] - fd) <= 1e-6 * abs(fd) + 1e-7 * np.max(np.abs(g)), (r, g[r], fd)
    vb.close()
