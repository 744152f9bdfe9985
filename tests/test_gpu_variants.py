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


def test_fit_step_reduced_64_variants_against_oracle():
    """One gradient step of the tied-rule fit at reduced size: 64 clamped variants in 4
    micro-batches of 16 interleaved copies, dL/dp accumulated on the device
    over all of them; dL/dtheta against the sum of the oracle's gradients
    of the 64 variants, and every variant's T-sweep alarm marginals bit for
    bit."""
    n_rules, K, T, k, damping = 16, 16, 12, 4, 0.3
    gen = fl.generate(1500, 41, n_rules=n_rules, n_alarms=60)
    rng = np.random.default_rng(41)
    theta = rng.normal(2.5, 0.3, n_rules)
    vs = make_variants(gen, 64, rng, max_clamps=12)
    ys = rng.integers(0, 2, (64, len(gen.alarms))).astype(np.float64)
    vb = VariantBatch(gen.factors, gen.rule, n_rules, gen.alarms, K, T, k, damping)
    vb.set_theta(theta)
    vb.clear_grad()
    bs = []
    for m in range(64 // K):
        print(f"[variants] fit, reduced: micro-batch {m}", flush=True)
        vb.set_clamps(vs[m * K:(m + 1) * K])
        b = vb.forward()
        vb.backward(2.0 * (b - ys[m * K:(m + 1) * K]))
        bs.append(b)
    g = vb.rule_grad()
    vb.close()
    b = np.concatenate(bs)
    g_ref = np.zeros(n_rules)
    L = L_ref = 0.0
    for c, ((lab, val), y) in enumerate(zip(vs, ys)):
        fc = fl.clamp(gen.factors, lab, val)
        o = oracle.OracleGraph(fc.with_p(tied_p(fc.p, gen.rule, theta)))
        o.init()
        o.run_fixed(T, damping)
        assert same_bits(b[c], o.beliefs()[fl.Graph(fc).index_of(gen.alarms)]), c
        Lo, go = oracle.grad(fc, gen.rule, n_rules, theta, gen.alarms, y, T, damping)
        L_ref += Lo
        L += seq_loss(b[c], y)
        g_ref += go
    assert L == L_ref
    assert close(g, g_ref, 1e-9), np.max(np.abs(g - g_ref))


def test_fit_step_full_size_properties():
    """One gradient step of the tied-rule fit on its full graph (20M variables, 30M factors,
    89.7M edges; 2 interleaved variants, T = 40, damping 0.3): marginals
    finite and in [0, 1], every clamped alarm at its clamp value, a small
    step against the gradient lowers the loss, and the gradient matches
    central differences on the two rules with the largest components."""
    n_rules, K, T, k, damping = 200, 2, 40, 8, 0.3
    print("[variants] fit, full size: generating 20M variables", flush=True)
    gen = fl.generate(20_000_000, 5, n_rules=n_rules, n_alarms=1000)
    rng = np.random.default_rng(5)
    vs = make_variants(gen, K, rng, max_clamps=50)
    ys = rng.integers(0, 2, (K, len(gen.alarms))).astype(np.float64)
    theta = rng.normal(2.5, 0.3, n_rules)
    print("[variants] fit, full size: building 2 interleaved copies", flush=True)
    vb = VariantBatch(gen.factors, gen.rule, n_rules, gen.alarms, K, T, k, damping)
    assert vb.graph.n_edges == 2 * 89_709_144
    vb.set_clamps(vs)

    def loss(th):
        vb.set_theta(th)
        b = vb.forward()
        return float(np.sum((b - ys) ** 2)), b

    print("[variants] fit, full size: forward + backward", flush=True)
    L0, b = loss(theta)
    assert np.all(np.isfinite(b)) and np.all((b >= 0) & (b <= 1))
    pos = {int(a): i for i, a in enumerate(gen.alarms)}
    for c, (lab, val) in enumerate(vs):   # clamped alarms sit at their values
        idx = np.array([pos[int(a)] for a in lab])
        assert np.allclose(b[c, idx], val, atol=1e-12)
    vb.clear_grad()
    vb.backward(2.0 * (b - ys))
    g = vb.rule_grad()
    assert np.all(np.isfinite(g)) and np.max(np.abs(g)) > 0
    L1, _ = loss(theta - 1e-3 * g / np.max(np.abs(g)))
    print(f"[variants] fit, full size: loss {L0:.9g} -> {L1:.9g} after a step against the gradient", flush=True)
    assert L1 < L0
    h = 1e-6
    for r in np.argsort(-np.abs(g))[:2]:
        e = np.zeros(n_rules)
        e[r] = h
        fd = (loss(theta + e)[0] - loss(theta - e)[0]) / (2 * h)
        print(f"[variants] fit, full size: rule {r}: adjoint {g[r]:.9e} central difference {fd:.9e}", flush=True)
        assert abs(g[r] - fd) <= 1e-5 * abs(fd) + 1e-7 * np.max(np.abs(g)), (r, g[r], fd)
    vb.close()
