"""Interleaved copies of a graph (fastlbp_amd.variants.interleave) on the
CPU, with the oracle: copy c of the K-copy graph, clamped as variant c,
solves to the same bits as variant c's own graph (the base factor list
clamped and rebuilt, examples/fl_wrapper.cu:26-45)."""
import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from fastlbp_amd.variants import interleave
from tests.helpers import same_bits


def variants(gen, graph, K, rng):
    aidx = graph.index_of(gen.alarms)
    free = gen.alarms[graph.body_uses(aidx) == 0]
    out = []
    for c in range(K):
        n = int(rng.integers(0, 6))     # variant 0 may clamp nothing
        lab = np.sort(rng.choice(free, size=n, replace=False)).astype(np.uint64)
        out.append((lab, rng.integers(0, 2, n).astype(np.uint8)))
    return out


def test_interleave_layout():
    gen = fl.generate(500, 3)
    K = 3
    u = interleave(gen.factors, K)
    F = gen.factors.n_factors
    assert u.n_factors == K * F
    deg = np.diff(gen.factors.body_off.astype(np.int64))
    assert np.array_equal(np.diff(u.body_off.astype(np.int64)), np.repeat(deg, K))
    for f in (0, 1, F // 2, F - 1):
        b = gen.factors.body[gen.factors.body_off[f]:gen.factors.body_off[f + 1]]
        for c in range(K):
            uf = f * K + c
            assert u.head[uf] == gen.factors.head[f] * K + c
            assert np.array_equal(u.body[u.body_off[uf]:u.body_off[uf + 1]], b * np.uint64(K) + np.uint64(c))
            assert u.p[uf] == gen.factors.p[f] and u.type[uf] == gen.factors.type[f]


@pytest.mark.parametrize("K,damping,T", [(4, 0.3, 12), (3, 0.0, 9), (2, 0.5, 15)])
def test_copies_solve_as_their_variants(K, damping, T):
    gen = fl.generate(2500, 17, n_rules=10, n_alarms=80)
    g = fl.Graph(gen.factors)
    rng = np.random.default_rng(K)
    vs = variants(gen, g, K, rng)
    u = interleave(gen.factors, K)
    lab = np.concatenate([l * np.uint64(K) + np.uint64(c) for c, (l, _) in enumerate(vs)])
    val = np.concatenate([x for _, x in vs])
    order = np.argsort(lab)
    uc = fl.clamp(u, lab[order], val[order])
    gu = fl.Graph(uc)
    ou = oracle.OracleGraph(uc)
    ou.init()
    ou.run_fixed(T, damping)
    bu = ou.beliefs()
    for c, (l, x) in enumerate(vs):
        fc = fl.clamp(gen.factors, l, x)
        o = oracle.OracleGraph(fc)
        o.init()
        o.run_fixed(T, damping)
        idx = gu.index_of(fl.Graph(fc).var_label * np.uint64(K) + np.uint64(c))
        assert same_bits(bu[idx], o.beliefs()), c
