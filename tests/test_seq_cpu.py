"""Sequential update schedules (updates = SEQFIX / TOPO) on the CPU.

Pins of the oracle's restatement and of the engine's host-side schedule
builder (flbp_schedule_*, no GPU needed):

- std::mt19937 against its standard known answer; the restated libstdc++
  std::shuffle against std::shuffle itself (both of its branches).
- Wave splits of putUpdateSeqToKernelFused (csrc/ext/causal_bp.cu:583-722)
  worked by hand on a 4-variable graph (plain order, repeated edges with
  and without respect_preorder, the TOPO order of buildTopoUpdateSeq
  203-304), checked against both the oracle and the engine.
- Engine orders and waves equal the oracle's on generated graphs, the
  SEQFIX generator advancing across schedules like the reference's
  thread_local rnd_gen.
- The engine's execution items (per wave one group item per factor, with
  separate v->f / f->v items where needed) run level by level, in any
  order within a level, give the wave execution's bits; merging levels
  does not.
- Sequential BP arithmetic: exact marginals on trees; the PARALL fixed point
  on loopy graphs; the in-place products agree with the messages' product;
  each f->v code with the sequential scale() against enumeration.
"""
import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from tests.test_oracle_pins import (close, code_of, exact_fv, exact_marginals, random_msgs,
                                    random_tree_list)

# vars 0..3; F0 prior(0), F1 prior(1), F2 AND(0,1)->2, F3 OR(2)->3.
# var-major edges: e0=(0,F0) e1=(0,F2) e2=(1,F1) e3=(1,F2) e4=(2,F2) e5=(2,F3) e6=(3,F3)
# factor-major slots: s0=F0:0 s1=F1:1 s2=F2:0 s3=F2:1 s4=F2:2 s5=F3:2 s6=F3:3
SMALL = [(0, 0, [], 0.3), (0, 1, [], 0.6), (1, 2, [0, 1], 0.9, 0.05), (2, 3, [2], 0.8, 0.1)]


def small():
    return fl.FactorList.build(SMALL)


def engine_waves(order, respect):
    s = fl.Schedule(fl.Graph(small()), order=order, respect_preorder=respect)
    return [np.asarray(x, np.int64).tolist() for x in s.waves()]


def oracle_waves(order, respect):
    o = oracle.OracleGraph(small())
    o.set_sequence(np.asarray(order, np.uint64), respect_preorder=respect)
    return [np.asarray(x, np.int64).tolist() for x in o.waves()]


def test_small_graph_numbering():
    g = fl.Graph(small())
    assert g.fac_off.tolist() == [0, 1, 2, 5, 7]
    assert g.edge_var.tolist() == [0, 1, 0, 1, 2, 2, 3]
    assert g.var_off.tolist() == [0, 2, 4, 6, 7]
    assert g.var_edge.tolist() == [0, 2, 1, 3, 4, 5, 6]


# (order, respect_preorder) -> (fv_off, fv_edge, vf_off, vf_slot), worked by hand
HAND = [
    # e4 hits v_set (var 2 already updated by e5); e6 needs v->f 2->F3 while
    # the wave updates e4 = F2->2, so it opens wave 3
    (([1, 3, 5, 4, 0, 6, 2], False),
     ([0, 3, 5, 7], [1, 3, 5, 4, 0, 6, 2], [0, 4, 6, 7], [3, 4, 2, 6, 2, 3, 5])),
    (([1, 3, 5, 4, 0, 6, 2], True),
     ([0, 3, 5, 7], [1, 3, 5, 4, 0, 6, 2], [0, 4, 6, 7], [3, 4, 2, 6, 2, 3, 5])),
    # repeated e4: its rank is its last position (2), after e6 (1), so with
    # respect_preorder the F2->2 update does not block e6
    (([4, 6, 4], True), ([0, 2, 3], [4, 6, 4], [0, 3, 5], [2, 3, 5, 2, 3])),
    (([4, 6, 4], False), ([0, 1, 3], [4, 6, 4], [0, 2, 5], [2, 3, 5, 2, 3])),
    # the empty order is one empty wave (the final push of 720-721)
    (([], False), ([0, 0], [], [0, 0], [])),
]


@pytest.mark.parametrize("case", range(len(HAND)))
def test_hand_worked_wave_splits(case):
    (order, respect), want = HAND[case]
    want = [list(x) for x in want]
    assert oracle_waves(order, respect) == want
    assert engine_waves(order, respect) == want


def test_hand_worked_topo_order_and_waves():
    # BFS from var 0: ranks var0 0, F0 1, F2 2, var1 3, var2 4, F1 5, F3 6,
    # var3 7; edges by (later rank, earlier rank, ...)
    want_order = [0, 1, 3, 4, 2, 5, 6]
    o = oracle.OracleGraph(small())
    assert o.order_topo().tolist() == want_order
    s = fl.Schedule(fl.Graph(small()), "TOPO")
    assert s.order().tolist() == want_order
    want = [[0, 1, 4, 7], [0, 1, 3, 4, 2, 5, 6], [0, 0, 3, 5], [3, 4, 2, 6, 5]]
    assert [np.asarray(x, np.int64).tolist() for x in s.waves()] == want
    info = s.info()
    assert info.updates == 2 and info.respect_preorder and info.n_waves == 3


def test_mt19937_known_answer():
    r = oracle.Rng(5489)
    for _ in range(9999):
        r.next()
    assert r.next() == 4123659995   # the standard's required 10000th output


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 7, 8, 1000, 65535, 65536, 70001])
def test_restated_shuffle_equals_std_shuffle(n):
    for seed in (42, 7):
        assert np.array_equal(oracle.Rng(seed).shuffle(n), oracle.std_shuffle(seed, n))


@pytest.mark.parametrize("n_vars", [500, 20000])
def test_engine_orders_and_waves_equal_oracle(n_vars):
    gen = fl.generate(n_vars, 11)
    g = fl.Graph(gen.factors)
    o = oracle.OracleGraph(gen.factors)
    rng = oracle.Rng(42)
    fl.seed(42)
    for k in range(2):   # the generator advances from one SEQFIX schedule to the next
        s = fl.Schedule(g, "SEQFIX")
        order = o.order_seqfix(rng)
        assert np.array_equal(s.order(), order), k
        o.set_sequence(order)
        for a, b in zip(s.waves(), o.waves()):
            assert np.array_equal(a, b)
    s = fl.Schedule(g, "TOPO")
    order = o.order_topo()
    assert np.array_equal(s.order(), order)
    o.set_sequence(order, respect_preorder=True)
    for a, b in zip(s.waves(), o.waves()):
        assert np.array_equal(a, b)
    # an explicit order with repeats
    r = np.random.default_rng(n_vars)
    order = r.integers(0, g.n_edges, size=g.n_edges + 50).astype(np.uint64)
    for respect in (False, True):
        s = fl.Schedule(g, order=order, respect_preorder=respect)
        o.set_sequence(order, respect_preorder=respect)
        for a, b in zip(s.waves(), o.waves()):
            assert np.array_equal(a, b)


def test_seed_restarts_the_sequence():
    g = fl.Graph(fl.generate(300, 2).factors)
    fl.seed(9)
    a = fl.Schedule(g, "SEQFIX").order()
    fl.seed(9)
    b = fl.Schedule(g, "SEQFIX").order()
    c = fl.Schedule(g, "SEQFIX").order()
    assert np.array_equal(a, b) and not np.array_equal(b, c)


def test_schedule_argument_errors():
    g = fl.Graph(small())
    with pytest.raises(ValueError):
        fl.Schedule(g, "PARALL")
    with pytest.raises(fl.FlbpError):
        fl.Schedule(g, order=[0, 7])   # edge 7 does not exist
    c = g.csr_c()
    import ctypes as C
    h = C.c_void_p()
    assert fl.lib().flbp_schedule_create(C.byref(c), 1, C.byref(h)) == -1   # PARALL


def _high_degree_graph():
    gen = fl.generate(2500, 21, k_min=6, k_max=14, b_min=4, b_max=10, leaf_frac=0.02)
    assert np.diff(fl.Graph(gen.factors).fac_off).max() > 6   # wide factors keep per-message items
    return gen


@pytest.mark.parametrize("mode,graph", [("SEQFIX", "plain"), ("TOPO", "plain"), ("SEQFIX", "wide"),
                                        ("TOPO", "wide"), ("order", "plain")])
def test_execution_items_reproduce_wave_bits(mode, graph):
    """The engine's execution items -- per wave one group item per factor
    (its v->f messages recomputed from the states before its own updates,
    then its f->v updates), separate items where another factor updates a
    variable in the same wave or the factor is wide -- run level by level
    in a shuffled order within each level give the wave execution's bits;
    merging levels pairwise does not."""
    gen = fl.generate(3000, 5) if graph == "plain" else _high_degree_graph()
    g = fl.Graph(gen.factors)
    fl.seed(42)
    if mode == "order":
        order = np.random.default_rng(2).integers(0, g.n_edges, size=g.n_edges + 99).astype(np.uint64)
        s = fl.Schedule(g, order=order)
    else:
        s = fl.Schedule(g, mode)
    kind, a, b, c, lev = s.exec_items()
    lev = lev.astype(np.int64)
    assert lev.min() >= 1 and lev.max() == s.info().n_levels
    assert (kind == 2).any()
    if graph == "wide":
        assert (kind == 0).any() and (kind == 1).any()
    if mode == "order":
        assert (kind == 0).any()   # v->f messages of variables another factor updates in the same wave
    waves = oracle.OracleGraph(gen.factors)
    waves.init()
    waves.set_sequence(s.order(), respect_preorder=s.info().respect_preorder)
    levels = oracle.OracleGraph(gen.factors)
    levels.init()
    merged = oracle.OracleGraph(gen.factors)
    merged.init()
    rng = np.random.default_rng(1)
    for sweep in range(5):
        perm = np.lexsort((rng.random(len(lev)), lev))
        assert waves.sweep() == levels.sweep_exec(kind[perm], a[perm], b[perm], c[perm])
        perm2 = np.lexsort((rng.random(len(lev)), (lev + 1) // 2))   # a wrong split
        merged.sweep_exec(kind[perm2], a[perm2], b[perm2], c[perm2])
    for a, b in zip(waves.messages() + (waves.beliefs(),), levels.messages() + (levels.beliefs(),)):
        assert np.array_equal(a, b)
    for a, b in zip(waves.products(), levels.products()):
        assert np.array_equal(a, b)
    assert not np.array_equal(waves.beliefs(), merged.beliefs())


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("mode", ["SEQFIX", "TOPO"])
def test_sequential_bp_exact_on_trees(seed, mode):
    rng = np.random.default_rng(seed)
    flist = random_tree_list(rng)
    if seed % 2:
        flist = fl.clamp(flist, [int(flist.head[1])], [seed % 4 // 2])
    o = oracle.OracleGraph(flist)
    o.init()
    order = o.order_seqfix(oracle.Rng(42 + seed)) if mode == "SEQFIX" else o.order_topo()
    o.set_sequence(order, respect_preorder=(mode == "TOPO"))
    it, md = o.run(1e-15, 300)
    assert md <= 1e-15
    np.testing.assert_allclose(o.beliefs(), exact_marginals(flist), rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", ["SEQFIX", "TOPO"])
def test_sequential_fixed_point_equals_parallel(mode):
    """BP fixed points do not depend on the schedule: where both converge,
    SEQ and PARALL beliefs agree to the tolerance."""
    gen = fl.generate(3000, 4)
    par = oracle.OracleGraph(gen.factors)
    par.init()
    _, md = par.run(1e-13, 5000)
    assert md <= 1e-13
    o = oracle.OracleGraph(gen.factors)
    o.init()
    order = o.order_seqfix(oracle.Rng(42)) if mode == "SEQFIX" else o.order_topo()
    o.set_sequence(order, respect_preorder=(mode == "TOPO"))
    it, md = o.run(1e-13, 5000)
    assert md <= 1e-13
    np.testing.assert_allclose(o.beliefs(), par.beliefs(), rtol=0, atol=1e-9)


def test_in_place_products_match_messages():
    """update_message keeps prod/nz equal (up to scale) to the product of
    each variable's f->v messages with zeros counted."""
    gen = fl.generate(2000, 6)
    o = oracle.OracleGraph(gen.factors)
    o.init()
    o.set_sequence(o.order_seqfix(oracle.Rng(1)))
    for _ in range(4):
        o.sweep()
    fv, _ = o.messages()
    p0, p1, z0, z1 = o.products()
    d = o.dump()
    off = d["var_off"].astype(np.int64)
    for v in range(o.n_vars):
        m = fv[off[v]:off[v + 1]]
        assert z0[v] == np.count_nonzero(m[:, 0] == 0) and z1[v] == np.count_nonzero(m[:, 1] == 0)
        a = np.prod(m[:, 0][m[:, 0] != 0])
        b = np.prod(m[:, 1][m[:, 1] != 0])
        assert abs(p1[v] / (p0[v] + p1[v]) - b / (a + b)) < 1e-12


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("clamped", [False, True])
@pytest.mark.parametrize("d", [1, 2, 3, 4, 6])
def test_sequential_fv_codes_match_enumeration(kind, clamped, d):
    """calcMessageFVFusedKernel (causal_bp_seq.cu:149-459): the PARALL
    arithmetic with the sequential scale().  Where the head-side mass Q is
    zero and P > 0 the sequential scale gives (0, 1) for AND and (1, 0) for
    OR (no 0/0); the P == Q == 0 corner is checked for a normalized result."""
    rng = np.random.default_rng(2000 * kind + 10 * d + clamped)
    masks = [(1.0, 1.0), (0.0, 1.0), (1.0, 0.0), (0.0, 0.0)]
    for trial in range(60):
        m = random_msgs(rng, d, specials=trial % 2 == 1)
        h = int(rng.integers(0, d))
        p = float(rng.random()) if trial % 3 else float(rng.choice([0.0, 1.0]))
        q = float(rng.random()) if trial % 4 else 0.0
        mask = masks[trial % 4] if clamped else (1.0, 1.0)
        for t in range(d):
            code = code_of(kind, clamped, t == h)
            got = oracle.fv_message(code, [a for a, _ in m], [b for _, b in m], t, h, p, q, *mask, seq=True)
            if t != h:
                hx, hy = (m[h][1], m[h][0]) if kind == 1 else (m[h][0], m[h][1])
                mx, my = (mask[1], mask[0]) if kind == 1 else (mask[0], mask[1])
                Q = q * hx * mx + (1 - q) * hy * my
                P = (p - q) * (hx * mx - hy * my)
                if Q == 0:
                    if P > 0:
                        assert got == ((0.0, 1.0) if kind == 1 else (1.0, 0.0))
                    else:
                        assert np.isfinite(got).all() and abs(got[0] + got[1] - 1) < 1e-15
                    continue
            want = exact_fv(kind, m, t, h, p, q, clamped, mask)
            assert close(got, want, rel=1e-12), (kind, clamped, d, t, h, got, want)
