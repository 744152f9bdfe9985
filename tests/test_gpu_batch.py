"""Batched solve against the oracle: every variant of a batch must stop at
the oracle's iteration count for its graph (the graph rebuilt with the
variant's clamps, examples/fl_wrapper.cu:26-45) with bit-identical beliefs
and last max change, whatever the group sizes, tile bounds, grid size,
launch mode or hipGraph use; and the 768-graph batch (512 graphs with V
log-uniform in [1k, 200k] plus 256 clamped variants of a 200k-variable
graph) as one batch."""
import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from fastlbp_amd.batch import Batch, Group, split_variants
from tests.helpers import oracle_clamp, same_bits

pytestmark = pytest.mark.gpu


def alarm_clamps(rng, g, alarms, max_n):
    """(variable indices, values) of 1..max_n alarms of graph g (alarms have
    no body uses, so the clamps keep the structure)."""
    idx = g.index_of(alarms)
    n = int(rng.integers(1, max_n + 1))
    pick = rng.choice(len(idx), size=min(n, len(idx)), replace=False)
    return idx[pick], rng.integers(0, 2, size=pick.shape[0]).astype(np.uint8)


def oracle_solve(flist, tol, maxiter, damping, threads=8, logdomain=False):
    o = oracle.OracleGraph(flist)
    o.init(logdomain)
    it, md = o.run(tol, maxiter, damping, threads=threads)
    return it, md, o.beliefs()


def variant_flist(flist, g, clamp):
    """the factor list of the variant: the graph clamped at its variables"""
    v, x = clamp
    if len(v) == 0:
        return flist
    return oracle_clamp(flist, g.var_label[np.asarray(v, np.int64)], np.asarray(x, np.uint8))


def mixed_batch(seed, n_graphs, variants_per_group, n_groups):
    """independent graphs (groups of one) followed by n_groups groups of
    clamped variants of one base graph"""
    rng = np.random.default_rng(seed)
    gens, groups = [], []
    for k in range(n_graphs):
        n = int(np.exp(rng.uniform(np.log(300), np.log(20000))))
        gen = fl.generate(n, seed * 1000 + k)
        gens.append((gen.factors, fl.Graph(gen.factors), [((), ())]))
    base = fl.generate(8000, seed * 1000 + 999)
    bg = fl.Graph(base.factors)
    for _ in range(n_groups):
        cl = [alarm_clamps(rng, bg, base.alarms, 40) for _ in range(variants_per_group)]
        cl[0] = ((), ())   # the unclamped graph is a variant too
        gens.append((base.factors, bg, cl))
    for _, g, cl in gens:
        groups.append(Group(g, cl))
    return gens, groups


def check_against_oracle(b, gens, tol, maxiter, damping, iters, md, logdomain=False):
    v = 0
    for flist, g, cl in gens:
        for clamp in cl:
            it, omd, bo = oracle_solve(variant_flist(flist, g, clamp), tol, maxiter, damping, logdomain=logdomain)
            assert iters[v] == it, (v, iters[v], it)
            assert same_bits([md[v]], [omd]) or (np.isinf(md[v]) and np.isinf(omd)), (v, md[v], omd)
            gb = b.beliefs(v)
            rebuilt = fl.Graph(variant_flist(flist, g, clamp))
            pos = rebuilt.index_of(g.var_label[np.isin(g.var_label, rebuilt.var_label)])
            mine = gb[np.isin(g.var_label, rebuilt.var_label)]
            assert same_bits(mine, bo[pos]), v
            v += 1


@pytest.mark.parametrize("opts,K", [(dict(), 4), (dict(tile_factors=256, tile_vars=512, grid_blocks=37), 8),
                                    (dict(use_graph=0), 2), (dict(tile_factors=64, tile_vars=96, grid_blocks=3), 4),
                                    (dict(chunk=5), 16), (dict(chunk=7, use_graph=0), 1), (dict(chunk=0), 4)])
def test_batch_matches_oracle(opts, K):
    gens, groups = mixed_batch(1, 6, K, 2)
    tol, maxiter, damping = 1e-9, 300, 0.3
    b = Batch(groups, **opts)
    b.init()
    iters, md, tot = b.run(tol, maxiter, damping)
    check_against_oracle(b, gens, tol, maxiter, damping, iters, md)
    assert tot.sweeps == int(iters.max())


def test_maxiter_cap_undamped_and_continued_runs():
    gens, groups = mixed_batch(2, 5, 8, 1)
    b = Batch(groups)
    b.init()
    iters, md, _ = b.run(1e-14, 23, 0.0)
    check_against_oracle(b, gens, 1e-14, 23, 0.0, iters, md)
    # a continued run (counters and messages persist), damped now
    iters, md, _ = b.run(1e-9, 200, 0.4)
    v = 0
    for flist, g, cl in gens:
        for clamp in cl:
            o = oracle.OracleGraph(variant_flist(flist, g, clamp))
            o.init()
            o.run(1e-14, 23, 0.0)
            it, omd = o.run(1e-9, 200, 0.4)
            assert iters[v] == it and same_bits([md[v]], [omd]), v
            v += 1


@pytest.mark.parametrize("K", [1, 2, 8])
def test_new_tile_bounds_between_runs(K):
    """New tile bounds between two runs re-cut the tiles, and with them the
    order the factor tiles store their messages in: the stored messages
    move along, so the continued run ends exactly where one run of the
    same sweeps does (the oracle's 9 + rest sweeps)."""
    gens, groups = mixed_batch(5, 4, K, 2)
    b = Batch(groups)
    b.init()
    b.run(1e-14, 9, 0.3)
    b.set_opts(tile_factors=128, tile_vars=256)
    iters, md, _ = b.run(1e-9, 300, 0.3)
    b.set_opts(tile_factors=512, tile_vars=512)
    check_against_oracle(b, gens, 1e-9, 300, 0.3, iters, md)


def test_logdomain_init_fixed_sweeps():
    """logdomain init and tol = -inf (fixed sweeps through NaN beliefs)"""
    gens, groups = mixed_batch(3, 3, 4, 1)
    b = Batch(groups)
    b.init(logdomain=True)
    iters, md, _ = b.run(-np.inf, 6, 0.3)
    v = 0
    for flist, g, cl in gens:
        for clamp in cl:
            o = oracle.OracleGraph(variant_flist(flist, g, clamp))
            o.init(True)
            o.run_fixed(6, 0.3)
            assert iters[v] == 6
            rebuilt = fl.Graph(variant_flist(flist, g, clamp))
            keep = np.isin(g.var_label, rebuilt.var_label)
            assert same_bits(b.beliefs(v)[keep], o.beliefs()[rebuilt.index_of(g.var_label[keep])]), v
            v += 1


def wide_graph():
    """A leaf used by 1500 ANDs (a variable with > 512 edges), an OR over
    700 clauses (a factor with > 512 slots), and ordinary structure."""
    facs = [(0, 0, [], 0.7), (0, 1, [], 0.6)]
    nxt = 2
    clauses = []
    for i in range(1500):
        c = nxt
        nxt += 1
        facs.append((1, c, [0, 1] if i % 3 else [0], 0.9 + 0.00005 * i, 0.01))
        if i < 700:
            clauses.append(c)
    facs.append((2, nxt, clauses, 1.0, 0.0))
    return fl.FactorList.build(facs), nxt


@pytest.mark.parametrize("K", [1, 2, 4, 64])
def test_wide_factor_and_variable_in_groups(K):
    flist, head = wide_graph()
    g = fl.Graph(flist)
    top = g.index_of([head])
    cl = [((), ())] + [(top, np.array([k % 2], np.uint8)) for k in range(K - 1)]
    b = Batch([Group(g, cl)], tile_factors=256, tile_vars=256)
    b.init()
    iters, md, _ = b.run(1e-12, 200, 0.2)
    check_against_oracle(b, [(flist, g, cl)], 1e-12, 200, 0.2, iters, md)


def test_body_used_variables_and_bad_groups_are_rejected():
    gen = fl.generate(3000, 5)
    g = fl.Graph(gen.factors)
    used = np.flatnonzero(g.body_uses(np.arange(g.n_vars)) > 0)[:1]
    with pytest.raises(fl.FlbpError):
        Batch([Group(g, [((), ()), (used, np.zeros(1, np.uint8))])])
    with pytest.raises(ValueError):
        Group(g, [((), ())] * 3)
    for bad in (dict(tile_factors=4096, tile_vars=4096), dict(tile_factors=32), dict(tile_vars=1024), dict(layout=1)):
        with pytest.raises(fl.FlbpError):
            Batch([g], **bad)
    b = Batch([g])
    with pytest.raises(fl.FlbpError):
        b.set_opts(tile_factors=1024)


def test_autotune_keeps_results_and_caches(tmp_path):
    gens, groups = mixed_batch(4, 4, 8, 1)
    b = Batch(groups)
    cache = str(tmp_path / "tune.json")
    best = b.autotune(0.3, trial_sweeps=4, cache=cache)
    assert best["tile_factors"] in (128, 256, 512) and best["tile_vars"] in (128, 256, 512)
    iters, md, _ = b.run(1e-9, 300, 0.3)
    check_against_oracle(b, gens, 1e-9, 300, 0.3, iters, md)
    again = Batch(groups)
    assert again.autotune(0.3, trial_sweeps=4, cache=cache)["tile_factors"] == best["tile_factors"]
    iters2, _, _ = again.run(1e-9, 300, 0.3)
    assert np.array_equal(iters, iters2)


def test_set_p_and_beliefs_at():
    gen = fl.generate(6000, 7, n_rules=20)
    g = fl.Graph(gen.factors)
    rng = np.random.default_rng(7)
    cl = [((), ())] + [alarm_clamps(rng, g, gen.alarms, 10) for _ in range(3)]
    b = Batch([Group(g, cl)])
    p = g.p.copy()
    p[gen.rule >= 0] = 0.7 + 0.2 * rng.random(int(np.sum(gen.rule >= 0)))
    b.set_p(0, p)
    b.init()
    iters, md, _ = b.run(1e-9, 300, 0.3)
    flist = gen.factors.with_p(p)
    aidx = g.index_of(gen.alarms)
    for v, clamp in enumerate(cl):
        it, omd, bo = oracle_solve(variant_flist(flist, g, clamp), 1e-9, 300, 0.3)
        assert iters[v] == it
        assert same_bits(b.beliefs_at(v, aidx), b.beliefs(v)[aidx])


def test_768_graph_batch():
    """512 graphs with V log-uniform in [1k, 200k] (seeds 10000 + g) plus
    256 clamped variants (1-50 alarms each to 0/1) of one 200k-variable graph,
    one batch, damping 0.3, tol 1e-9, at most 500 sweeps.  Every variant
    against the oracle where it solves within the test's budget (the graphs
    below 6k variables and 8 of the variants), the rest against plain GPU
    solves of the same graphs; all bit-identical, with property checks."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(200000), 512)).astype(np.int64)
    with ThreadPoolExecutor(16) as ex:
        gens = list(ex.map(lambda a: fl.generate(int(a[1]), 10_000 + a[0]), enumerate(sizes)))
        graphs = list(ex.map(lambda gn: fl.Graph(gn.factors), gens))
    base = fl.generate(200_000, 99_000)
    bg = fl.Graph(base.factors)
    clamps = [alarm_clamps(rng, bg, base.alarms, 50) for _ in range(256)]
    groups = [Group(g) for g in graphs] + split_variants(bg, clamps, 32)
    b = Batch(groups)
    tol, maxiter, damping = 1e-9, 500, 0.3
    b.init()
    iters, md, tot = b.run(tol, maxiter, damping)
    assert np.all(iters >= 1) and np.all(iters <= maxiter)
    assert np.all((md <= tol) | (iters == maxiter))
    checked_oracle = 0
    for i, (gen, g) in enumerate(zip(gens, graphs)):
        bb = b.beliefs(i)
        assert np.all((bb >= 0) & (bb <= 1))
        if g.n_vars < 6000:
            it, omd, bo = oracle_solve(gen.factors, tol, maxiter, damping, threads=16)
            assert iters[i] == it and same_bits(bb, bo) and same_bits([md[i]], [omd]), i
            checked_oracle += 1
        else:
            p = fl.Plan(g)
            p.init()
            r = p.run(tol, maxiter, damping)
            assert iters[i] == r.iters and same_bits(bb, p.beliefs()) and same_bits([md[i]], [r.maxdiff]), i
    assert checked_oracle >= 50
    plan = fl.Plan(bg)
    for k, (v, x) in enumerate(clamps):
        j = 512 + k
        bb = b.beliefs(j)
        assert np.all((bb >= 0) & (bb <= 1))
        if k < 8:
            it, omd, bo = oracle_solve(variant_flist(base.factors, bg, (v, x)), tol, maxiter, damping, threads=16)
            rebuilt = fl.Graph(variant_flist(base.factors, bg, (v, x)))
            keep = np.isin(bg.var_label, rebuilt.var_label)
            assert iters[j] == it and same_bits(bb[keep], bo[rebuilt.index_of(bg.var_label[keep])]), j
        plan.reset_head_clamps()
        plan.clamp_vars(v, x)
        plan.reset()
        plan.init()
        r = plan.run(tol, maxiter, damping)
        assert iters[j] == r.iters and same_bits(bb, plan.beliefs()) and same_bits([md[j]], [r.maxdiff]), j
