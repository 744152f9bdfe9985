"""GPU parity: the HIP engine against the oracle (the reference's FP64
PARALL arithmetic restated on the CPU) on the same seeded graphs.

The engine keeps the reference's per-message operation order, so marginals
must agree bit for bit (a stronger form of the 1e-9 bar), iteration counts
and the final max belief change must be identical, and the alarm ranking
(marginal descending, label ascending) must match.
"""
import math

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from tests.helpers import oracle_clamp, random_alarm_clamps, same_bits

pytestmark = pytest.mark.gpu

TOL_MARGINAL = 1e-9   # the stated bar; checked besides bitwise equality


def solve_both(flist, tol, maxiter, damping, threads=8):
    g = fl.Graph(flist)
    plan = fl.Plan(g)
    plan.init()
    r = plan.run(tol, maxiter, damping)
    bg = plan.beliefs()
    o = oracle.OracleGraph(flist)
    o.init()
    it, md = o.run(tol, maxiter, damping, threads=threads)
    bo = o.beliefs()
    return g, plan, r, bg, o, it, md, bo


def parity_of(g, plan, r, bg, o, it, md, bo):
    assert_parity(r, bg, it, md, bo)


def assert_parity(r, bg, it, md, bo):
    assert r.iters == it, f"iterations differ: engine {r.iters}, oracle {it}"
    assert np.max(np.abs(bg - bo), initial=0.0) <= TOL_MARGINAL
    assert same_bits(bg, bo), "marginals are not bit-identical"
    assert (math.isinf(md) and math.isinf(r.maxdiff)) or same_bits([r.maxdiff], [md])


def test_2k_variable_graph_plain_fp64():
    """2,000 variables, no damping, tol 1e-6, at most 1000 iterations."""
    gen = fl.generate(2000, 1)
    g, plan, r, bg, o, it, md, bo = solve_both(gen.factors, 1e-6, 1000, 0.0)
    assert_parity(r, bg, it, md, bo)
    assert r.iters < 1000 and md <= 1e-6
    idx = g.index_of(gen.alarms)
    assert np.array_equal(fl.rank_alarms(gen.alarms, bg[idx]), fl.rank_alarms(gen.alarms, bo[idx]))


@pytest.mark.parametrize("seed,n_vars,damping,tol", [
    (2, 5000, 0.3, 1e-9), (3, 20000, 0.0, 1e-9), (4, 777, 0.5, 1e-12), (5, 50000, 0.3, 1e-9)])
def test_seeded_graphs(seed, n_vars, damping, tol):
    gen = fl.generate(n_vars, seed)
    parity_of(*solve_both(gen.factors, tol, 500, damping))


def test_messages_bitwise_after_fixed_sweeps():
    """run(tol, maxIters, hist): exactly n sweeps; every message equal."""
    gen = fl.generate(3000, 11)
    g = fl.Graph(gen.factors)
    plan = fl.Plan(g)
    plan.init()
    plan.run_fixed(7, 0.2)
    fv, vf = plan.messages()
    o = oracle.OracleGraph(gen.factors)
    o.init()
    o.run_fixed(7, 0.2)
    ofv, ovf = o.messages()
    assert same_bits(fv, ofv) and same_bits(vf, ovf)
    assert same_bits(plan.beliefs(), o.beliefs())


def test_high_degree_paths():
    """Factors and variables above the register-path degree (8)."""
    gen = fl.generate(4000, 21, k_min=6, k_max=14, b_min=4, b_max=10, leaf_frac=0.02)
    g = fl.Graph(gen.factors)
    assert np.diff(g.fac_off).max() > 8 and np.diff(g.var_off).max() > 8
    parity_of(*solve_both(gen.factors, 1e-9, 300, 0.3))


@pytest.mark.parametrize("variant", range(6))
def test_clamped_variants(variant):
    """fl_wrapper-style clamps of 1-50 alarms, rebuilt graph."""
    gen = fl.generate(6000, 31)
    rng = np.random.default_rng(100 + variant)
    labels, values = random_alarm_clamps(rng, gen.alarms, 50)
    if variant == 5:   # head, body and leaf clamps mixed, duplicates included
        labels = np.concatenate([labels, [0, 1, labels[0]]]).astype(np.uint64)
        values = np.concatenate([values, [1, 0, values[0]]]).astype(np.uint8)
    ce = fl.clamp(gen.factors, labels, values)
    co = oracle_clamp(gen.factors, labels, values)
    for k in ("type", "head", "body_off", "body", "p", "q", "head_clamped", "mask0", "mask1"):
        assert np.array_equal(getattr(ce, k), getattr(co, k)), k
    parity_of(*solve_both(ce, 1e-9, 500, 0.3 if variant % 2 else 0.0))


def test_contradictory_clamps_and_tiny_graphs():
    # an AND head clamped true while its only body var is clamped false: zero mass
    flist = fl.FactorList.build([(0, 1, [], 0.8), (1, 2, [1], 0.9, 0.0), (0, 3, [], 0.5),
                                 (2, 4, [2, 3], 1.0, 0.0)])
    c = fl.clamp(flist, [1, 2, 4], [0, 1, 1])
    parity_of(*solve_both(c, 1e-9, 50, 0.0))
    single = fl.FactorList.build([(0, 5, [], 0.3)])
    parity_of(*solve_both(single, 1e-9, 50, 0.0))


def test_run_semantics_counter_and_init():
    """_iters persists across run() and init(); old beliefs are not reset."""
    gen = fl.generate(2500, 41)
    g = fl.Graph(gen.factors)
    plan = fl.Plan(g)
    o = oracle.OracleGraph(gen.factors)
    plan.init(); o.init()
    r1 = plan.run(1e-4, 1000, 0.0)
    i1, _ = o.run(1e-4, 1000, 0.0)
    assert r1.iters == i1
    plan.init(); o.init()
    r2 = plan.run(1e-9, 1000, 0.0)
    i2, _ = o.run(1e-9, 1000, 0.0)
    assert r2.iters == i2 and same_bits(plan.beliefs(), o.beliefs())
    # logdomain init (0, 0) messages: the reference's arithmetic, whatever it yields
    plan.init(logdomain=True); o.init(logdomain=True)
    plan.run_fixed(3); o.run_fixed(3)
    assert same_bits(plan.beliefs(), o.beliefs())


def test_1m_variable_graph_full_size_parity():
    """1M variables, damping 0.3, tol 1e-9, at most 500 iterations: the full
    configuration against the oracle (OpenMP)."""
    gen = fl.generate(1_000_000, 2)
    g, plan, r, bg, o, it, md, bo = solve_both(gen.factors, 1e-9, 500, 0.3, threads=16)
    assert_parity(r, bg, it, md, bo)
    assert np.all((bg >= 0) & (bg <= 1))


def solve_variant(flist, tol, maxiter, damping, variant):
    g = fl.Graph(flist)
    plan = fl.Plan(g)
    plan.set_variant(variant)
    plan.init()
    r = plan.run(tol, maxiter, damping)
    return r, plan.beliefs(), plan.messages()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_both_kernel_variants_match_oracle(variant):
    gen = fl.generate(30000, 61, back_frac=0.3)
    g, plan, r, bg, o, it, md, bo = solve_both(gen.factors, 1e-9, 400, 0.3)
    rv, bv, (fv, vf) = solve_variant(gen.factors, 1e-9, 400, 0.3, variant)
    assert rv.iters == it and same_bits(bv, bo)
    ofv, ovf = o.messages()
    assert same_bits(fv, ofv) and same_bits(vf, ovf)


def test_variant_switch_mid_run_keeps_state():
    """Switching the kernel shape between sweeps (the push layout reorders
    the stored v->f messages) continues the same iteration exactly."""
    gen = fl.generate(20000, 62, back_frac=0.2)
    o = oracle.OracleGraph(gen.factors)
    o.init()
    plan = fl.Plan(fl.Graph(gen.factors))
    plan.init()
    segments = ((0, 3, 0.3), (2, 4, 0.3), (3, 2, 0.3), (1, 2, 0.3), (3, 2, 0.0), (3, 2, 0.3), (2, 1, 0.0),
                (3, 1, 0.0), (0, 2, 0.3), (4, 3, 0.3), (3, 1, 0.3), (4, 2, 0.0), (4, 1, 0.3), (0, 1, 0.3),
                (4, 2, 0.3), (2, 1, 0.3), (4, 1, 0.0), (0, 1, 0.0), (4, 2, 0.3), (3, 1, 0.3))
    for variant, n, damping in segments:
        plan.set_variant(variant)
        plan.run_fixed(n, damping)
        o.run_fixed(n, damping)
    fv, vf = plan.messages()
    ofv, ovf = o.messages()
    assert same_bits(plan.beliefs(), o.beliefs())
    assert same_bits(fv, ofv) and same_bits(vf, ovf)


def test_layout_default_follows_footprint():
    """A new plan starts in the scatter layout (variant 3) up to a 64 MB
    footprint, in the state layout (variant 4) above; the scatter and pull
    layouts give the same bits."""
    small = fl.Plan(fl.Graph(fl.generate(5000, 71).factors))
    assert small.variant == 3
    gen = fl.generate(1_600_000, 72)
    g = fl.Graph(gen.factors)
    big = fl.Plan(g)
    assert big.variant == 4 and big.device_bytes > 320 * 2**20
    other = fl.Plan(g)
    other.set_variant(3)
    pull = fl.Plan(g)
    pull.set_variant(0)
    for p in (big, other, pull):
        p.init()
        p.run_fixed(6, 0.3)
    fa, va = big.messages()
    for p in (other, pull):
        assert same_bits(big.beliefs(), p.beliefs())
        fb, vb = p.messages()
        assert same_bits(fa, fb) and same_bits(va, vb)


def test_head_clamps_in_place_equal_rebuilt_graph():
    """Head clamps applied to a plan equal a graph rebuilt by clamp() at
    alarms with no body uses; reset restores the unclamped solve."""
    gen = fl.generate(30000, 81)
    g = fl.Graph(gen.factors)
    idx = g.index_of(gen.alarms)
    free = gen.alarms[g.body_uses(idx) == 0]
    rng = np.random.default_rng(81)
    for trial in range(3):
        labels = rng.choice(free, size=int(rng.integers(1, 40)), replace=False)
        values = rng.integers(0, 2, labels.shape[0]).astype(np.uint8)
        rebuilt = fl.Graph(fl.clamp(gen.factors, labels, values))
        ref = fl.Plan(rebuilt)
        ref.init()
        r_ref = ref.run(1e-9, 300, 0.3)
        plan = fl.Plan(g)
        plan.clamp_vars(g.index_of(labels), values)
        plan.init()
        r = plan.run(1e-9, 300, 0.3)
        assert r.iters == r_ref.iters
        b = plan.beliefs()
        # the rebuilt graph orders variables the same way (labels unchanged)
        assert same_bits(b, ref.beliefs()[rebuilt.index_of(g.var_label)])
    plan.reset_head_clamps()
    plan.reset()
    plan.init()
    r = plan.run(1e-9, 300, 0.3)
    plain = fl.Plan(g)
    plain.init()
    assert r.iters == plain.run(1e-9, 300, 0.3).iters and same_bits(plan.beliefs(), plain.beliefs())
    with pytest.raises(fl.FlbpError):
        prior = np.flatnonzero(g.fac_type == 0)[:1]
        plan.set_head_clamps(prior, np.zeros(1, np.uint8))
    # a head that appears in some factor body needs the rebuild
    used = np.flatnonzero((g.body_uses(np.arange(g.n_vars)) > 0) & (g.n_head_factors(np.arange(g.n_vars)) > 0))
    fac = g.head_factors(used)
    fac = fac[g.fac_type[fac] != 0][:1]
    with pytest.raises(fl.FlbpError):
        plan.set_head_clamps(fac, np.zeros(1, np.uint8))
    with pytest.raises(fl.FlbpError):
        plan.clamp_vars(used[:1], np.zeros(1, np.uint8))


def wide_graph():
    """A leaf used by 3000 ANDs (a variable with > 1024 edges), an OR over
    1500 clauses (a factor with > 1024 slots), and ordinary structure."""
    facs = [(0, 0, [], 0.7), (0, 1, [], 0.6)]
    nxt = 10
    clauses = []
    for i in range(3000):
        c = nxt; nxt += 1
        facs.append((1, c, [0, 1] if i % 3 else [0], 0.9 + 0.0001 * (i % 50), 0.01 * (i % 4)))
        if i < 1500:
            clauses.append(c)
    facs.append((2, nxt, clauses, 1.0, 0.0))
    return fl.FactorList.build(facs)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_wide_factor_and_variable(variant):
    flist = wide_graph()
    g = fl.Graph(flist)
    assert np.diff(g.fac_off).max() > 1024 and np.diff(g.var_off).max() > 1024
    o = oracle.OracleGraph(flist)
    o.init()
    it, md = o.run(1e-12, 200, 0.2)
    rv, bv, (fv, vf) = solve_variant(flist, 1e-12, 200, 0.2, variant)
    assert rv.iters == it and same_bits(bv, o.beliefs())
    ofv, ovf = o.messages()
    assert same_bits(fv, ofv) and same_bits(vf, ovf)


def test_graph_replays_match_direct_launches():
    """Chunks replayed from captured hipGraphs, solve after solve (eager work
    between the replays), give the bits of direct launches and the oracle."""
    gen = fl.generate(20000, 91)
    g = fl.Graph(gen.factors)
    o = oracle.OracleGraph(gen.factors)
    o.init()
    it, md = o.run(1e-9, 300, 0.3, threads=8)
    bo = o.beliefs()
    for variant in (3, 0, 4):
        plan = fl.Plan(g)
        plan.set_variant(variant)
        plan.use_graphs(True)
        for rep in range(3):
            plan.reset()
            plan.init()
            r = plan.run(1e-9, 300, 0.3)
            assert r.iters == it and same_bits(plan.beliefs(), bo), (variant, rep)
        plan.reset()
        plan.init()
        plan.run(1e-14, 7, 0.0)           # undamped chunks, then damped ones replayed
        plan.run(1e-14, 19, 0.3)
        direct = fl.Plan(g)
        direct.set_variant(variant)
        direct.use_graphs(False)
        direct.init()
        direct.run(1e-14, 7, 0.0)
        direct.run(1e-14, 19, 0.3)
        assert same_bits(plan.beliefs(), direct.beliefs())


@pytest.mark.parametrize("logdomain", [False, True])
def test_state_layout_messages_in_and_out(logdomain):
    """The state layout keeps products instead of v->f messages: messages
    read back after any sweep (also straight after init), messages set
    explicitly (both directions, or f->v only, which keeps the current v->f
    messages), and the solve continued afterwards all equal the oracle."""
    gen = fl.generate(15000, 63, back_frac=0.2)
    g = fl.Graph(gen.factors)
    o = oracle.OracleGraph(gen.factors)
    o.init(logdomain)
    plan = fl.Plan(g)
    plan.set_variant(4)
    plan.init(logdomain)
    fv, vf = plan.messages()
    ofv, ovf = o.messages()
    assert same_bits(fv, ofv) and same_bits(vf, ovf)
    plan.run_fixed(3, 0.3)
    o.run_fixed(3, 0.3)
    fv, vf = plan.messages()
    ofv, ovf = o.messages()
    assert same_bits(fv, ofv) and same_bits(vf, ovf) and same_bits(plan.beliefs(), o.beliefs())
    # a second plan takes the state over explicitly and both continue
    other = fl.Plan(g)
    other.set_variant(4)
    other.init(logdomain)
    other.set_messages(fv, vf)
    other.run_fixed(4, 0.3)
    plan.run_fixed(4, 0.3)
    o.run_fixed(4, 0.3)
    ofv, ovf = o.messages()
    for p in (plan, other):
        f2, v2 = p.messages()
        assert same_bits(f2, ofv) and same_bits(v2, ovf)
    assert same_bits(plan.beliefs(), o.beliefs())
    # f->v set alone: the v->f messages stay what the last sweep left
    rng = np.random.default_rng(3)
    pert = ofv.copy()
    pert[rng.integers(0, len(pert), 50)] = [0.25, 0.75]
    plan.set_messages(pert, None)
    o.set_messages(pert, ovf)
    plan.run_fixed(3, 0.3)
    o.run_fixed(3, 0.3)
    f3, v3 = plan.messages()
    ofv, ovf = o.messages()
    assert same_bits(f3, ofv) and same_bits(v3, ovf) and same_bits(plan.beliefs(), o.beliefs())


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_fixed_sweeps_run_through_nan_beliefs(variant):
    """run(tol, maxIters, hist) does exactly maxIters sweeps (causal_bp.cu:
    935-938), whatever the beliefs: after a logdomain init every belief is
    NaN from the first sweep on, and no later sweep may be skipped or
    repeated (the prior messages' damped sums count the sweeps: 1 - d^k)."""
    gen = fl.generate(15000, 63, back_frac=0.2)
    g = fl.Graph(gen.factors)
    o = oracle.OracleGraph(gen.factors)
    o.init(True)
    plan = fl.Plan(g)
    plan.set_variant(variant)
    plan.init(True)
    for n in (3, 1, 5):
        plan.run_fixed(n, 0.3)
        o.run_fixed(n, 0.3)
        fv, vf = plan.messages()
        ofv, ovf = o.messages()
        assert same_bits(fv, ofv) and same_bits(vf, ovf), n
        assert same_bits(plan.beliefs(), o.beliefs())
    assert np.isnan(plan.beliefs()).any()
    assert abs(fv[0].sum() - (1 - 0.3 ** 9)) < 1e-12
