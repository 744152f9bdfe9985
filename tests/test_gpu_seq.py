"""GPU parity of the sequential schedules (updates = SEQFIX / TOPO): the HIP
level kernels against the oracle's wave-by-wave restatement of
putUpdateSeqToKernelFused + calcMessageVFKernel + calcMessageFVFusedKernel
(csrc/ext/causal_bp.cu:583-722, csrc/kernel/causal_bp_seq.cu) on the same
seeded graphs and the same orders.  Iteration counts, the last max belief
change, every marginal and every message must be bit-identical.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from tests.helpers import oracle_clamp, random_alarm_clamps, same_bits

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fastlbp_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def schedule_for(g, mode, seed=42):
    if mode == "SEQFIX":
        fl.seed(seed)
    return fl.Schedule(g, mode)


def oracle_for(flist, s):
    o = oracle.OracleGraph(flist)
    o.set_sequence(s.order(), respect_preorder=s.info().respect_preorder)
    return o


def assert_same_state(plan, o):
    fv, vf = plan.messages()
    ofv, ovf = o.messages()
    assert same_bits(fv, ofv), "f->v messages differ"
    assert same_bits(vf, ovf), "v->f messages differ"
    assert same_bits(plan.beliefs(), o.beliefs()), "marginals differ"


def run_both(flist, mode, tol, maxiter, chunk=0, tail=None):
    g = fl.Graph(flist)
    s = schedule_for(g, mode)
    plan = fl.Plan(g)
    plan.set_schedule(s)
    if tail is not None:
        plan.set_seq_tail(tail)
    assert plan.updates == s.info().updates and plan.variant == 3
    plan.init()
    r = plan.run(tol, maxiter, chunk=chunk)
    o = oracle_for(flist, s)
    o.init()
    it, md = o.run(tol, maxiter)
    assert r.iters == it, f"iterations differ: engine {r.iters}, oracle {it}"
    assert (math.isinf(md) and math.isinf(r.maxdiff)) or same_bits([r.maxdiff], [md])
    assert_same_state(plan, o)
    return g, plan, o, r


# tail None: the default split (small graphs run almost every level in the
# single-workgroup tail kernel); 0: every level its own launch
@pytest.mark.parametrize("tail", [None, 0])
@pytest.mark.parametrize("mode", ["SEQFIX", "TOPO"])
@pytest.mark.parametrize("seed,n_vars,tol", [(1, 2000, 1e-6), (3, 20000, 1e-9), (4, 777, 1e-12)])
def test_seeded_graphs(mode, seed, n_vars, tol, tail):
    gen = fl.generate(n_vars, seed)
    _, _, _, r = run_both(gen.factors, mode, tol, 500, tail=tail)
    assert r.iters < 500


@pytest.mark.parametrize("tail", [None, 0, 64])
@pytest.mark.parametrize("mode", ["SEQFIX", "TOPO"])
def test_high_degree_factors_and_variables(mode, tail):
    """Factors above the register path (8 slots) and wide variables, mixed
    with group items (factors of at most 6 slots)."""
    gen = fl.generate(4000, 21, k_min=6, k_max=14, b_min=4, b_max=10, leaf_frac=0.02)
    g = fl.Graph(gen.factors)
    deg = np.diff(g.fac_off)
    assert deg.max() > 8 and np.diff(g.var_off).max() > 8 and (deg <= 6).any()
    run_both(gen.factors, mode, 1e-9, 300, tail=tail)


@pytest.mark.parametrize("tail", [None, 0])
@pytest.mark.parametrize("variant", range(3))
def test_clamped_variants(variant, tail):
    gen = fl.generate(6000, 31)
    rng = np.random.default_rng(300 + variant)
    labels, values = random_alarm_clamps(rng, gen.alarms, 50)
    ce = fl.clamp(gen.factors, labels, values)
    run_both(ce, "SEQFIX" if variant % 2 == 0 else "TOPO", 1e-9, 500, tail=tail)


def test_edge_case_graph_and_contradictions():
    flist = fl.read_causal_fg(os.path.join(GOLDEN, "edge_cases.causal_fg"))
    for mode in ("SEQFIX", "TOPO"):
        run_both(flist, mode, 1e-9, 200)
    # zero mass: an AND head clamped true with its body clamped false
    flist = fl.FactorList.build([(0, 1, [], 0.8), (1, 2, [1], 0.9, 0.0), (0, 3, [], 0.5),
                                 (2, 4, [2, 3], 1.0, 0.0)])
    c = fl.clamp(flist, [1, 2, 4], [0, 1, 1])
    for mode in ("SEQFIX", "TOPO"):
        run_both(c, mode, 1e-9, 50)
    run_both(fl.FactorList.build([(0, 5, [], 0.3)]), "SEQFIX", 1e-9, 10)


def test_fixed_sweeps_logdomain_chunks_and_graph_replay():
    """run(tol, maxIters, hist) sweeps; logdomain init (zero messages and
    products: the counts wrap as the reference's size_t does and NaNs
    appear); chunk sizes and hipGraph replay change nothing."""
    gen = fl.generate(3000, 11)
    g = fl.Graph(gen.factors)
    s = schedule_for(g, "SEQFIX", seed=5)
    for logdomain in (False, True):
        o = oracle_for(gen.factors, s)
        plan = fl.Plan(g)
        plan.set_schedule(s)
        plan.init(logdomain=logdomain)
        o.init(logdomain=logdomain)
        plan.run_fixed(6)
        o.run_fixed(6)
        assert_same_state(plan, o)
    ref = None
    for chunk, graphs in ((1, False), (7, False), (0, True)):
        plan = fl.Plan(g)
        plan.set_schedule(s)
        plan.use_graphs(graphs)
        plan.init()
        r = plan.run(1e-10, 400, chunk=chunk)
        out = (r.iters, plan.beliefs(), plan.messages()[0])
        if ref is None:
            ref = out
        else:
            assert out[0] == ref[0] and same_bits(out[1], ref[1]) and same_bits(out[2], ref[2])
    # graph replays again after eager work on the same plan
    plan = fl.Plan(g)
    plan.set_schedule(s)
    plan.use_graphs(True)
    plan.init()
    o = oracle_for(gen.factors, s)
    o.init()
    for k in range(3):
        plan.run_fixed(3)
        o.run_fixed(3)
        plan.beliefs_at([0, 1, 2])
        assert_same_state(plan, o)


def test_explicit_orders_with_repeats():
    gen = fl.generate(2500, 12)
    g = fl.Graph(gen.factors)
    rng = np.random.default_rng(4)
    order = rng.integers(0, g.n_edges, size=g.n_edges + 300).astype(np.uint64)
    for respect in (False, True):
        s = fl.Schedule(g, order=order, respect_preorder=respect)
        plan = fl.Plan(g)
        plan.set_schedule(s)
        assert plan.updates == -1
        plan.init()
        plan.run_fixed(5)
        o = oracle_for(gen.factors, s)
        o.init()
        o.run_fixed(5)
        assert_same_state(plan, o)
        # run to a tolerance: the max belief change of every sweep, from the
        # variables' last updates and from those no update reaches
        plan.reset()
        plan.init()
        r = plan.run(1e-9, 60)
        o = oracle_for(gen.factors, s)
        o.init()
        it, md = o.run(1e-9, 60)
        assert r.iters == it and same_bits([r.maxdiff], [md])
        assert_same_state(plan, o)


def test_in_place_head_clamps_under_a_schedule():
    """In-place head clamps with the schedule kept equal the rebuilt clamped
    graph run by the same order (clamps of variables with no body uses keep
    the graph's structure, so edge positions are unchanged)."""
    gen = fl.generate(5000, 17)
    g = fl.Graph(gen.factors)
    idx = g.index_of(gen.alarms)
    ok = idx[g.body_uses(idx) == 0][:20]
    vals = (np.arange(len(ok)) % 2).astype(np.uint8)
    s = schedule_for(g, "SEQFIX", seed=3)
    plan = fl.Plan(g)
    plan.set_schedule(s)
    plan.clamp_vars(ok, vals)
    plan.init()
    r = plan.run(1e-9, 300)
    c = oracle_clamp(gen.factors, g.var_label[ok], vals)
    o = oracle.OracleGraph(c)
    o.set_sequence(s.order())
    o.init()
    it, md = o.run(1e-9, 300)
    assert r.iters == it
    assert_same_state(plan, o)


def test_switching_schedules_mid_run():
    """PARALL -> SEQFIX -> PARALL on one plan, continuing from the stored
    messages (the products are reset to init's values when a schedule is
    set, as the reference's construct + init would leave them)."""
    gen = fl.generate(3000, 13)
    g = fl.Graph(gen.factors)
    s = schedule_for(g, "SEQFIX", seed=8)
    plan = fl.Plan(g)
    o = oracle.OracleGraph(gen.factors)
    plan.init()
    o.init()
    plan.run_fixed(3)
    o.run_fixed(3)
    plan.set_schedule(s)
    o.set_sequence(s.order())
    o.init()
    plan.init()
    plan.run_fixed(4)
    o.run_fixed(4)
    assert_same_state(plan, o)
    plan.set_schedule(None)
    o.clear_sequence()
    assert plan.updates == 1
    plan.run_fixed(3, 0.3)   # damped PARALL reads the factor-major copy: rebuilt from var-major
    o.run_fixed(3, 0.3)
    assert_same_state(plan, o)


def test_tail_setting_rejects_what_the_kernel_cannot_hold():
    g = fl.Graph(fl.generate(500, 2).factors)
    plan = fl.Plan(g)
    with pytest.raises(fl.FlbpError):
        plan.set_seq_tail(1025)
    plan.set_seq_tail(0)          # before a schedule: kept for the next one
    plan.set_schedule(schedule_for(g, "SEQFIX"))
    plan.init()
    plan.run(1e-9, 100)


def test_messages_under_a_schedule():
    """Export works under a schedule (the f->v messages are kept
    factor-major there); setting f->v messages is refused, v->f accepted;
    without the schedule the plan takes both again."""
    gen = fl.generate(1500, 5)
    g = fl.Graph(gen.factors)
    plan = fl.Plan(g)
    plan.set_schedule(schedule_for(g, "SEQFIX"))
    plan.init()
    plan.run(1e-9, 50)
    fv, vf = plan.messages()
    with pytest.raises(fl.FlbpError):
        plan.set_messages(fv, None)
    plan.set_messages(None, vf)
    assert same_bits(plan.messages()[1], vf) and same_bits(plan.messages()[0], fv)
    plan.set_schedule(None)                     # var-major copy rebuilt for PARALL
    assert same_bits(plan.messages()[0], fv)
    plan.set_messages(fv, vf)
    assert same_bits(plan.messages()[0], fv)


def test_rejections_under_a_schedule():
    gen = fl.generate(500, 2)
    g = fl.Graph(gen.factors)
    plan = fl.Plan(g)
    plan.set_schedule(fl.Schedule(g, "TOPO"))
    plan.init()
    with pytest.raises(fl.FlbpError):
        plan.run(1e-9, 10, damping=0.2)
    with pytest.raises(fl.FlbpError):
        plan.run_fixed(2, 0.1)
    with pytest.raises(fl.FlbpError):
        plan.set_variant(0)
    from fastlbp_amd.adjoint import Adjoint
    with pytest.raises(fl.FlbpError):
        Adjoint(plan, 5, 2, 0.0)
    other = fl.Graph(fl.generate(600, 2).factors)
    with pytest.raises(fl.FlbpError):
        plan.set_schedule(fl.Schedule(other, "TOPO"))


@pytest.mark.parametrize("mode", ["SEQFIX", "TOPO"])
def test_full_size_graph(mode):
    """The benchmarked configuration: the 1M-variable graph (seed 2),
    tol 1e-9, at most 500 sweeps.  Iteration count, last max-diff, every
    marginal and every message equal the oracle's wave-by-wave solve (about
    20 s of oracle time each)."""
    gen = fl.generate(1_000_000, 2)
    _, plan, o, r = run_both(gen.factors, mode, 1e-9, 500)
    assert r.maxdiff <= 1e-9 and r.iters < 500


def test_200k_seqfix_converges_like_the_oracle():
    gen = fl.generate(200000, 9)
    _, plan, o, r = run_both(gen.factors, "SEQFIX", 1e-9, 200)
    assert r.maxdiff <= 1e-9


@pytest.mark.parametrize("mode,skip", [("SEQFIX", 0), ("SEQFIX", 2), ("TOPO", 0)])
def test_causal_bp_drop_in(tmp_path, mode, skip):
    """CausalBP with updates = SEQFIX / TOPO through the C++ drop-in; each
    earlier SEQFIX construction in the process advances rnd_gen."""
    gen = fl.generate(2000, 1)
    path = str(tmp_path / "g.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    r = subprocess.run([os.path.join(BIN, "flbp_run"), path, "1e-9", "1000", "0", mode, str(skip)],
                       check=True, capture_output=True, text=True, timeout=120)
    lines = r.stdout.strip().split("\n")
    iters = int(lines[0].split()[1])
    flist = fl.read_causal_fg(path)
    o = oracle.OracleGraph(flist)
    if mode == "SEQFIX":
        rng = oracle.Rng(42)
        for _ in range(skip):
            o.order_seqfix(rng)
        o.set_sequence(o.order_seqfix(rng))
    else:
        o.set_sequence(o.order_topo(), respect_preorder=True)
    o.init()
    it, md = o.run(1e-9, 1000)
    assert iters == it and float(lines[1].split()[1]) == md
    g = fl.Graph(flist)
    marg = {int(a): float(b) for a, b in (ln.split() for ln in lines[2:])}
    assert same_bits(np.array([marg[int(l)] for l in g.var_label]), o.beliefs())
