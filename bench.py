#!/usr/bin/env python3
"""Benchmark: plain FP64 loopy BP on the 1M-variable causal graph, one GPU.

One step = one full solve of the graph already resident in HBM: init()
(messages to (1,1)) and run() to convergence (tol 1e-9, damping 0.3, at most
500 sweeps).  value = edge-message updates/s = sum over steps of
2 * E * sweeps / wall time of the timed steps.

Prints ONE JSON line with the metric, a roofline object for the dominant
kernel (algorithmic bytes per launch / mean launch time from HIP events on
the engine's stream, against MI355X's 8 TB/s) and a CPU baseline (the
oracle, an FP64 restatement of the reference's PARALL sweep, timed on a
bounded number of sweeps of the same graph on this host's cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
    python bench.py --workload batched   # 512 graphs + 256 clamped variants vs sequential solves
    python bench.py --workload adjoint   # 100k-variable tied-rule gradient, backward vs forward
    python bench.py --workload fit --steps 20  # tied-rule fit over 64 clamped variants of a 20M graph
    python bench.py --workload seq [--updates TOPO]   # SEQFIX / TOPO solve of the 1M graph vs PARALL
    python bench.py --workload bingo     # Bingo protocol: BP command latency on the 1M graph
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "msg updates/s + % HBM roofline; batched & adjoint vs plain LBP, 1 GPU"
HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(name="plain-fp64-1M", n_vars=1_000_000, seed=2, damping=0.3, tol=1e-9, maxiter=500)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(n_vars, n_factors, n_edges, damping):
    """Algorithmic HBM bytes per launch of each sweep kernel.

    factor pass: read v->f pair 16 E, write f->v pair 16 E, (+16 E old f->v
                 when damped), per factor offset 4 + meta 4 + p 8 + q 8.
    var pass:    slot map 4 E, gather f->v 16 E, scatter v->f 16 E, per
                 variable offset 4 + old belief 8 + new belief 8.
    """
    fac = (32 + (16 if damping else 0)) * n_edges + 24 * n_factors
    var = 36 * n_edges + 20 * n_vars
    return fac, var


def load_traffic(n_edges, variant):
    """Measured HBM bytes per launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), if
    it was taken on this graph (same edge count) with the same kernels."""
    path = os.path.join(ROOT, "profiles", "traffic_plain_fp64_1M.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("n_edges") != n_edges or t.get("variant") != variant:
        return None
    return t


def cpu_baseline(flist, damping, budget_s=15.0):
    import oracle
    threads = len(os.sched_getaffinity(0))
    env_t = os.environ.get("OMP_NUM_THREADS")
    if env_t and env_t.isdigit():
        threads = min(threads, int(env_t))
    g = oracle.OracleGraph(flist)
    g.init()
    g.sweep(damping, threads)          # warm
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        g.sweep(damping, threads)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": 2.0 * g.n_edges * n / dt, "unit": "msg-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} PARALL sweeps (damping {damping}) of the same {g.n_vars}-variable graph, "
                      f"{dt:.1f} s, OpenMP over {threads} threads"}


def build_graphs_parallel(specs, workers=16):
    """specs: callables returning a Graph; run in a thread pool (the native
    generator and builder release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda f: f(), specs))


def batch768_workload(variants_per_group=32):
    """The 768-graph batch: 512 graphs with V log-uniform in [1k, 200k] (generator seeds
    10000 + g) plus 256 clamped variants of one 200k-variable graph (seed
    99000), each clamping 1-50 of its alarms to 0/1 (alarms head no factor
    body, so a clamp keeps the structure).  Returns (generated lists,
    graphs, base graph, clamps, batch groups)."""
    import numpy as np
    import fastlbp_amd as fl
    from fastlbp_amd.batch import Group, split_variants
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(7)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(200000), 512)).astype(np.int64)
    with ThreadPoolExecutor(16) as ex:
        gens = list(ex.map(lambda a: fl.generate(int(a[1]), 10_000 + a[0]), enumerate(sizes)))
        graphs = list(ex.map(lambda gn: fl.Graph(gn.factors), gens))
    base = fl.generate(200_000, 99_000)
    bg = fl.Graph(base.factors)
    aidx = bg.index_of(base.alarms)
    clamps = []
    for _ in range(256):
        n = int(rng.integers(1, 51))
        pick = rng.choice(len(aidx), size=n, replace=False)
        clamps.append((aidx[pick], rng.integers(0, 2, n).astype(np.uint8)))
    groups = [Group(g) for g in graphs] + split_variants(bg, clamps, variants_per_group)
    return gens, graphs, bg, clamps, groups


def run_batch768(variants_per_group=32, tune_cache=None, with_cpu=False, cpu_seconds=15.0, batch_only=False,
                 batch_opts=None):
    """Sequential plain solves (one plan per graph, one plan for the 256
    variants with their clamps applied in place; each solve init + run,
    timed alone) against one batched solve (init excluded).  batch_only:
    the batched solve alone, with the default tile bounds (for counter
    runs: no sequential solves, no autotune trials)."""
    import numpy as np
    import torch
    import fastlbp_amd as fl
    from fastlbp_amd.batch import Batch

    t0 = time.perf_counter()
    gens, graphs, bg, clamps, groups = batch768_workload(variants_per_group)
    log(f"[bench] batch: built {len(graphs)} graphs + 256 variants of V={bg.n_vars} in {time.perf_counter() - t0:.1f}s; "
        f"sum V={sum(g.n_vars for g in graphs) + 256 * bg.n_vars} "
        f"sum E={sum(g.n_edges for g in graphs) + 256 * bg.n_edges}")
    tol, maxiter, damp = 1e-9, 500, 0.3
    plans = [] if batch_only else [fl.Plan(g) for g in graphs]
    vplan = None if batch_only else fl.Plan(bg)
    t_seq, seq_iters = 0.0, []
    seq_bel = {}
    sample = (0, 255, 511, 512, 512 + 100, 512 + 255)
    torch.cuda.synchronize()
    for i, p in enumerate(plans):
        p.reset()
        t1 = time.perf_counter()
        p.init()
        seq_iters.append(p.run(tol, maxiter, damp).iters)
        t_seq += time.perf_counter() - t1
        if i in sample:
            seq_bel[i] = p.beliefs()
    for k, (v, x) in enumerate([] if batch_only else clamps):
        vplan.reset_head_clamps()
        vplan.clamp_vars(v, x)
        vplan.reset()
        t1 = time.perf_counter()
        vplan.init()
        seq_iters.append(vplan.run(tol, maxiter, damp).iters)
        t_seq += time.perf_counter() - t1
        if 512 + k in sample:
            seq_bel[512 + k] = vplan.beliefs()
    del plans, vplan
    b = Batch(groups)
    best = b.opts() if batch_only else b.autotune(damp, trial_sweeps=8, cache=tune_cache)
    if batch_opts:
        b.set_opts(**batch_opts)
        best = b.opts()
    log(f"[bench] batch: batch opts {best}, device bytes {b.device_bytes / 1e9:.2f} GB")
    b.init()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, md, tot = b.run(tol, maxiter, damp)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t0
    vgraphs = graphs + [bg] * 256
    msgs = sum(2.0 * g.n_edges * int(it) for g, it in zip(vgraphs, iters))
    alg = sum((84.0 * g.n_edges + 24.0 * g.n_factors + 20.0 * g.n_vars) * int(it) for g, it in zip(vgraphs, iters))
    if batch_only:
        log(f"[bench] batch: batched {t_bat * 1e3:.1f} ms, {msgs / t_bat:.4g} msg-updates/s (no sequential solves)")
        return {"seconds": t_bat, "msg_updates_per_s": msgs / t_bat, "alg_bytes": alg,
                "frac_hbm": alg / t_bat / 1e9 / HBM_PEAK_GBS, "max_sweeps": int(iters.max()),
                "n_variants": len(vgraphs), "variants_per_group": variants_per_group, "batch_opts": best,
                "device_bytes": b.device_bytes, "sum_edges": int(sum(g.n_edges for g in vgraphs))}
    same = list(iters) == seq_iters and all(np.array_equal(b.beliefs(i), bel) for i, bel in seq_bel.items())
    log(f"[bench] batch: sequential {t_seq * 1e3:.1f} ms, batched {t_bat * 1e3:.1f} ms, "
        f"speedup {t_seq / t_bat:.2f}x, {msgs / t_bat:.4g} msg-updates/s, identical {same}")
    out = {"seconds": t_bat, "sequential_seconds": t_seq, "speedup_vs_sequential": t_seq / t_bat,
           "msg_updates_per_s": msgs / t_bat, "sequential_msg_updates_per_s": msgs / t_seq,
           "alg_bytes": alg, "frac_hbm": alg / t_bat / 1e9 / HBM_PEAK_GBS,
           "iterations_identical": bool(same), "max_sweeps": int(iters.max()), "n_variants": len(vgraphs),
           "variants_per_group": variants_per_group, "batch_opts": best, "device_bytes": b.device_bytes,
           "sum_edges": int(sum(g.n_edges for g in vgraphs))}
    if with_cpu:
        import oracle
        threads = cpu_threads()
        done_msgs, n_done, t0 = 0.0, 0, time.perf_counter()
        for gen in gens:   # the batch's graphs in order, solved to convergence
            if time.perf_counter() - t0 >= cpu_seconds:
                break
            o = oracle.OracleGraph(gen.factors)
            o.init()
            it_o, _ = o.run(tol, maxiter, damp, threads)
            done_msgs += 2.0 * o.n_edges * it_o
            n_done += 1
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done_msgs / cdt, "unit": "msg-updates/s", "cores": threads, "kind": "port",
                               "sample": f"the batch's first {n_done} graphs solved to convergence, {cdt:.1f} s, "
                                         f"OpenMP over {threads} threads"}
    return out


def cpu_threads():
    threads = len(os.sched_getaffinity(0))
    env_t = os.environ.get("OMP_NUM_THREADS")
    if env_t and env_t.isdigit():
        threads = min(threads, int(env_t))
    return threads


def bench_batched(args):
    """The 768-graph batch alone: one JSON line for the batched solve."""
    r = run_batch768(args.variants_per_group, args.tune_cache, not args.no_cpu and not args.batch_only, args.cpu_seconds,
                     args.batch_only, json.loads(args.batch_opts) if args.batch_opts else None)
    if args.batch_only:
        print(json.dumps({"batched": r}), flush=True)
        return
    cpu = r.pop("cpu_baseline", None)
    roofline = {"bound": "hbm", "kernel": "batched factor + var passes (whole run)",
                "achieved": round(r["alg_bytes"] / r["seconds"] / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(r["frac_hbm"], 4), "traffic": None, "algorithmic_bytes": r["alg_bytes"],
                "timing": "wall time of the batched run (init excluded) over the algorithmic bytes of every "
                          "sweep of every variant, 84 E + 24 F + 20 V per damped sweep"}
    out = {"metric": METRIC, "value": r["msg_updates_per_s"], "unit": "msg-updates/s", "n_gpus": 1, "steps": 1,
           "warmup": 0, "ms_per_step": r["seconds"] * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "batched-512-graphs+256-variants", "n_variants": r["n_variants"],
                      "sum_edges": r["sum_edges"], "damping": 0.3, "tol": 1e-9, "maxiter": 500,
                      "variants_per_group": r["variants_per_group"], "batch_opts": r["batch_opts"]},
           "batched": r, "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def run_adjoint(steps=10, warmup=2):
    """The tied-rule adjoint: 100k variables, 200 tied rules, 1000 alarms, T = 50, damping 0.5,
    checkpoint every 10: forward and backward device time (HIP events,
    median over the timed steps)."""
    import numpy as np
    import fastlbp_amd as fl
    from fastlbp_amd.adjoint import TiedProblem

    gen = fl.generate(100_000, 4, n_rules=200, n_alarms=1000)
    rng = np.random.default_rng(4)
    theta = rng.normal(2.5, 0.3, 200)
    y = rng.integers(0, 2, len(gen.alarms)).astype(np.float64)
    g = fl.Graph(gen.factors)
    prob = TiedProblem(g, gen.rule, 200, gen.alarms, 50, 10, 0.5)
    fw, bw = [], []
    for i in range(warmup + steps):
        b = prob.marginals(theta)
        prob.grad(2.0 * (b - y))
        f_ms, b_ms = prob.adj.last_timing()
        if i >= warmup:
            fw.append(f_ms)
            bw.append(b_ms)
    f_ms, b_ms = float(np.median(fw)), float(np.median(bw))
    T, k = 50, 10
    b_iter = 84.0 * g.n_edges + 24.0 * g.n_factors + 20.0 * g.n_vars   # one damped sweep
    alg_fwd = T * b_iter + -(-T // k) * 32.0 * g.n_edges              # sweeps + checkpoints
    alg_bwd = 3.0 * T * b_iter                                           # recompute + reverse
    log(f"[bench] adjoint: forward {f_ms:.3f} ms, backward {b_ms:.3f} ms, ratio {b_ms / f_ms:.2f}")
    return {"forward_ms": f_ms, "backward_ms": b_ms, "backward_over_forward": b_ms / f_ms,
            "gradients_per_s": 1e3 / (f_ms + b_ms), "device_bytes": prob.adj.device_bytes,
            "n_vars": g.n_vars, "n_factors": g.n_factors, "n_edges": g.n_edges, "T": T, "checkpoint_every": k,
            "alg_bytes_forward": alg_fwd, "alg_bytes_backward": alg_bwd,
            "msg_updates_per_s_forward": 2.0 * g.n_edges * T / (f_ms * 1e-3),
            "_problem": (gen, theta, y)}


def bench_adjoint(args):
    """The tied-rule adjoint alone: one JSON line for the adjoint."""
    r = run_adjoint(args.steps, args.warmup)
    gen, theta, y = r.pop("_problem")
    f_ms, b_ms = r["forward_ms"], r["backward_ms"]
    achieved = r["alg_bytes_backward"] / (b_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": "adjoint backward (segment recompute + reverse sweeps)",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "algorithmic_bytes": r["alg_bytes_backward"],
                "forward_achieved": round(r["alg_bytes_forward"] / (f_ms * 1e-3) / 1e9, 1),
                "timing": "HIP events around the adjoint's forward and backward (median of the timed steps); "
                          "bytes: T sweeps of 84 E + 24 F + 20 V forward plus 32 E per checkpoint, three times "
                          "the sweeps backward"}
    cpu = None
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        oracle.grad(gen.factors, gen.rule, 200, theta, gen.alarms, y, r["T"], 0.5)
        cdt = time.perf_counter() - t0
        cpu = {"value": 1.0 / cdt, "unit": "gradients/s", "cores": 1, "kind": "port",
               "sample": f"one orc_grad (unrolled forward + reverse mode, T = {r['T']}) of the same problem, "
                         f"{cdt:.2f} s, one thread"}
        log(f"[bench] cpu baseline: one gradient in {cdt:.2f} s (1 thread); GPU {f_ms + b_ms:.2f} ms")
    out = {"metric": METRIC, "value": r["msg_updates_per_s_forward"], "unit": "msg-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": f_ms + b_ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "adjoint-100k-T50", "n_vars": r["n_vars"], "n_factors": r["n_factors"],
                      "n_edges": r["n_edges"], "n_rules": 200, "n_alarms": int(len(gen.alarms)), "T": r["T"],
                      "checkpoint_every": r["checkpoint_every"], "damping": 0.5},
           "adjoint": r, "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_fit(args):
    """Tied-rule fitting over clamped variants of one large graph, on one GPU.

    A generated graph (20M variables by default, 200 tied AND rules, 1,000
    alarms) and 64 variants, each clamping 1-50 alarms without body uses to
    0 or 1.  The variants run in micro-batches of K: one plan holds K
    interleaved copies of the graph, each clamped as one variant, and one
    adjoint differentiates the T-sweep damped solve of all K copies
    (fastlbp_amd.variants.VariantBatch).  Targets are the variants' T-sweep
    alarm marginals at a reference theta; the fit starts from a perturbed
    theta.  A gradient step runs, per micro-batch, the forward and the
    backward, sums dL/dp over all 64 variants on the device (sparse dL/db
    in) and reduces it to dL/dtheta once, then takes an Adam step.
    Afterwards the batched solver (flbp_batch, K-interleaved groups, each
    variant stopping on its own) runs plain LBP to convergence at the fitted
    theta for all 64 variants, against the T-sweep marginals."""
    import gc
    import numpy as np
    import fastlbp_amd as fl
    from fastlbp_amd.adjoint import tied_p
    from fastlbp_amd.batch import Batch, split_variants
    from fastlbp_amd.variants import VariantBatch

    n_vars = args.n_vars or 20_000_000
    n_var, K = args.fit_variants, args.fit_micro
    if n_var % K:
        raise SystemExit("fit: the variant count must be a multiple of the micro-batch size")
    T, k, damp, n_rules = args.fit_sweeps, args.fit_checkpoint, args.fit_damping, 200
    rng = np.random.default_rng(5)
    t0 = time.perf_counter()
    gen = fl.generate(n_vars, 5, n_rules=n_rules, n_alarms=1000)
    g = fl.Graph(gen.factors)
    aidx = g.index_of(gen.alarms)
    free = gen.alarms[g.body_uses(aidx) == 0]
    variants = []
    for _ in range(n_var):
        m = int(rng.integers(1, min(51, len(free) + 1)))
        lab = np.sort(rng.choice(free, size=m, replace=False)).astype(np.uint64)
        variants.append((lab, rng.integers(0, 2, m).astype(np.uint8)))
    micro = [variants[i:i + K] for i in range(0, n_var, K)]
    log(f"[bench] fit graph V={g.n_vars} F={g.n_factors} E={g.n_edges}, {n_var} variants in micro-batches of {K}, "
        f"generated in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    vb = VariantBatch(gen.factors, gen.rule, n_rules, gen.alarms, K, T, k, damp)
    log(f"[bench] fit {K} interleaved copies: E={vb.graph.n_edges}, built and uploaded in "
        f"{time.perf_counter() - t0:.1f}s, device bytes {vb.device_bytes / 1e9:.1f} GB "
        f"(plan {vb.plan.device_bytes / 1e9:.1f} + adjoint {vb.adj.device_bytes / 1e9:.1f})")

    def forward_all(theta):
        vb.set_theta(theta)
        out = []
        for mb in micro:
            vb.set_clamps(mb)
            out.append(vb.forward())
        return np.concatenate(out)

    theta_ref = rng.normal(2.5, 0.3, n_rules)
    targets = forward_all(theta_ref)
    theta = theta_ref + rng.normal(0.0, 0.5, n_rules)
    m1, m2 = np.zeros(n_rules), np.zeros(n_rules)
    lr, b1, b2, eps = 0.05, 0.9, 0.999, 1e-8
    losses, fwd_ms, bwd_ms = [], 0.0, 0.0
    t_fit = time.perf_counter()
    for step in range(args.steps):
        vb.set_theta(theta)
        vb.clear_grad()
        loss = 0.0
        for i, mb in enumerate(micro):
            vb.set_clamps(mb)
            b = vb.forward()
            y = targets[i * K:(i + 1) * K]
            vb.backward(2.0 * (b - y))   # dL/dp of the K variants added on the device
            a, c = vb.last_timing()
            fwd_ms += a
            bwd_ms += c
            loss += float(np.sum((b - y) ** 2))
        grad = vb.rule_grad()
        m1 = b1 * m1 + (1 - b1) * grad
        m2 = b2 * m2 + (1 - b2) * grad * grad
        mh, vh = m1 / (1 - b1 ** (step + 1)), m2 / (1 - b2 ** (step + 1))
        theta = theta - lr * mh / (np.sqrt(vh) + eps)
        losses.append(loss)
        log(f"[bench] fit step {step}: loss {loss:.6g} ({time.perf_counter() - t_fit:.1f}s)")
    t_fit = time.perf_counter() - t_fit
    bT = forward_all(theta)   # T-sweep marginals at the fitted theta
    dev_bytes = vb.device_bytes
    vb.close()
    del vb
    gc.collect()
    # plain LBP to convergence at the fitted theta, all variants batched
    t0 = time.perf_counter()
    g_fit = fl.Graph(gen.factors.with_p(tied_p(gen.factors.p, gen.rule, theta)))
    clamps = [(g_fit.index_of(lab), val) for lab, val in variants]
    batch = Batch(split_variants(g_fit, clamps, args.fit_eval_group))
    batch.init()
    iters, _, tot = batch.run(1e-9, 500, damp)
    errs = np.array([np.max(np.abs(batch.beliefs_at(v, aidx) - bT[v])) for v in range(n_var)])
    t_eval = time.perf_counter() - t0
    log(f"[bench] fit plain LBP at the fitted theta: {tot.seconds:.2f}s batched ({n_var} variants, "
        f"median {np.median(iters):.0f} sweeps, max {iters.max()}), max |T-sweep - converged| {errs.max():.3g}")
    sweeps = args.steps * n_var * T
    out = {"metric": METRIC, "value": 2.0 * g.n_edges * sweeps / t_fit, "unit": "msg-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 0, "ms_per_step": t_fit * 1e3 / max(args.steps, 1),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"tied-fit-{g.n_vars}-x{n_var}", "n_vars": g.n_vars, "n_factors": g.n_factors,
                      "n_edges": g.n_edges, "n_variants": n_var, "micro_batch": K, "n_rules": n_rules,
                      "n_alarms": int(len(aidx)), "T": T, "checkpoint_every": k, "damping": damp,
                      "optimizer": "adam lr 0.05"},
           "fit": {"steps_per_s": args.steps / t_fit, "loss_first": losses[0] if losses else None,
                  "loss_last": losses[-1] if losses else None, "losses": losses,
                  "forward_ms_per_micro_batch": fwd_ms / max(1, args.steps * len(micro)),
                  "backward_ms_per_micro_batch": bwd_ms / max(1, args.steps * len(micro)),
                  "plain_vs_T_sweep_max_abs_err": float(errs.max()),
                  "plain_vs_T_sweep_mean_max_abs_err": float(errs.mean()),
                  "plain_iters_median": float(np.median(iters)), "plain_iters_max": int(iters.max()),
                  "plain_batched_seconds": tot.seconds, "plain_eval_seconds": t_eval,
                  "device_bytes_fit": dev_bytes, "device_bytes_eval": batch.device_bytes}}
    print(json.dumps(out), flush=True)


def seq_level_bytes(graph, sched):
    """Algorithmic HBM bytes of the level kernels of one sequential sweep.

    v->f item: item 16, f->v message 16, product + counts 32, write 16.
    f->v item: item 32, the factor's d v->f messages 16 d, the factor's
               p / q / mask record 32, old message 16, product + counts 32,
               and writes of product, counts and message 48.
    beliefs:   product + counts 32, old belief 8, new belief 8 per variable.
    """
    import numpy as np
    fo, fe, vo, vs = sched.waves()
    deg = np.diff(graph.fac_off.astype(np.int64))
    slot_fac = np.repeat(np.arange(graph.n_factors), deg)
    d = deg[slot_fac[graph.var_edge[fe.astype(np.int64)].astype(np.int64)]]
    level = 80 * len(vs) + int(np.sum(160 + 16 * d))
    return level, 48 * graph.n_vars


def bench_seq(args):
    """SEQFIX / TOPO: the reference's sequential schedule on the 1M graph,
    run as dependency levels; beside it the PARALL solve of the same graph
    at the same tolerance (undamped: the sequential schedules have no
    damping)."""
    import torch
    import fastlbp_amd as fl

    n_vars = args.n_vars or WORKLOAD["n_vars"]
    tol, maxiter = WORKLOAD["tol"], WORKLOAD["maxiter"]
    gen = fl.generate(n_vars, WORKLOAD["seed"])
    graph = fl.Graph(gen.factors)
    V, F, E = graph.n_vars, graph.n_factors, graph.n_edges
    t0 = time.perf_counter()
    fl.seed(42)
    sched = fl.Schedule(graph, args.updates)
    t_sched = time.perf_counter() - t0
    info = sched.info()
    log(f"[bench] {args.updates}: V={V} E={E} waves {info.n_waves} levels {info.n_levels} "
        f"f->v items {info.n_fv_items} v->f items {info.n_vf_items} (schedule built in {t_sched:.2f}s)")
    t0 = time.perf_counter()
    plan = fl.Plan(graph, device=0)
    t_plan = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan.set_schedule(sched)
    t_attach = time.perf_counter() - t0
    log(f"[bench] plan upload {t_plan:.2f}s, schedule attached (items laid out and uploaded) in {t_attach:.2f}s")
    plan.use_graphs(args.graphs)

    def solve():
        plan.reset()
        plan.init()
        return plan.run(tol, maxiter)

    for i in range(args.warmup):
        r = solve()
        log(f"[bench] warmup {i}: {r.sweeps} sweeps {r.seconds * 1e3:.2f} ms maxdiff {r.maxdiff:.3g}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sweeps = 0
    for i in range(args.steps):
        r = solve()
        sweeps += r.sweeps
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    per_sweep = info.n_fv_items + info.n_vf_items
    value = per_sweep * sweeps / dt
    ms_solve = dt * 1e3 / args.steps
    log(f"[bench] {args.steps} solves, {sweeps / args.steps:.1f} sweeps each, {ms_solve:.3f} ms/solve, "
        f"{value:.4g} msg-updates/s, {dt * 1e6 / max(sweeps, 1):.1f} us/sweep")
    plan.profile(True)
    lvl_ms = bel_ms = 0.0
    n_sw = 0
    for i in range(3):
        solve()
        a, b, na, nb = plan.last_timing()
        lvl_ms += a; bel_ms += b; n_sw += na
    plan.profile(False)
    lb, bb = seq_level_bytes(graph, sched)
    lvl_us = lvl_ms * 1e3 / max(n_sw, 1)
    bel_us = bel_ms * 1e3 / max(n_sw, 1)
    achieved = lb / (lvl_us * 1e-6) / 1e9
    log(f"[bench] levels {lvl_us:.1f} us/sweep ({info.n_levels} launches, {achieved:.0f} GB/s), "
        f"beliefs {bel_us:.1f} us/sweep")
    # PARALL on the same graph, same tolerance, undamped and damped
    par = {}
    for damping in (0.0, WORKLOAD["damping"]):
        pp = fl.Plan(graph, device=0)
        pp.init()
        r = pp.run(tol, maxiter, damping)
        t0 = time.perf_counter()
        for i in range(3):
            pp.reset()
            pp.init()
            r = pp.run(tol, maxiter, damping)
        torch.cuda.synchronize()
        par[str(damping)] = {"sweeps": r.sweeps, "maxdiff": r.maxdiff,
                             "ms_per_solve": (time.perf_counter() - t0) * 1e3 / 3}
        pp.close()
        log(f"[bench] PARALL damping {damping}: {r.sweeps} sweeps, maxdiff {r.maxdiff:.3g}, "
            f"{par[str(damping)]['ms_per_solve']:.2f} ms/solve")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_seq_seqfix_1M.json")
    if args.updates == "SEQFIX" and os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("n_edges") == E:   # measured on this graph
            traffic = t.get("seq_level")
    out = {"metric": METRIC, "value": value, "unit": "msg-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_solve,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"{args.updates.lower()}-fp64-{n_vars // 1000}k", "n_vars": V, "n_factors": F,
                      "n_edges": E, "tol": tol, "maxiter": maxiter, "seed": WORKLOAD["seed"],
                      "updates": args.updates, "waves": info.n_waves, "levels": info.n_levels,
                      "fv_items_per_sweep": info.n_fv_items, "vf_items_per_sweep": info.n_vf_items,
                      "sweeps_per_solve": sweeps / args.steps, "last_maxdiff": r.maxdiff,
                      "graphs": bool(args.graphs),
                      "step": "init + run() to convergence under the sequential schedule"},
           "roofline": {"bound": "hbm", "kernel": "seq_level (all levels of a sweep)",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "algorithmic_bytes_per_sweep": lb, "level_us_per_sweep": round(lvl_us, 2),
                        "belief_us_per_sweep": round(bel_us, 2),
                        "timing": "HIP events around the level launches of each sweep in profiled solves"},
           "parall_same_graph": par, "cpu_baseline": None}
    if not args.no_cpu:
        import oracle
        o = oracle.OracleGraph(gen.factors)
        o.set_sequence(sched.order(), respect_preorder=info.respect_preorder)
        o.init()
        o.sweep()
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            o.sweep()
            n += 1
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": per_sweep * n / cdt, "unit": "msg-updates/s", "cores": 1, "kind": "port",
                               "sample": f"{n} {args.updates} sweeps (wave by wave, one thread) of the same "
                                         f"{V}-variable graph, {cdt:.1f} s"}
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.4g} msg-updates/s (1 thread)")
    print(json.dumps(out), flush=True)


def bench_bingo(args):
    """The Bingo protocol (flbp_bingo, the reference's bingo_wrapper.cu):
    BP command latency on the 1M-variable graph.  The session: NL (answered
    once the file is read and the graph built), BP with no clamps, clamps on
    ten alarms, BP again, Q on those alarms, BYE.  Each BP builds a new
    CausalBP (updates SEQFIX) on the clamped graph and runs init + the
    4-argument run; the tool logs its sweeps and its construct and run
    times.  The oracle runs the first BP's solve (same shuffle, wave by
    wave, one thread) and its yet-to-converge fraction must equal the tool's
    at 17 digits."""
    import re
    import subprocess
    import tempfile
    import fastlbp_amd as fl

    n_vars = args.n_vars or WORKLOAD["n_vars"]
    tol, min_it, max_it, hist = 1e-6, 50, 500, 10
    gen = fl.generate(n_vars, WORKLOAD["seed"])
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "bingo.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    graph = fl.Graph(gen.factors)
    V, F, E = graph.n_vars, graph.n_factors, graph.n_edges
    alarms = [int(i) for i in graph.index_of(gen.alarms[:10])]   # the tool takes label ranks
    errf = open(os.path.join(tmp, "bingo.log"), "w+")
    t_start = time.perf_counter()
    proc = subprocess.Popen([os.path.join(ROOT, "fastlbp_amd", "bin", "flbp_bingo"), path, "42", "--auto",
                             "--digits", "17"], stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=errf,
                            text=True, bufsize=1)

    def ask(line):
        t0 = time.perf_counter()
        proc.stdin.write(line + "\n")
        proc.stdin.flush()
        out = proc.stdout.readline()
        if not out:
            raise RuntimeError(f"flbp_bingo ended at {line!r}")
        return out.strip(), time.perf_counter() - t0

    ask("NL")
    t_load = time.perf_counter() - t_start
    yet1, t_bp1 = ask(f"BP {tol} {min_it} {max_it} {hist}")
    q1 = [ask(f"Q {a}")[0] for a in alarms]
    for a in alarms:
        ask(f"O {a} true")
    yet2, t_bp2 = ask(f"BP {tol} {min_it} {max_it} {hist}")
    qs = [ask(f"Q {a}")[0] for a in alarms]
    proc.stdin.write("BYE\n")
    proc.stdin.flush()
    proc.wait(timeout=60)
    errf.seek(0)
    logs = errf.read()
    done = [tuple(float(x) for x in m) for m in
            re.findall(r"BP done: (\d+) sweeps over \d+ edges; construct \+ init ([\d.]+) ms, run ([\d.]+) ms", logs)]
    if proc.returncode != 0 or len(done) != 2:
        raise RuntimeError(f"flbp_bingo failed (rc {proc.returncode}):\n{logs[-2000:]}")
    (sw1, c1, r1), (sw2, c2, r2) = done
    log(f"[bench] bingo: V={V} E={E}; file read + graph build {t_load:.2f}s")
    log(f"[bench] BP (no clamps): {sw1:.0f} sweeps, {t_bp1 * 1e3:.1f} ms (construct + init {c1:.1f} ms, run {r1:.1f} ms), "
        f"yet-to-converge {yet1}")
    log(f"[bench] BP (10 alarms clamped): {sw2:.0f} sweeps, {t_bp2 * 1e3:.1f} ms (construct + init {c2:.1f} ms, "
        f"run {r2:.1f} ms), yet-to-converge {yet2}")
    value = 2.0 * E * sw1 / (r1 * 1e-3)
    out = {"metric": METRIC, "value": value, "unit": "msg-updates/s", "n_gpus": 1, "steps": 2, "warmup": 0,
           "ms_per_step": t_bp1 * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"bingo-seqfix-fp64-{n_vars // 1000}k", "n_vars": V, "n_factors": F, "n_edges": E,
                      "bp": [tol, min_it, max_it, hist], "seed": 42, "load_s": t_load,
                      "bp_ms": [t_bp1 * 1e3, t_bp2 * 1e3], "construct_init_ms": [c1, c2], "run_ms": [r1, r2],
                      "sweeps": [sw1, sw2], "yet_to_converge": [yet1, yet2], "q_alarms": qs,
                      "value": "2 E x sweeps / run time of the first BP (the 4-argument run)"},
           "roofline": None, "cpu_baseline": None}
    if not args.no_cpu:
        import oracle
        flist = fl.read_causal_fg(path)        # the factors exactly as the tool read them
        fl.seed(42)
        sched = fl.Schedule(fl.Graph(flist), "SEQFIX")   # the shuffle the tool's first BP drew
        o = oracle.OracleGraph(flist)
        o.set_sequence(sched.order())
        o.init()
        t0 = time.perf_counter()
        yet_o, low_o, _, _ = o.run_hist(tol, min_it, max_it, hist)
        cdt = time.perf_counter() - t0
        same = format(yet_o, ".17g") == yet1 and [format(low_o[a], ".17g") for a in alarms] == q1
        out["cpu_baseline"] = {"value": 2.0 * E * sw1 / cdt, "unit": "msg-updates/s", "cores": 1, "kind": "port",
                               "sample": f"the first BP's 4-argument run ({sw1:.0f} SEQFIX sweeps, wave by wave, "
                                         f"one thread), {cdt:.1f} s; yet-to-converge fraction and the "
                                         f"alarms' low-pass beliefs {'equal' if same else 'DIFFERENT'} to the "
                                         f"tool's at 17 digits"}
        log(f"[bench] oracle: first BP {cdt:.1f} s, yet-to-converge {yet_o!r}; fraction and queried beliefs "
            f"{'equal to' if same else 'DIFFERENT from'} the tool's")
    errf.close()
    import shutil
    shutil.rmtree(tmp, ignore_errors=True)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="plain workload only: skip the batched and adjoint summaries of the default line")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=["plain", "batched", "adjoint", "fit", "seq", "bingo"], default="plain")
    ap.add_argument("--updates", choices=["SEQFIX", "TOPO"], default="SEQFIX", help="seq: the schedule")
    ap.add_argument("--graphs", action="store_true", help="replay chunks of sweeps from captured hipGraphs")
    ap.add_argument("--fit-variants", type=int, default=64, help="fit: clamped variants per gradient step")
    ap.add_argument("--variants-per-group", type=int, default=32,
                    help="batched: variants of one graph per K-interleaved group (1, 2, 4, ..., 64)")
    ap.add_argument("--tune-cache", default=None, help="batched: autotune cache file (JSON, by size class)")
    ap.add_argument("--batch-opts", default=None,
                    help='batched: batch options as JSON, set after the autotune (e.g. \'{"chunk": 0}\')')
    ap.add_argument("--batch-only", action="store_true",
                    help="batched: the batched solve alone, default tile bounds (for counter runs; prints a "
                         "summary line, not the bench line)")
    ap.add_argument("--fit-sweeps", type=int, default=40, help="fit: sweeps T of the differentiated solve")
    ap.add_argument("--fit-checkpoint", type=int, default=8, help="fit: checkpoint interval of the adjoint")
    ap.add_argument("--fit-damping", type=float, default=0.3, help="fit: damping of the differentiated solve")
    ap.add_argument("--fit-micro", type=int, default=4, help="fit: variants per micro-batch (interleaved copies)")
    ap.add_argument("--fit-eval-group", type=int, default=16,
                    help="fit: variants per K-interleaved group of the converged evaluation")
    ap.add_argument("--variant", type=int, default=-1,
                    help="sweep kernels: -1 the plan's own choice by footprint (default), 0 LDS-staged (pull), "
                         "1 one lane per item, 2 LDS-staged (push), 3 LDS-staged (scatter), 4 LDS-staged (state)")
    ap.add_argument("--n-vars", type=int, default=0,
                    help="plain workload at another graph size (a sizing study, not the headline workload)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy a 512 MiB buffer once, so HBM counters can be scaled by a known size")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(it runs one GPU's share of a sharded job); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 536870912 bytes read, 536870912 bytes written")
        del src, dst
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "fit":
        return bench_fit(args)
    if args.workload == "seq":
        return bench_seq(args)
    if args.workload == "bingo":
        return bench_bingo(args)

    import torch
    import fastlbp_amd as fl

    w = dict(WORKLOAD)
    if args.n_vars:
        w.update(n_vars=args.n_vars, name=f"plain-fp64-{args.n_vars}")
    t0 = time.perf_counter()
    gen = fl.generate(w["n_vars"], w["seed"])
    graph = fl.Graph(gen.factors)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = fl.Plan(graph, device=0)
    if args.variant >= 0:
        plan.set_variant(args.variant)
    t_up = time.perf_counter() - t0
    V, F, E = graph.n_vars, graph.n_factors, graph.n_edges
    log(f"[bench] graph V={V} F={F} E={E} generate+build {t_gen:.2f}s upload {t_up:.2f}s "
        f"device bytes {plan.device_bytes / 1e6:.1f} MB")

    def step():
        plan.reset()   # a fresh solve each step: counter and old beliefs as constructed
        plan.init()
        return plan.run(w["tol"], w["maxiter"], w["damping"])

    for i in range(args.warmup):
        r = step()
        log(f"[bench] warmup {i}: {r.sweeps} sweeps {r.seconds * 1e3:.2f} ms")
    # the timed solves: no profiling events between the kernels
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sweeps = 0
    for i in range(args.steps):
        r = step()
        sweeps += r.sweeps
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    iters_per_solve = sweeps / max(args.steps, 1)
    value = 2.0 * E * sweeps / dt
    log(f"[bench] {args.steps} solves, {iters_per_solve:.1f} sweeps each, {dt * 1e3 / args.steps:.3f} ms/solve, "
        f"{value:.4g} msg-updates/s")
    # per-kernel launch times from HIP events on the engine's stream, in
    # separate profiled solves of the same workload
    plan.profile(True)
    fac_ms = var_ms = 0.0
    fac_n = var_n = 0
    for i in range(max(3, min(args.steps, 5))):
        step()
        a, b, na, nb = plan.last_timing()
        fac_ms += a; var_ms += b; fac_n += na; var_n += nb
    plan.profile(False)
    fb, vb = kernel_bytes(V, F, E, w["damping"])
    fac_us = fac_ms * 1e3 / max(fac_n, 1)
    var_us = var_ms * 1e3 / max(var_n, 1)
    log(f"[bench] factor pass {fac_us:.2f} us/launch ({fb / fac_us / 1e3:.0f} GB/s), "
        f"var pass {var_us:.2f} us/launch ({vb / var_us / 1e3:.0f} GB/s)")
    if var_us >= fac_us:
        dom, dbytes, dus = "var_pass", vb, var_us
    else:
        dom, dbytes, dus = "factor_pass", fb, fac_us
    achieved = dbytes / (dus * 1e-6) / 1e9
    traffic = load_traffic(E, plan.variant)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic.get(dom) if traffic else None,
                "algorithmic_bytes_per_launch": dbytes, "mean_launch_us": round(dus, 3),
                "sweep_bytes": fb + vb, "sweep_us": round(fac_us + var_us, 3),
                "sweep_frac": round((fb + vb) / ((fac_us + var_us) * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                "timing": "HIP events on the engine's stream around every launch of profiled solves of the same "
                          "workload, run after the timed solves (events between kernels would slow the timed ones); "
                          "each launch's interval includes its dispatch gap"}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(gen.factors, w["damping"], args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.4g} msg-updates/s on {cpu['cores']} threads")
    variant = plan.variant
    extra = {}
    if not args.plain_only and not args.n_vars:
        # the other two headline forms, measured in the same run: the
        # 768-graph batched solve against sequential plain solves, and the
        # 100k-variable tied-rule adjoint
        del plan
        b768 = run_batch768(args.variants_per_group, args.tune_cache)
        extra["batched"] = {k: b768[k] for k in ("speedup_vs_sequential", "seconds", "sequential_seconds",
                                                "msg_updates_per_s", "iterations_identical", "n_variants",
                                                "variants_per_group", "sum_edges", "frac_hbm", "max_sweeps")}
        extra["batched"]["workload"] = "512 graphs (V log-uniform in [1k, 200k]) + 256 clamped variants of a " \
                                       "200k-variable graph, damping 0.3, tol 1e-9, maxiter 500"
        ad = run_adjoint()
        ad.pop("_problem")
        extra["adjoint"] = {k: ad[k] for k in ("backward_over_forward", "forward_ms", "backward_ms",
                                               "gradients_per_s", "n_vars", "n_edges", "T", "checkpoint_every")}
        extra["adjoint"]["workload"] = "100k variables, 200 tied rules, 1000 alarms, damping 0.5"
    out = {"metric": METRIC, "value": value, "unit": "msg-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": w["name"], "n_vars": V, "n_factors": F, "n_edges": E,
                      "damping": w["damping"], "tol": w["tol"], "maxiter": w["maxiter"],
                      "seed": w["seed"], "sweeps_per_solve": iters_per_solve,
                      "kernels": ["lds-staged-pull", "lane-per-item", "lds-staged-push", "lds-staged-scatter",
                                  "lds-staged-state"][variant],
                      "step": "init + run() to convergence of the resident graph"},
           "roofline": roofline, "cpu_baseline": cpu, **extra}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
