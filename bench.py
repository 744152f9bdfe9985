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


def load_traffic():
    """Measured HBM bytes per launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", "traffic_plain_fp64_1M.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import fastlbp_amd as fl

    w = WORKLOAD
    t0 = time.perf_counter()
    gen = fl.generate(w["n_vars"], w["seed"])
    graph = fl.Graph(gen.factors)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = fl.Plan(graph, device=0)
    t_up = time.perf_counter() - t0
    V, F, E = graph.n_vars, graph.n_factors, graph.n_edges
    log(f"[bench] graph V={V} F={F} E={E} generate+build {t_gen:.2f}s upload {t_up:.2f}s "
        f"device bytes {plan.device_bytes / 1e6:.1f} MB")

    def step():
        plan.init()
        return plan.run(w["tol"], w["maxiter"], w["damping"])

    for i in range(args.warmup):
        r = step()
        log(f"[bench] warmup {i}: {r.sweeps} sweeps {r.seconds * 1e3:.2f} ms")
    plan.profile(True)
    fac_ms = var_ms = 0.0
    fac_n = var_n = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sweeps = 0
    for i in range(args.steps):
        r = step()
        sweeps += r.sweeps
        a, b, na, nb = plan.last_timing()
        fac_ms += a; var_ms += b; fac_n += na; var_n += nb
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    iters_per_solve = sweeps / max(args.steps, 1)
    value = 2.0 * E * sweeps / dt
    log(f"[bench] {args.steps} solves, {iters_per_solve:.1f} sweeps each, {dt * 1e3 / args.steps:.3f} ms/solve, "
        f"{value:.4g} msg-updates/s")

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
    traffic = load_traffic()
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic.get(dom) if traffic else None,
                "algorithmic_bytes_per_launch": dbytes, "mean_launch_us": round(dus, 3),
                "sweep_bytes": fb + vb, "sweep_us": round(fac_us + var_us, 3),
                "sweep_frac": round((fb + vb) / ((fac_us + var_us) * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(gen.factors, w["damping"], args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.4g} msg-updates/s on {cpu['cores']} threads")
    out = {"metric": METRIC, "value": value, "unit": "msg-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": w["name"], "n_vars": V, "n_factors": F, "n_edges": E,
                      "damping": w["damping"], "tol": w["tol"], "maxiter": w["maxiter"],
                      "seed": w["seed"], "sweeps_per_solve": iters_per_solve,
                      "step": "init + run() to convergence of the resident graph"},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
