"""The 4-argument run and the Bingo protocol on the CPU.

- The oracle's run(tol, minIters, maxIters, histLength) restatement
  (csrc/ext/causal_bp.cu:1006-1143, HostBeliefHistory 28-78) against what
  its definition fixes: it stops after the same sweeps as run() when no
  other exit comes first; its low-pass result is the plain mean of each
  variable's last histLength finite beliefs, taken from independently run
  sweeps; its BIG_FRAC exit comes at the first pass where the fraction of
  variables still moving is below (n - minIters) / (maxIters - minIters);
  non-finite beliefs never enter the history.
- flbp_bingo's protocol paths that need no GPU: echoes of O / UC / NL,
  BYE, and its error exits.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINGO = os.path.join(ROOT, "fastlbp_amd", "bin", "flbp_bingo")


def graph(n=1500, seed=3):
    return fl.generate(n, seed).factors


def sweep_history(flist, n, order=None, logdomain=False):
    """Beliefs before and after each of n plain sweeps."""
    o = oracle.OracleGraph(flist)
    if order is not None:
        o.set_sequence(order)
    o.init(logdomain=logdomain)
    out = [o.beliefs().copy()]
    for _ in range(n):
        o.sweep()
        out.append(o.beliefs().copy())
    return out


@pytest.mark.parametrize("seq", [False, True])
def test_converging_run_matches_plain_run(seq):
    flist = graph()
    a = oracle.OracleGraph(flist)
    b = oracle.OracleGraph(flist)
    if seq:
        order = a.order_seqfix(oracle.Rng(42))
        a.set_sequence(order)
        b.set_sequence(order)
    a.init()
    b.init()
    it, md = a.run(1e-9, 100000)
    yet, low, reason, md2 = b.run_hist(1e-9, 5000, 10000, 4)
    assert reason == 0 and b.iters == it and md2 == md
    assert np.array_equal(low, a.beliefs()) and np.array_equal(b.beliefs(), a.beliefs())


@pytest.mark.parametrize("hist", [1, 3, 7])
def test_low_pass_is_the_mean_of_the_last_beliefs(hist):
    flist = graph(800, 5)
    n_max = 6                                     # passes 0..6 sweep, pass 7 exits (n > maxIters)
    o = oracle.OracleGraph(flist)
    o.init()
    yet, low, reason, _ = o.run_hist(1e-300, 10, n_max, hist)
    assert reason == 2 and o.iters == n_max + 1
    hs = sweep_history(flist, n_max + 1)[1:]     # beliefs after sweeps 1..n_max+1
    want = np.mean(np.stack(hs[-hist:]), axis=0)
    np.testing.assert_allclose(low, want, rtol=1e-15, atol=0)


def test_big_fraction_exit_rule():
    flist = graph(1200, 8)
    lo, hi, tol = 3, 40, 1e-5
    hs = sweep_history(flist, hi + 2)
    frac = [np.count_nonzero(np.abs(hs[k] - hs[k + 1]) > tol) / len(hs[0]) for k in range(len(hs) - 1)]
    md = [np.max(np.abs(hs[k] - hs[k + 1])) for k in range(len(hs) - 1)]
    # first pass n (after n sweeps) that exits, by the reference's tests in order
    expect = None
    for n in range(1, hi + 2):
        node = (n - lo) / (hi - lo) if n >= lo else 0.0
        if md[n - 1] <= tol:
            expect = (n, 0)
            break
        if n > lo and frac[n - 1] < node:
            expect = (n, 1)
            break
    assert expect is not None and expect[1] == 1, expect
    o = oracle.OracleGraph(flist)
    o.init()
    yet, low, reason, _ = o.run_hist(tol, lo, hi, 5)
    assert (o.iters, reason) == expect
    assert yet == frac[expect[0] - 1]


def test_non_finite_beliefs_stay_out_of_the_history():
    """logdomain init: zero messages make most beliefs 0/0 (NaN), and they
    stay NaN; NaN beliefs never enter the history, which keeps each
    variable's last histLength finite beliefs (none: the mean is 0)."""
    flist = graph(600, 9)
    hist, n_max = 3, 4
    hs = sweep_history(flist, n_max + 1, logdomain=True)[1:]
    assert np.any(np.isnan(hs[0])) and np.any(np.isfinite(hs[0])) and np.all(np.isnan(hs[-1]))
    o = oracle.OracleGraph(flist)
    o.init(logdomain=True)
    yet, low, reason, _ = o.run_hist(1e-300, 10, n_max, hist)
    assert reason == 2
    want = np.zeros(len(low))
    for v in range(len(low)):
        vals = [h[v] for h in hs if np.isfinite(h[v])][-hist:]
        want[v] = np.mean(vals) if vals else 0.0
    np.testing.assert_allclose(low, want, rtol=1e-15, atol=0)


def test_history_length_zero_is_rejected():
    o = oracle.OracleGraph(graph(100, 1))
    o.init()
    with pytest.raises(RuntimeError):
        o.run_hist(1e-6, 1, 5, 0)


# ------------------------------------------------------------ the REPL

def bingo(args, script, check=False):
    return subprocess.run([BINGO] + args, input=script, capture_output=True, text=True, timeout=60, check=check)


def write_graph(tmp_path, n=300, seed=2):
    path = str(tmp_path / "g.causal_fg")
    fl.write_causal_fg(fl.generate(n, seed).factors, path)
    return path


def test_echoes_and_bye_without_a_gpu(tmp_path):
    path = write_graph(tmp_path)
    r = bingo([path, "7"], "O 3 true\nO 5 false\nUC 3\nNL\nBYE\nO 9 true\n", check=True)
    assert r.stdout == "O 3 true\nO 5 false\nUC 3\n\n"


def test_protocol_errors_exit_nonzero(tmp_path):
    path = write_graph(tmp_path)
    assert bingo([path], "UC 4\n").returncode != 0            # not clamped
    assert bingo([path], "O 4 yes\n").returncode != 0         # bad value
    assert bingo([path], "Q 1\n").returncode != 0             # no BP yet
    assert bingo([path], "O 99999999 true\n").returncode != 0  # no such variable
    assert bingo([path], "XYZ\n").returncode != 0
    other = str(tmp_path / "g.fg")
    open(other, "w").write("1\n")
    assert bingo([other], "BYE\n").returncode != 0           # only .causal_fg graphs
    assert bingo([], "").returncode != 0
