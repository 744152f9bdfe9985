"""The Bingo protocol on the GPU: flbp_bingo sessions (clamps, several BP
rounds each drawing the next SEQFIX shuffle, low-pass queries) against the
oracle's 4-argument run on the same clamped graphs and orders.  With
--digits 17 every printed number must equal the oracle's exactly; with the
reference's default 6 significant digits the text must equal the oracle's
value printed the same way."""
import os
import subprocess

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from tests.helpers import oracle_clamp

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINGO = os.path.join(ROOT, "fastlbp_amd", "bin", "flbp_bingo")


class OracleSession:
    """The session the reference's bingo_wrapper would run, on the oracle."""

    def __init__(self, flist, seed):
        self.flist = flist
        self.g = fl.Graph(flist)
        self.rng = oracle.Rng(seed)
        self.clamps = {}
        self.low = None

    def bp(self, tol, lo, hi, hist):
        idx = sorted(self.clamps)
        if idx:
            c = oracle_clamp(self.flist, self.g.var_label[idx], [int(self.clamps[i]) for i in idx])
        else:
            c = self.flist
        o = oracle.OracleGraph(c)
        o.set_sequence(o.order_seqfix(self.rng))   # one CausalBP construction per BP
        o.init()
        yet, self.low, reason, _ = o.run_hist(tol, lo, hi, hist)
        return yet, reason


def run_session(path, seed, script, digits):
    args = [BINGO, path, str(seed)] + (["--digits", str(digits)] if digits else [])
    r = subprocess.run(args, input=script, capture_output=True, text=True, timeout=300, check=True)
    return r.stdout.split("\n")


def fmt(x, digits):
    """std::ostream's default float format at a precision: C's %.<digits>g."""
    return format(float(x), f".{digits}g")


@pytest.mark.parametrize("digits", [17, None])
def test_session_matches_oracle(tmp_path, digits):
    gen = fl.generate(3000, 14)
    path = str(tmp_path / "g.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    flist = fl.read_causal_fg(path)
    g = fl.Graph(flist)
    alarms = g.index_of(gen.alarms[:6])
    queries = [int(x) for x in alarms] + [0, 1, 2, g.n_vars - 1]
    session = OracleSession(flist, 11)
    lines = []
    want = []
    def bp(tol, lo, hi, hist):
        lines.append(f"BP {tol} {lo} {hi} {hist}")
        yet, _ = session.bp(tol, lo, hi, hist)
        want.append(yet)
        for q in queries:
            lines.append(f"Q {q}")
            want.append(session.low[q])
    def clamp(i, v):
        lines.append(f"O {i} {'true' if v else 'false'}")
        session.clamps[i] = v
    bp(1e-9, 20, 400, 5)                      # converges: low-pass = beliefs
    clamp(int(alarms[0]), True)
    clamp(int(alarms[1]), False)
    clamp(0, True)                            # a body variable: the graph is rebuilt
    bp(1e-6, 2, 6, 3)                         # stops early: low-pass = history mean
    lines.append(f"UC {int(alarms[1])}")
    del session.clamps[int(alarms[1])]
    bp(1e-12, 3, 40, 4)
    lines.append("NL")
    lines.append("BYE")
    out = run_session(path, 11, "\n".join(lines) + "\n", digits)
    got = [l for l in out if l and not l.startswith(("O ", "UC "))]
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert a == fmt(b, digits or 6), (a, b)
    assert "" in out   # NL


def test_head_clamps_equal_the_reference_in_place_clamp(tmp_path):
    """Clamps of variables with no body uses only: there the reference's
    un-rebuilt clamped copy is consistent and is the rebuilt graph, so this
    is the session the reference itself defines."""
    gen = fl.generate(2500, 15)
    path = str(tmp_path / "h.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    flist = fl.read_causal_fg(path)
    g = fl.Graph(flist)
    idx = g.index_of(gen.alarms)
    idx = [int(i) for i in idx[g.body_uses(idx) == 0][:8]]
    session = OracleSession(flist, 3)
    lines = [f"O {i} {'true' if k % 2 else 'false'}" for k, i in enumerate(idx)]
    for k, i in enumerate(idx):
        session.clamps[i] = bool(k % 2)
    lines += ["BP 1e-9 10 300 4"] + [f"Q {i}" for i in range(0, g.n_vars, 97)] + ["BYE"]
    yet, _ = session.bp(1e-9, 10, 300, 4)
    out = [l for l in run_session(path, 3, "\n".join(lines) + "\n", 17) if l and not l.startswith("O ")]
    assert out[0] == fmt(yet, 17)
    assert out[1:] == [fmt(session.low[i], 17) for i in range(0, g.n_vars, 97)]


def same_or_both_nan(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


# (tol, minIters, maxIters, histLen): converges; stops on the moved-fraction
# test; runs out of iterations with a history longer than the sweeps
@pytest.mark.parametrize("args", [(1e-9, 20, 400, 5), (1e-5, 3, 40, 5), (1e-300, 10, 6, 3), (1e-300, 2, 3, 9)])
@pytest.mark.parametrize("mode", ["PARALL", "PARALL-damped", "SEQFIX"])
def test_device_history_run_equals_the_oracle(args, mode):
    """flbp_plan_run_hist (history on the device) against the oracle's
    restatement of the 4-argument run: sweeps, exit reason, moved fraction,
    last max-diff and every low-pass belief bit for bit."""
    tol, lo, hi, hist = args
    gen = fl.generate(1200, 8)
    g = fl.Graph(gen.factors)
    damping = 0.3 if mode == "PARALL-damped" else 0.0
    plan = fl.Plan(g)
    o = oracle.OracleGraph(gen.factors)
    if mode == "SEQFIX":
        fl.seed(42)
        s = fl.Schedule(g, "SEQFIX")
        plan.set_schedule(s)
        o.set_sequence(s.order())
    plan.init()
    o.init()
    yet, low, reason, md, iters = plan.run_hist(tol, lo, hi, hist, damping=damping)
    oyet, olow, oreason, omd = o.run_hist(tol, lo, hi, hist, damping)
    assert (iters, reason) == (o.iters, oreason)
    assert same_or_both_nan([yet, md], [oyet, omd])
    assert same_or_both_nan(low, olow)
    assert same_or_both_nan(plan.beliefs(), o.beliefs())


def test_device_history_keeps_non_finite_beliefs_out():
    gen = fl.generate(600, 9)
    plan = fl.Plan(fl.Graph(gen.factors))
    o = oracle.OracleGraph(gen.factors)
    plan.init(logdomain=True)
    o.init(logdomain=True)
    yet, low, reason, md, iters = plan.run_hist(1e-300, 10, 4, 3)
    oyet, olow, oreason, omd = o.run_hist(1e-300, 10, 4, 3)
    assert reason == oreason == 2 and iters == o.iters
    assert np.any(np.isnan(plan.beliefs()))
    assert same_or_both_nan(low, olow) and same_or_both_nan([yet, md], [oyet, omd])
    with pytest.raises(fl.FlbpError):
        plan.run_hist(1e-6, 1, 5, 0)
