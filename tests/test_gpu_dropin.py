"""The C++ drop-in API on the GPU: CausalBP::run() and the fl_wrapper CLI
(clamp -> rebuild -> run(1e-6, 500, 1)) against the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from tests.helpers import oracle_clamp

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fastlbp_amd", "bin")


def parse_marginals(lines):
    out = {}
    for ln in lines:
        a, b = ln.split()
        out[int(a)] = float(b)
    return out


def test_causal_bp_run_matches_oracle(tmp_path):
    gen = fl.generate(2000, 1)
    path = str(tmp_path / "g2k.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    r = subprocess.run([os.path.join(BIN, "flbp_run"), path, "1e-6", "1000", "0"], check=True,
                       capture_output=True, text=True, timeout=120)
    lines = r.stdout.strip().split("\n")
    iters = int(lines[0].split()[1])
    marg = parse_marginals(lines[2:])
    o = oracle.OracleGraph(fl.read_causal_fg(path))
    o.init()
    it, md = o.run(1e-6, 1000)
    assert iters == it
    assert float(lines[1].split()[1]) == md
    bo = o.beliefs()
    g = fl.Graph(fl.read_causal_fg(path))
    got = np.array([marg[int(l)] for l in g.var_label])
    assert np.array_equal(got, bo)   # %.17g round-trips doubles exactly


def test_fl_wrapper_cli_matches_oracle(tmp_path):
    gen = fl.generate(3000, 8)
    path = str(tmp_path / "g.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    labels = [int(x) for x in gen.alarms[:12]] + [0, 3]
    values = [1, 0] * 7
    obs = tmp_path / "g.obs"
    obs.write_text("".join(f"{l} {'true' if v else 'false'}\n" for l, v in zip(labels, values)))
    r = subprocess.run([os.path.join(BIN, "flbp_fl_wrapper"), path, str(obs), "60"], check=True,
                       capture_output=True, text=True, timeout=120)
    marg = parse_marginals(r.stdout.strip().split("\n"))
    c = oracle_clamp(fl.read_causal_fg(path), labels, values)
    o = oracle.OracleGraph(c)
    o.init()
    o.run_fixed(60)
    bo = o.beliefs()
    labs = o.dump()["var_label"]
    clamps = dict(zip(labels, values))
    expect = {}
    for l, b in zip(labs, bo):
        expect[int(l)] = b
    for l in range(int(labs.max()) + 1):   # fl_wrapper.cu:80-88
        b = expect.get(l, -1.0)
        want = b if 0 <= b <= 1 else (float(clamps[l]) if l in clamps else 0.5)
        assert marg[l] == want, (l, marg[l], want)
    assert len(marg) == int(labs.max()) + 1


@pytest.mark.gpu
def test_maxdiff_member_follows_the_reference(tmp_path):
    """maxDiff() is the member the 4-argument run raises to its last max
    belief change (causal_bp.cu:1101-1102); run() keeps its own maxDiff
    local and leaves the member alone (causal_bp.cu:864-890).  A fresh
    CausalBP holds 0."""
    gen = fl.generate(2000, 3)
    path = str(tmp_path / "g.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    tool = os.path.join(BIN, "flbp_run")
    r = subprocess.run([tool, path, "1e-6", "1000", "0"], check=True, capture_output=True, text=True, timeout=120)
    md = [float(ln.split()[1]) for ln in r.stderr.split("\n") if ln.startswith("maxDiff ")]
    assert md == [0.0, 0.0]
    r = subprocess.run([tool, path, "1e-6", "1000", "0", "PARALL", "0", "5,50,3"], check=True,
                       capture_output=True, text=True, timeout=120)
    md = [float(ln.split()[1]) for ln in r.stderr.split("\n") if ln.startswith("maxDiff ")]
    o = oracle.OracleGraph(fl.read_causal_fg(path))
    o.init()
    _, _, _, hist_md = o.run_hist(1e-6, 5, 50, 3)
    assert md[0] == hist_md and md[1] == hist_md      # run() leaves it alone
    it, last = o.run(1e-6, 1000)
    lines = r.stdout.strip().split("\n")
    assert int(lines[0].split()[1]) == it and float(lines[1].split()[1]) == last
