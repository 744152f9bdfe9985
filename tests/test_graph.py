"""Graph construction, the .causal_fg reader and clamping (CPU only).

The engine's flat builder (flbp_graph_build), reader (flbp_read_causal_fg)
and clamp (flbp_clamp) are checked against:
  - the committed dumps of the reference's own graph model
    (tests/golden/*.refdump.txt, written by tests/golden/make_golden.py from
    oracle/_ref/ref_dump, i.e. csrc/ext/causal_fg.cpp, causal_factor.cpp and
    csrc/lbp/bipgraph.cpp compiled from the reference's sources);
  - oracle/_ref/ref_dump itself on freshly generated graphs, when built;
  - the oracle's independent restatement of the same construction.
"""
import glob
import os

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle
from tests.helpers import oracle_clamp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("type", "head", "body_off", "body", "p", "q", "head_clamped", "mask0", "mask1")


def read_obs(path):
    labels, values = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 2:
                labels.append(int(parts[0]))
                values.append(1 if parts[1] == "true" else 0)
    return labels, values


def engine_as_dump(flist):
    """The engine's graph in ref_dump's shape."""
    g = fl.Graph(flist)
    fac_of_slot = np.repeat(np.arange(g.n_factors), np.diff(g.fac_off).astype(np.int64))
    pos_in_var = np.empty(g.n_edges, np.int64)
    for i in range(g.n_vars):
        a, b = int(g.var_off[i]), int(g.var_off[i + 1])
        pos_in_var[g.var_edge[a:b].astype(np.int64)] = np.arange(b - a)
    facs = []
    for I in range(g.n_factors):
        a, b = int(g.fac_off[I]), int(g.fac_off[I + 1])
        facs.append(dict(type=int(g.fac_type[I]), p=float(g.p[I]), q=float(g.q[I]),
                         clamped=int(g.head_clamped[I]), mask0=float(g.mask0[I]), mask1=float(g.mask1[I]),
                         head=int(g.var_label[g.edge_var[g.head_slot[I]]]),
                         nbF=[(int(g.edge_var[s]), int(pos_in_var[s])) for s in range(a, b)]))
    nbv = []
    for i in range(g.n_vars):
        a, b = int(g.var_off[i]), int(g.var_off[i + 1])
        row = []
        for s in g.var_edge[a:b].astype(np.int64):
            I = int(fac_of_slot[s])
            row.append((I, int(s - g.fac_off[I])))
        nbv.append(row)
    return dict(V=g.n_vars, F=g.n_factors, E=g.n_edges, labels=[int(x) for x in g.var_label],
                factors=facs, nbV=nbv)


def assert_same_dump(got, want):
    assert (got["V"], got["F"], got["E"]) == (want["V"], want["F"], want["E"])
    assert got["labels"] == want["labels"]
    for I, (a, b) in enumerate(zip(got["factors"], want["factors"])):
        assert a == b, (I, a, b)
    assert got["nbV"] == want["nbV"]


def golden_cases():
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN, "*.causal_fg"))):
        base = path[: -len(".causal_fg")]
        out.append((path, None, base + ".refdump.txt"))
        if os.path.exists(base + ".obs"):
            out.append((path, base + ".obs", base + ".obs.refdump.txt"))
    return out


@pytest.mark.parametrize("path,obs,dump", golden_cases(),
                         ids=lambda x: os.path.basename(x) if isinstance(x, str) else "-")
def test_engine_matches_reference_graph_model(path, obs, dump):
    with open(dump) as f:
        want = oracle.parse_dump(f.read())
    flist = fl.read_causal_fg(path)
    if obs:
        flist = fl.clamp(flist, *read_obs(obs))
    assert_same_dump(engine_as_dump(flist), want)


@pytest.mark.skipif(not os.path.exists(oracle.REF_DUMP), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_engine_matches_live_reference(tmp_path, seed):
    gen = fl.generate(400 * seed, seed, back_frac=0.3)
    path = str(tmp_path / "g.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    obs = str(tmp_path / "g.obs")
    rng = np.random.default_rng(seed)
    labels = rng.choice(gen.factors.head, size=20, replace=False)
    vals = rng.integers(0, 2, size=20)
    with open(obs, "w") as f:
        for l, v in zip(labels, vals):
            f.write(f"{int(l)} {'true' if v else 'false'}\n")
    assert_same_dump(engine_as_dump(fl.read_causal_fg(path)), oracle.ref_dump(path))
    clamped = fl.clamp(fl.read_causal_fg(path), labels, vals)
    assert_same_dump(engine_as_dump(clamped), oracle.ref_dump(path, obs))


@pytest.mark.parametrize("seed", [1, 5, 9])
def test_builder_matches_oracle_construction(seed):
    gen = fl.generate(3000, seed, back_frac=0.2)
    rng = np.random.default_rng(seed)
    labels = rng.choice(gen.factors.head, size=40, replace=False)
    vals = rng.integers(0, 2, size=40)
    for flist in (gen.factors, fl.clamp(gen.factors, labels, vals)):
        g = fl.Graph(flist)
        d = oracle.OracleGraph(flist).dump()
        assert np.array_equal(d["var_label"], g.var_label)
        assert np.array_equal(d["fac_off"], g.fac_off)
        assert np.array_equal(d["edge_var"], g.edge_var)
        assert np.array_equal(d["var_off"], g.var_off)
        assert np.array_equal(d["var_src"], g.var_edge)
        assert np.array_equal(d["head_slot"], g.head_slot)
    co = oracle_clamp(gen.factors, labels, vals)
    ce = fl.clamp(gen.factors, labels, vals)
    for k in FIELDS:
        assert np.array_equal(getattr(ce, k), getattr(co, k)), k


def test_writer_reader_round_trip(tmp_path):
    gen = fl.generate(5000, 3)
    path = str(tmp_path / "rt.causal_fg")
    fl.write_causal_fg(gen.factors, path)
    back = fl.read_causal_fg(path)
    for k in FIELDS:
        assert np.array_equal(getattr(back, k), getattr(gen.factors, k)), k


def test_reader_errors_are_reported(tmp_path):
    bad = tmp_path / "bad.causal_fg"
    bad.write_text("2\n\n1\nX\n")
    with pytest.raises(fl.FlbpError):
        fl.read_causal_fg(str(bad))
    with pytest.raises(fl.FlbpError):
        fl.read_causal_fg(str(tmp_path / "missing.causal_fg"))


@pytest.mark.skipif(not os.path.exists(oracle.REF_DUMP), reason="oracle/_ref not built")
def test_body_lines_parse_like_the_reference(tmp_path):
    """Body lines as `is >> long` reads them (causal_fg.cpp's body loop):
    tabs and runs of blanks, a leading '+', duplicates, tokens past the
    declared length and trailing junk after it -- against the reference's
    own reader."""
    path = tmp_path / "odd.causal_fg"
    path.write_text("5\n\n"
                    "10\n*0.9;0.05\n2\n\t+3 \t  1\n\n"
                    "11\n+0.7;0.1\n3\n 4 4 2 99 junk\n\n"
                    "12\n*;0.2\n1\n7 and more\n\n"
                    "13\nI\n0.25\n\n"
                    "14\n+\n2\n013 +10\n")
    assert_same_dump(engine_as_dump(fl.read_causal_fg(str(path))), oracle.ref_dump(str(path)))


def test_body_line_shorter_than_declared_is_an_error(tmp_path):
    for body in ("3 x", "3 -", "3 99999999999999999999999"):
        path = tmp_path / "short.causal_fg"
        path.write_text(f"1\n\n10\n*0.9;0.05\n2\n{body}\n")
        with pytest.raises(fl.FlbpError):
            fl.read_causal_fg(str(path))


def test_implicit_singletons_and_label_ranks():
    """Labels that head no factor get Singleton(0.5) appended in ascending
    order (causal_fg.cpp:264-269); var index = rank of label."""
    gen = fl.generate(200, 4)
    g = fl.Graph(gen.factors)
    assert np.all(np.diff(g.var_label.astype(np.int64)) > 0)
    flist = fl.read_causal_fg(os.path.join(GOLDEN, "edge_cases.causal_fg"))
    assert flist.n_factors == 9 and list(flist.head[6:]) == [1, 2, 3]
    assert np.all(flist.type[6:] == 0) and np.all(flist.p[6:] == 0.5)


def test_generator_is_deterministic_and_sized():
    a = fl.generate(10000, 77)
    b = fl.generate(10000, 77)
    for k in FIELDS:
        assert np.array_equal(getattr(a.factors, k), getattr(b.factors, k))
    assert np.array_equal(a.alarms, b.alarms)
    g = fl.Graph(a.factors)
    assert g.n_vars == 10000 and len(a.alarms) == 100
    t = fl.generate(10000, 77, n_rules=200)
    assert t.rule.max() == 199 and np.all(t.rule[t.factors.type != 1] == -1)


@pytest.mark.parametrize("n,seed", [(2000, 1), (100000, 4)])
def test_generator_realizes_the_stated_sizes(n, seed):
    """F = 1.5 V (one defining factor per variable plus V/2 further AND
    rule instances on existing heads) and E ~ 4.5 V; the further factors
    have q = 0.5 and a head of layers 1..; alarms are heads without body
    uses, so every alarm clamp keeps the structure.  With extra_frac = 0 and
    alarm_free = 0 the one-factor-per-variable graph of the same seed comes
    back (F = V)."""
    gen = fl.generate(n, seed)
    g = fl.Graph(gen.factors)
    assert g.n_vars == n and g.n_factors == n + round(0.5 * n)
    assert 4.3 * n < g.n_edges < 4.7 * n
    f = gen.factors
    extra = np.arange(n, f.n_factors)
    assert np.all(f.type[extra] == 1) and np.all(f.q[extra] == 0.5)
    assert np.all((f.p[extra] >= 0.90) & (f.p[extra] <= 0.99))
    heads_of_or = set(int(h) for h in f.head[:n][f.type[:n] == 2])
    assert all(int(h) in heads_of_or for h in f.head[extra])
    a = g.index_of(gen.alarms)
    assert len(gen.alarms) == 100 and np.all(g.body_uses(a) == 0)
    assert np.all(g.n_head_factors(a) >= 1)
    base = fl.generate(n, seed, extra_frac=0.0, alarm_free=0)
    assert base.factors.n_factors == n


def test_rank_alarms_ties_by_label():
    labels = np.array([40, 10, 30, 20], np.uint64)
    marg = np.array([0.5, 0.9, 0.5, float("nan")])
    assert list(fl.rank_alarms(labels, marg)) == [1, 2, 0, 3]
