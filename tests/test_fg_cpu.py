"""The generic factor graph of the drop-in API on the CPU: libDAI's .fg reader
(the reference's csrc/lbp/factorgraph.cpp:106-212) and the table index
helpers (csrc/lbp/index.cpp), through flbp_graphdump.

Expected tables follow from the format's definition: a listed entry's
linear index runs over the members as listed, first fastest; the table is
ordered by ascending label, first fastest.
"""
import itertools
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMP = os.path.join(ROOT, "tests", "native", "_bin", "flbp_graphdump")


def dump(path):
    return subprocess.run([DUMP, path], capture_output=True, text=True, timeout=60)


def parse(out):
    lines = out.strip().split("\n")
    V, F = map(int, lines[0].split())
    vars_ = [tuple(map(int, l.split())) for l in lines[1:1 + V]]
    facs = []
    for l in lines[1 + V:1 + V + F]:
        head, table = l.split("|")
        h = list(map(int, head.split()))
        facs.append((h[1:], [float(x) for x in table.split()]))
    return vars_, facs


def table_index(members, dims, li):
    """the table position of file index li: digits over the listed members
    (first fastest), reassembled over ascending labels (first fastest)"""
    digit = {}
    for m, d in zip(members, dims):
        digit[m] = li % d
        li //= d
    out, radix = 0, 1
    for m, d in sorted(zip(members, dims)):
        out += digit[m] * radix
        radix *= d
    return out


def write_fg(path, factors, header="# generated\n"):
    with open(path, "w") as f:
        f.write(header + f"{len(factors)}\n\n")
        for members, dims, entries in factors:
            f.write(f"{len(members)}\n{' '.join(map(str, members))}\n{' '.join(map(str, dims))}\n{len(entries)}\n")
            for li, v in entries:
                f.write(f"{li} {v!r}\n")
            f.write("\n")


def test_hand_worked_permuted_table(tmp_path):
    p = str(tmp_path / "a.fg")
    write_fg(p, [([5, 2], [3, 2], [(0, 1.5), (1, 2.5), (3, 4.0), (5, 7.0)]), ([5], [3], [(0, 0.25), (2, 0.75)])])
    r = dump(p)
    assert r.returncode == 0, r.stderr
    vars_, facs = parse(r.stdout)
    assert vars_ == [(2, 2), (5, 3)]
    # x5 + 3*x2 in the file -> x2 + 2*x5 in the table
    assert facs[0] == ([2, 5], [1.5, 4.0, 2.5, 0.0, 0.0, 7.0])
    assert facs[1] == ([5], [0.25, 0.0, 0.75])


def test_random_tables_round_to_their_positions(tmp_path):
    rng = np.random.default_rng(3)
    factors = []
    dims_of = {}
    for _ in range(12):
        k = int(rng.integers(1, 5))
        members = [int(x) for x in rng.choice(20, size=k, replace=False)]
        dims = [dims_of.setdefault(m, int(rng.integers(2, 5))) for m in members]
        n = int(np.prod(dims))
        lis = sorted(set(int(x) for x in rng.integers(0, n, size=n // 2 + 1)))
        factors.append((members, dims, [(li, float(rng.random())) for li in lis]))
    p = str(tmp_path / "r.fg")
    write_fg(p, factors)
    r = dump(p)
    assert r.returncode == 0, r.stderr
    vars_, facs = parse(r.stdout)
    assert vars_ == sorted(dims_of.items())
    for (members, dims, entries), (labels, table) in zip(factors, facs):
        assert labels == sorted(members)
        want = [0.0] * int(np.prod(dims))
        for li, v in entries:
            want[table_index(members, dims, li)] = v
        assert table == want


def test_index_helpers():
    out = subprocess.run([DUMP, "--index"], capture_output=True, text=True, timeout=60, check=True).stdout
    idx, fwd, rev = [list(map(int, l.split())) for l in out.strip().split("\n")]
    # IndexFor of {x1(2), x3(3)} over {x1(2), x2(2), x3(3)}: x1 + 2*x3 as x1, x2, x3 run (x1 fastest)
    assert idx == [x1 + 2 * x3 for x3, x2, x1 in itertools.product(range(3), range(2), range(2))]
    # Permute of (x7(2), x3(3), x5(2)): table index x3 + 3*x5 + 6*x7
    assert fwd == [table_index([7, 3, 5], [2, 3, 2], li) for li in range(12)]
    assert rev == [table_index([5, 3, 7], [2, 3, 2], li) for li in range(12)]


@pytest.mark.parametrize("text,why", [
    ("2 junk\n", "first line"),
    ("2\n\n2\n1 2\n2 3\n0\n1\n1\n4\n0\n", "dims"),          # label 1 with 2 and 4 states
    ("1\n\n1\nx\n", "member list"),
    ("1\n\n# a comment after a blank line is not skipped\n1\n3\n2\n0\n", "comment"),
])
def test_malformed_files_exit_with_an_error(tmp_path, text, why):
    p = str(tmp_path / "m.fg")
    open(p, "w").write(text)
    r = dump(p)
    assert r.returncode != 0, why
    assert "FATAL" in r.stderr


def test_index_past_the_table_keeps_the_digits_that_fit(tmp_path):
    p = str(tmp_path / "w.fg")
    open(p, "w").write("1\n\n1\n3\n2\n1\n9 0.5\n")
    r = dump(p)
    assert r.returncode == 0
    assert parse(r.stdout)[1] == [([3], [0.0, 0.5])]   # 9 mod 2 = 1


def test_missing_file_exits_with_an_error(tmp_path):
    assert dump(str(tmp_path / "none.fg")).returncode != 0


WRAP = os.path.join(ROOT, "fastlbp_amd", "bin", "flbp_fl_wrapper")
BINGO = os.path.join(ROOT, "fastlbp_amd", "bin", "flbp_bingo")


def small_fg(tmp_path):
    path = str(tmp_path / "g.fg")
    write_fg(path, [([3, 7], [2, 2], [(0, 0.25), (1, 0.5), (3, 0.125)]), ([7], [2], [(0, 0.6), (1, 0.4)])])
    return path


def test_fl_wrapper_reads_a_generic_graph_and_solves_nothing(tmp_path):
    """A file that is not .causal_fg goes to the generic reader
    (fl_wrapper.cu:106-115), and the reference runs no inference on it:
    without observations the empty causal graph reports the single line
    "0 0.5" (fl_wrapper.cu:57-88); an observation ends the program with
    "We don't support clamp for BP" (fl_wrapper.cu:39-42).  No GPU call is
    made on this path."""
    path = small_fg(tmp_path)
    obs = tmp_path / "none.obs"
    obs.write_text("")
    r = subprocess.run([WRAP, path, str(obs)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout == "0 0.5\n"
    obs.write_text("7 true\n")
    r = subprocess.run([WRAP, path, str(obs)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "We don't support clamp for BP" in r.stderr
    bad = tmp_path / "bad.fg"
    bad.write_text("1\n\n2\n3 7\n2 2\nx\n")
    r = subprocess.run([WRAP, str(bad), str(obs)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and r.stdout == ""


def test_bingo_on_a_generic_graph(tmp_path):
    """bingo_wrapper.cu:25-53 on a generic graph: BP prints the
    yet-to-converge fraction 1 and Q prints 0, as no inference runs; O and
    UC echo; a BP with a clamp set ends the program."""
    path = small_fg(tmp_path)
    r = subprocess.run([BINGO, path], input="BP 1e-6 1 10 1\nQ 0\nO 1 true\nUC 1\nBP 1e-6 1 10 1\nNL\nBYE\n",
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout == "1\n0\nO 1 true\nUC 1\n1\n\n"
    r = subprocess.run([BINGO, path], input="O 1 false\nBP 1e-6 1 10 1\n", capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "We don't support clamp for BP" in r.stderr and r.stdout == "O 1 false\n"
