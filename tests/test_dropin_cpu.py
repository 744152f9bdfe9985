"""The C++ drop-in classes (include/ext, include/lbp) on the CPU: graph
construction, reader and clamp through CausalFactorGraph, compared with the
dumps of the reference's own graph model (tests/golden/*.refdump.txt), and
PropertySet's exact-type lookup."""
import os
import subprocess

import pytest

import oracle
from tests.test_graph import golden_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "native", "_bin")


def run(*args):
    return subprocess.run([os.path.join(BIN, args[0])] + list(args[1:]), check=True,
                          capture_output=True, text=True)


@pytest.mark.parametrize("path,obs,dump", golden_cases(),
                         ids=lambda x: os.path.basename(x) if isinstance(x, str) else "-")
def test_causal_factor_graph_matches_reference_dump(path, obs, dump):
    out = run("flbp_graphdump", path, *([obs] if obs else [])).stdout
    with open(dump) as f:
        assert oracle.parse_dump(out) == oracle.parse_dump(f.read())


def test_property_set_needs_exact_type():
    """opts.set("maxiter", 1000) stores an int: getStringAs<size_t> prints
    "get properties error" and returns 0 (properties.h:44-53)."""
    r = run("flbp_graphdump", "--properties")
    assert r.stdout.split() == ["1000", "0", "9.9999999999999995e-07", "PARALL", "1", "0"]
    assert "get properties error" in r.stderr
