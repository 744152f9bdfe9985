"""Binary graph files (flbp_write_graph / flbp_read_graph) on the CPU: exact
round trips of the CSR, and every damaged file refused."""
import os
import struct

import numpy as np
import pytest

import fastlbp_amd as fl
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
FIELDS = ["var_label", "fac_off", "edge_var", "var_off", "var_edge", "fac_type", "head_slot", "head_clamped",
          "p", "q", "mask0", "mask1"]


def same(a, b):
    return all(np.array_equal(getattr(a, k), getattr(b, k)) for k in FIELDS)


@pytest.mark.parametrize("case", ["generated", "clamped", "edge_cases", "single"])
def test_round_trip(tmp_path, case):
    if case == "generated":
        flist = fl.generate(20000, 4).factors
    elif case == "clamped":
        gen = fl.generate(5000, 6)
        flist = fl.clamp(gen.factors, gen.alarms[:20], np.arange(20) % 2)
    elif case == "edge_cases":
        flist = fl.read_causal_fg(os.path.join(GOLDEN, "edge_cases.causal_fg"))
    else:
        flist = fl.FactorList.build([(0, 7, [], 0.25)])
    g = fl.Graph(flist)
    path = str(tmp_path / "g.flbpcsr")
    g.write(path)
    h = fl.Graph.read(path)
    assert same(g, h)
    # the read graph is the reference's CSR of the factor list (oracle's construct)
    d = oracle.OracleGraph(flist).dump()
    assert np.array_equal(h.fac_off, d["fac_off"]) and np.array_equal(h.var_edge, d["var_src"])
    assert np.array_equal(h.var_label, d["var_label"]) and np.array_equal(h.head_slot, d["head_slot"])


def write_small(tmp_path):
    g = fl.Graph(fl.generate(300, 2).factors)
    path = str(tmp_path / "s.flbpcsr")
    g.write(path)
    return g, path, open(path, "rb").read()


def refused(path):
    with pytest.raises(fl.FlbpError) as e:
        fl.Graph.read(path)
    return e.value.code


def test_damaged_files_are_refused(tmp_path):
    g, path, data = write_small(tmp_path)
    bad = str(tmp_path / "bad.flbpcsr")
    for name, blob in [("flipped", data[:500] + bytes([data[500] ^ 4]) + data[501:]),
                       ("truncated", data[:-9]),
                       ("extended", data + b"\0"),
                       ("magic", b"FLBPCSR2" + data[8:]),
                       ("huge header", data[:8] + struct.pack("<Q", 1 << 40) + data[16:]),
                       ("empty", b"")]:
        open(bad, "wb").write(blob)
        assert refused(bad) == -6, name          # FLBP_EIO
    assert refused(str(tmp_path / "missing.flbpcsr")) == -6


def word_hash(data):
    """The file's integrity check, restated to forge a well-formed file."""
    M = (1 << 64) - 1
    h = 0x6a09e667f3bcc909
    n = len(data)
    pad = data + b"\0" * (-n % 8)
    for i in range(0, len(pad), 8):
        h ^= int.from_bytes(pad[i:i + 8], "little")
        h = ((h << 29) | (h >> 35)) & M
        h = (h * 0x9e3779b97f4a7c15) & M
    h ^= 0x8000000000000000 | (n % 8)
    h = ((h << 29) | (h >> 35)) & M
    return (h * 0x9e3779b97f4a7c15) & M


def test_check_matches_its_definition_and_structure_is_validated(tmp_path):
    g, path, data = write_small(tmp_path)
    body, check = data[:-8], int.from_bytes(data[-8:], "little")
    assert word_hash(body) == check
    # a well-formed file whose var_edge is no longer a permutation
    V, F, E = struct.unpack("<3Q", body[8:32])
    off = 32 + 8 * V + 4 * (F + 1) + 4 * E + 4 * (V + 1)
    forged = bytearray(body)
    forged[off:off + 4] = forged[off + 4:off + 8]
    forged = bytes(forged)
    bad = str(tmp_path / "forged.flbpcsr")
    open(bad, "wb").write(forged + word_hash(forged).to_bytes(8, "little"))
    assert refused(bad) == -1                     # FLBP_EINVAL from the CSR checks


def test_graphs_beyond_32_bit_indices_are_refused(tmp_path):
    g = fl.Graph(fl.generate(200, 1).factors)
    import ctypes as C
    c = g.csr_c()
    c.n_edges = 1 << 33
    assert fl.lib().flbp_write_graph(C.byref(c), str(tmp_path / "x").encode()) < 0
