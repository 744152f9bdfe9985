"""Replays damaged input files through a sanitizer build of flbp_graphdump.

usage: python3 tests/native/replay_damaged.py <flbp_graphdump_san> [--seed S] [--keep DIR]

The corpus is made here, from a fixed seed, out of small valid inputs:
  - binary graph files (flbp_write_graph of a generated graph, a clamped
    graph and tests/golden/edge_cases.causal_fg): truncations, bit flips,
    and edits of the header sizes and of the CSR arrays with the integrity
    check recomputed, so that the structure checks behind it are reached;
  - .causal_fg files (tests/golden): truncations and byte / token edits,
    read by the C-ABI reader with graph build and both sequential
    schedules (--schedules) and by the drop-in CausalFactorGraph reader;
  - generic .fg files: truncations and token edits.
Every file must end in a clean exit (status 0, or 1 for an error the reader
reports) with no AddressSanitizer or UndefinedBehaviorSanitizer report.
Exits 1 if any run does not.
"""
import argparse
import os
import random
import shutil
import struct
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)

M64 = (1 << 64) - 1


def word_hash(data: bytes) -> int:
    """The graph file's integrity check: h = rotl(h ^ w, 29) * K over the
    little-endian 64-bit words, the tail zero-padded, then a final word of
    0x8000000000000000 | (tail length)."""
    h = 0x6A09E667F3BCC909

    def word(h, w):
        h ^= w
        h = ((h << 29) | (h >> 35)) & M64
        return (h * 0x9E3779B97F4A7C15) & M64

    n8 = len(data) // 8 * 8
    for (w,) in struct.iter_unpack("<Q", data[:n8]):
        h = word(h, w)
    tail = data[n8:]
    if tail:
        h = word(h, int.from_bytes(tail + bytes(8 - len(tail)), "little"))
    return word(h, 0x8000000000000000 | len(tail))


def resealed(body: bytes) -> bytes:
    return body + struct.pack("<Q", word_hash(body))


def graph_files(out_dir):
    import numpy as np
    import fastlbp_amd as fl

    gen = fl.generate(400, 3)
    sources = {
        "gen": gen.factors,
        "clamped": fl.clamp(gen.factors, gen.alarms[:10], np.arange(10) % 2),
        "edge": fl.read_causal_fg(os.path.join(GOLDEN, "edge_cases.causal_fg")),
    }
    paths = []
    for name, flist in sources.items():
        p = os.path.join(out_dir, f"{name}.flbpcsr")
        fl.Graph(flist).write(p)
        paths.append(p)
    return paths


def damage_graph(data: bytes, rng: random.Random):
    """(tag, bytes) variants of one valid graph file."""
    out = []
    body = data[:-8]
    assert word_hash(body) == struct.unpack("<Q", data[-8:])[0], "integrity check restated wrongly"
    n = len(data)
    for k in range(24):   # truncations
        out.append((f"trunc{k}", data[:rng.randrange(0, n)]))
    for k in range(24):   # bit flips (caught by the check)
        i = rng.randrange(0, n)
        out.append((f"flip{k}", data[:i] + bytes([data[i] ^ (1 << rng.randrange(8))]) + data[i + 1:]))
    V, F, E = struct.unpack_from("<QQQ", body, 8)
    # header sizes edited, check recomputed
    for k, (v, f, e) in enumerate([(V + 1, F, E), (V, F - 1, E), (V, F, E + 7), (0, F, E), (V, 0, 0), (M64, F, E),
                                   (V, 1 << 40, E), (V, F, 1 << 33), (1 << 31, 1 << 31, 1 << 31)]):
        out.append((f"hdr{k}", resealed(body[:8] + struct.pack("<QQQ", v, f, e) + body[32:])))
    # array words edited at random positions past the header, check recomputed
    for k in range(48):
        i = 32 + rng.randrange(0, (len(body) - 32) // 4) * 4
        val = rng.choice([0, 1, 0xFFFFFFFF, 0x7FFFFFFF, V, F, E, E + 1, rng.randrange(1 << 32)])
        out.append((f"word{k}", resealed(body[:i] + struct.pack("<I", val) + body[i + 4:])))
    # trailing bytes and a resealed truncation
    out.append(("extra", data + b"\0" * 8))
    out.append(("short", resealed(body[: len(body) // 2])))
    return out


def damage_text(text: str, rng: random.Random, n_trunc=16, n_edit=40):
    out = []
    for k in range(n_trunc):
        out.append((f"trunc{k}", text[:rng.randrange(0, len(text))]))
    tokens = ["-1", "0", "1e308", "nan", "inf", "x", "", "4294967296", "18446744073709551616", "-0", "99999999",
              "0.5", "2", "\n", " ", "#"]
    for k in range(n_edit):
        i = rng.randrange(0, len(text))
        j = min(len(text), i + rng.randrange(1, 6))
        out.append((f"edit{k}", text[:i] + rng.choice(tokens) + text[j:]))
    lines = text.split("\n")
    for k in range(8):   # a line dropped or repeated
        i = rng.randrange(0, len(lines))
        out.append((f"drop{k}", "\n".join(lines[:i] + lines[i + 1:])))
        out.append((f"dup{k}", "\n".join(lines[:i + 1] + lines[i:])))
    return out


def fg_text(rng: random.Random) -> str:
    """A small generic factor graph: 6 factors over 8 variables."""
    lines = ["# replay source", "6", ""]
    for _ in range(6):
        members = rng.sample(range(8), rng.randrange(1, 4))
        dims = [rng.choice([2, 3]) for _ in members]
        size = 1
        for d in dims:
            size *= d
        entries = sorted(rng.sample(range(size), rng.randrange(1, size + 1)))
        lines += [str(len(members)), " ".join(map(str, members)), " ".join(map(str, dims)), str(len(entries))]
        lines += [f"{li} {rng.random()!r}" for li in entries]
        lines.append("")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("binary")
    ap.add_argument("--seed", type=int, default=20240611)
    ap.add_argument("--keep", default=None, help="write the corpus here and keep it")
    args = ap.parse_args()
    rng = random.Random(args.seed)
    work = args.keep or tempfile.mkdtemp(prefix="flbp_replay_")
    os.makedirs(work, exist_ok=True)
    runs = []   # (argv, file)
    for src in graph_files(work):
        data = open(src, "rb").read()
        base = os.path.splitext(os.path.basename(src))[0]
        runs.append((["--graph", src], src))
        for tag, blob in damage_graph(data, rng):
            p = os.path.join(work, f"{base}.{tag}.flbpcsr")
            with open(p, "wb") as f:
                f.write(blob)
            runs.append((["--graph", p], p))
    for name in ("gen120.causal_fg", "edge_cases.causal_fg"):
        text = open(os.path.join(GOLDEN, name)).read()
        base = name.split(".")[0]
        for tag, t in [("orig", text)] + damage_text(text, rng):
            p = os.path.join(work, f"{base}.{tag}.causal_fg")
            with open(p, "w") as f:
                f.write(t)
            runs.append((["--schedules", p], p))
            runs.append(([p], p))
    text = fg_text(rng)
    crafted = [("dup_member", "1\n\n2\n3 3\n2 2\n1\n3 0.5\n"),
               ("zero_states", "1\n\n1\n3\n0\n1\n0 0.5\n"),
               ("huge_table", "1\n\n3\n1 2 3\n100000 100000 100000\n1\n0 0.5\n"),
               ("huge_counts", "18446744073709551615\n\n99999999999\n1\n"),
               ("negative_index", "1\n\n1\n3\n2\n1\n-1 0.5\n")]
    for tag, t in [("orig", text)] + damage_text(text, rng) + crafted:
        p = os.path.join(work, f"generic.{tag}.fg")
        with open(p, "w") as f:
            f.write(t)
        runs.append(([p], p))

    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:exitcode=86"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=86"

    def one(run):
        argv, path = run
        r = subprocess.run([args.binary] + argv, capture_output=True, text=True, env=env, timeout=120)
        bad = r.returncode not in (0, 1) or "Sanitizer" in r.stderr or "runtime error" in r.stderr
        return argv, r.returncode, bad, r.stderr

    failed, codes = [], {}
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for argv, rc, bad, err in ex.map(one, runs):
            codes[rc] = codes.get(rc, 0) + 1
            if bad:
                failed.append((argv, rc, err[-2000:]))
    print(f"replayed {len(runs)} runs over {len(set(p for _, p in runs))} files; exit codes {dict(sorted(codes.items()))}")
    for argv, rc, err in failed[:10]:
        print(f"FAILED {' '.join(argv)} (exit {rc})\n{err}")
    if not args.keep:
        shutil.rmtree(work, ignore_errors=True)
    if failed:
        print(f"{len(failed)} runs failed")
        return 1
    print("no sanitizer reports")
    return 0


if __name__ == "__main__":
    sys.exit(main())
