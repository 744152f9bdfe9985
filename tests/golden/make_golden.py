#!/usr/bin/env python3
"""Writes the graph-structure fixtures under tests/golden/.

For every .causal_fg file here (plus a generated one) and its optional
.obs observation file, runs oracle/_ref/ref_dump -- the reference's own
reader, clamp and graph construction (csrc/ext/causal_fg.cpp,
csrc/ext/causal_factor.cpp, csrc/lbp/bipgraph.cpp) compiled from its
sources -- and stores its output as <name>.refdump.txt (and
<name>.obs.refdump.txt with the clamps applied and the graph rebuilt as
examples/fl_wrapper.cu does).  The tests compare the engine's builder with
these dumps on any machine.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import fastlbp_amd as fl  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_dump")


def dump(path, obs=None):
    args = [REF, path] + ([obs] if obs else [])
    return subprocess.run(args, check=True, capture_output=True, text=True).stdout


def main():
    if not os.path.exists(REF):
        sys.exit("build oracle/_ref first: make -C oracle ref")
    gen = fl.generate(120, 9, n_alarms=10)
    fl.write_causal_fg(gen.factors, os.path.join(HERE, "gen120.causal_fg"))
    with open(os.path.join(HERE, "gen120.obs"), "w") as f:
        for lab, val in zip(gen.alarms[:6], [1, 0, 1, 1, 0, 0]):
            f.write(f"{int(lab)} {'true' if val else 'false'}\n")
        f.write("0 false\n5 true\n")
    for path in sorted(glob.glob(os.path.join(HERE, "*.causal_fg"))):
        base = path[: -len(".causal_fg")]
        with open(base + ".refdump.txt", "w") as f:
            f.write(dump(path))
        if os.path.exists(base + ".obs"):
            with open(base + ".obs.refdump.txt", "w") as f:
                f.write(dump(path, base + ".obs"))
        print("wrote", os.path.basename(base))


if __name__ == "__main__":
    main()
