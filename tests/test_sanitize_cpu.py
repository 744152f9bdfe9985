"""The readers of untrusted input under AddressSanitizer and
UndefinedBehaviorSanitizer: `make -C tests/native sanitize` builds the
library's host sources and flbp_graphdump with both sanitizers (g++, host
code only) and replays a seeded corpus of damaged binary graph files,
.causal_fg files and .fg files through them; every file must end in a clean
exit with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with libasan")
def test_damaged_inputs_replay_clean_under_sanitizers():
    if not os.path.isdir(os.path.join(ROOT, "_build", "fastlbp")):
        pytest.skip("the library's device objects are not built (make -C fastlbp_amd/csrc)")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "native"), "sanitize"], capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "no sanitizer reports" in r.stdout
