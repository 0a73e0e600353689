"""tri_decide, the build's shortcut for the quick triangulation test
(StereoMatch.cpp:397-407: triangulate, then err <= tolerance), compiled for
the host from the library's own header and run against the exact
triangulate_n (myMATH::triangulation's Gauss inverse) on line sets whose
tolerance lies within 1e-5 relative of the error: every decision the
shortcut takes is the exact one; the rest it leaves to triangulate_n."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_tri_decide_matches_exact_triangulation(tmp_path):
    exe = str(tmp_path / "tri_decide_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off",
                    f"-I{os.path.join(ROOT, 'include')}",
                    f"-I{os.path.join(ROOT, 'pyopenlpt_amd', 'csrc')}",
                    os.path.join(HERE, "tri_decide_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe, "300000"], check=True, capture_output=True, text=True).stdout
    rows = re.findall(r"N=(\d) trials (\d+) undecided (\d+) wrong (\d+)", out)
    assert len(rows) == 3, out
    for n, trials, undec, wrong in rows:
        assert int(wrong) == 0, out
        # most near-threshold cases are still decided by the shortcut
        assert int(undec) < 0.6 * int(trials), out
