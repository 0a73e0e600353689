"""run(config) end to end against the compiled reference on the same run
folder: 2,000 particles advected by an ABC flow, 512x512 uint8, 12 frames,
reduced-camera IPR passes on.  tests/golden/make_golden_run.py ran the
reference (oracle/_ref, built from the reference's sources: readConfig, the
OTF estimate from the first five frames, processFrame per frame) on this
folder and stored every file it wrote in run_abc_512.npz; this test rebuilds
the folder from the same seeds, runs the library's run(config) in FP64 and
compares PredField_<f>_<f+1>.csv, InitialTrack/*.csv and ConvergeTrack/*.csv:
same names, same rows, same integer columns, and floats within 1e-6 (the
files print 6 decimals for tracks, 8 significant digits for the field; the
FP64 runs agree to ~1e-12, so only a rounding tie of the text could differ).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden_run as G  # noqa: E402

pytestmark = pytest.mark.gpu


def test_run_config_matches_reference(tmp_path, gpu_available):
    from pyopenlpt_amd.stb import run
    gold = np.load(G.OUT)
    names = [str(n) for n in gold["names"]]
    cfg, _n = G.write_case(os.path.join(str(tmp_path), "gpu"))
    assert run(cfg, precision="f64", verbose=False) == 0
    gpu = G.csv_files(os.path.join(str(tmp_path), "gpu", "out"))
    assert sorted(gpu) == names and len(names) >= 8, (sorted(gpu), names)
    n_long = 0
    for i, name in enumerate(names):
        d = G.read_output(gpu[name], name)
        if name.startswith("PredField"):
            fa = gold[f"val_{i}"]
            assert d["val"].shape == fa.shape, name
            assert np.abs(d["val"] - fa).max() <= 1e-6 * max(1.0, np.abs(fa).max()), name
            continue
        assert d["hdr"] == str(gold[f"hdr_{i}"]), name
        assert np.array_equal(d["int"], gold[f"int_{i}"]), name
        fa = gold[f"fix_{i}"]
        assert d["fix"].shape == fa.shape, name
        if fa.size:
            assert np.abs(d["fix"] - fa).max() <= 1, name   # one unit of the 6th decimal
        if name.startswith("ConvergeTrack/LongTrackActive"):
            n_long = len(set(d["int"][:, 0].tolist()))
    assert n_long > 1000
