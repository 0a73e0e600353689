"""Pins the CPU oracle against golden vectors written by the reference's own
C++ IPR path (tests/golden/make_golden.py) - stage by stage, no GPU."""
import os

import numpy as np
import pytest

import oracle as O
from common import (GOLDEN, GOLDEN_CASES, Case, oracle_ipr, oracle_peaks, oracle_shake,
                    oracle_stereo)


@pytest.fixture(scope="module", params=GOLDEN_CASES)
def case(request):
    return Case(request.param)


def test_peaks_bit_exact(case):
    for c in range(case.n_cam):
        got = oracle_peaks(case.images[c].astype(np.float64), case.params)
        want = case.z[f"peaks_cam{c}"]
        assert got.shape == want.shape
        assert np.array_equal(got, want)


def test_stereo_bit_exact(case):
    rc, r = oracle_stereo(case.cams, case.peaks(), case.params)
    assert rc == 0
    assert np.array_equal(r["cand_ids"], case.z["sm_cand_ids"])
    assert np.array_equal(r["cand_e"], case.z["sm_cand_e"])
    assert np.array_equal(r["sel"], case.z["sm_sel"])
    assert np.array_equal(r["obj_xyz"], case.z["sm_obj_xyz"])
    assert np.array_equal(r["obj_ids"], case.z["sm_obj_ids"])


def test_shake_bit_exact(case):
    xo, fl, sc, res = oracle_shake(case.cams, case.images, case.params, case.otf(),
                                   case.z["sm_obj_xyz"])
    assert np.array_equal(xo, case.z["sh_xyz"])
    assert np.array_equal(fl, case.z["sh_flags"])
    assert np.array_equal(sc, case.z["sh_score"])
    flat = res.reshape(case.n_cam, -1)
    assert np.array_equal(flat[:, case.z["sh_res_idx"]], case.z["sh_res_sample"])
    assert np.allclose(flat.sum(1), case.z["sh_res_sum"], rtol=1e-12)


def test_shake_planted_flags(case):
    if "shf_flags" not in case.z:
        pytest.skip("no planted-flag case")
    xo, fl, sc, _ = oracle_shake(case.cams, case.images, case.params, case.otf(),
                                 case.z["shf_xyz_in"], residual=False)
    assert np.array_equal(fl, case.z["shf_flags"])
    assert (fl & 1).sum() > 0 and (fl & 2).sum() > 0   # both Ghost and Repeated occur
    assert np.array_equal(xo, case.z["shf_xyz"])
    assert np.array_equal(sc, case.z["shf_score"])


def test_run_ipr_bit_exact(case):
    xyz, uv, loops = oracle_ipr(case.cams, case.images, case.params, case.otf())
    assert np.array_equal(loops, case.z["ipr_loops"])
    assert np.array_equal(xyz, case.z["ipr_xyz"])
    assert np.array_equal(uv, case.z["ipr_uv"])


@pytest.mark.parametrize("name", ["random", "dense"])
def test_oracle_prune_matches_reference(name):
    """pruneMatch alone on hand-built candidate sets (exact and one-ulp e_check
    ties, shared points): the oracle's selection equals the compiled
    reference's (tests/golden/make_golden_prune.py)."""
    z = np.load(os.path.join(GOLDEN, "prune_cases.npz"))
    ids = np.ascontiguousarray(z[f"{name}_ids"])
    e = np.ascontiguousarray(z[f"{name}_e"])
    npts = np.ascontiguousarray(z[f"{name}_npts"])
    sel = np.zeros(len(e), np.uint8)
    rc = O.load_port().orc_prune(ids.shape[1], None, O.iptr(npts), O.iptr(ids), O.dptr(e),
                                 len(e), O.u8ptr(sel))
    assert rc == 0
    assert np.array_equal(sel, z[f"{name}_sel"])


def test_oracle_limit2d_matches_reference():
    """limit2DObjectNumber's std::shuffle restated in C, on its one-position-
    per-draw branch (~49,000 peaks per camera cut to 3,000): runIPR equals the
    compiled reference's (tests/golden/make_golden_full.py)."""
    import hashlib
    from pyopenlpt_amd import synth as S
    z = np.load(os.path.join(GOLDEN, "ref_noise_limit.npz"))
    w = S.workload_noise_limit(seed=int(z["seed"]))
    assert hashlib.sha256(np.ascontiguousarray(w.images).tobytes()).hexdigest() == str(z["image_sha256"])
    xyz, _uv, loops = oracle_ipr(w.cams, w.images, w.params, w.otf_abca)
    assert np.array_equal(loops, z["loops"])
    assert np.array_equal(xyz, z["xyz"])
