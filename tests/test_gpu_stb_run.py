"""run(config) end to end against the compiled reference on the same run
folder: 2,000 particles advected by an ABC flow, 512x512 uint8, 12 frames,
reduced-camera IPR passes on.  The reference (oracle/_ref, built from the
reference's sources) and this library each read their own copy of the
folder (cameras, TIFF lists, config.txt, tracerConfig.txt), estimate the
OTF from the first five frames and track all frames; every file both write
- PredField_<f>_<f+1>.csv, InitialTrack/*.csv, ConvergeTrack/*.csv - must
agree: same names, same rows, same integer columns, and floats within
1e-6 (the files print 6 decimals for tracks, 8 significant digits for the
field; the FP64 runs agree to ~1e-12, so only a rounding tie of the text
could differ).  Skipped where the reference library is not built.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TRACER = dict(radius_search_obj_vox=5.0, n_initial_frames=4, radius_search_track_vox=25.0,
              shake_width_vox=0.25, pf_grid=(21, 21, 21), pf_r_vox=25.0, n_loop_ipr=3,
              n_shake_loop=4, thred_ghost=0.1, tol_2d_px=1.0, tol_3d_vox=1.0, n_cam_reduced=1,
              n_loop_ipr_reduced=1, min_obj_int=30, radius_obj=2.0)


def _case(root):
    from pyopenlpt_amd import config as C
    from pyopenlpt_amd import synth as S
    size = 512
    vox = 40.0 / (0.95 * size)
    cams = S.make_rig(size, size, max_intensity=255)
    seq = S.abc_sequence(2000, 12, 17, max_disp_mm=2.5 * vox)
    frames = S.render_sequence(cams, seq, seed=18)
    cfgs = []
    for name in ("ref", "gpu"):
        cfgs.append(C.write_run_folder(os.path.join(root, name), cams, frames,
                                       axis_limit=(-20, 20, -20, 20, -20, 20), voxel_to_mm=vox,
                                       tracer=TRACER))
    return cfgs, frames.shape[0]


def _files(out):
    got = {}
    for dp, _, fs in os.walk(out):
        for f in fs:
            if f.endswith(".csv"):
                got[os.path.relpath(os.path.join(dp, f), out)] = os.path.join(dp, f)
    return got


def _rows(path):
    with open(path) as f:
        return [line.rstrip("\n").split(",") for line in f]


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built")
def test_run_config_matches_reference(tmp_path, gpu_available):
    from pyopenlpt_amd.stb import run
    (cfg_ref, cfg_gpu), n_frames = _case(str(tmp_path))
    L = O.load_ref()
    L.ref_stb_open.restype = ctypes.c_void_p
    L.ref_stb_open.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]
    L.ref_stb_process.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.ref_stb_close.argtypes = [ctypes.c_void_p]
    h = L.ref_stb_open(cfg_ref.encode(), None, 1)
    assert h
    for f in range(n_frames):
        assert L.ref_stb_process(ctypes.c_void_p(h), f, 1) == 0
    L.ref_stb_close(ctypes.c_void_p(h))
    assert run(cfg_gpu, precision="f64", verbose=False) == 0
    ref = _files(os.path.join(str(tmp_path), "ref", "out"))
    gpu = _files(os.path.join(str(tmp_path), "gpu", "out"))
    assert sorted(ref) == sorted(gpu) and len(ref) >= 8, (sorted(ref), sorted(gpu))
    n_long = 0
    for name in sorted(ref):
        a, b = _rows(ref[name]), _rows(gpu[name])
        assert len(a) == len(b), name
        if name.startswith("PredField"):
            fa = np.array(a, float)
            fb = np.array(b, float)
            assert np.abs(fa - fb).max() <= 1e-6 * max(1.0, np.abs(fa).max()), name
            continue
        assert a[0] == b[0], name      # header
        ia = np.array([r[:2] for r in a[1:]], int).reshape(-1, 2)
        ib = np.array([r[:2] for r in b[1:]], int).reshape(-1, 2)
        assert np.array_equal(ia, ib), name
        fa = np.array([r[2:] for r in a[1:]], float)
        fb = np.array([r[2:] for r in b[1:]], float)
        if fa.size:
            assert np.abs(fa - fb).max() <= 1.01e-6, name
        if name.startswith("ConvergeTrack/LongTrackActive"):
            n_long = len(set(ia[:, 0].tolist()))
    assert n_long > 1000
