#!/usr/bin/env python3
"""Write the tracking (STB) golden vectors from the reference's own C++.

Runs oracle/_ref/libolpt_ref.so (ref_harness_stb.cpp drives the compiled
reference: BasicSetting::readConfig, STB construction with the OTF estimate
from the first five frames, STB::processFrame per frame) on a seeded
synthetic sequence - particles advected by an ABC flow, rendered into four
256x256 uint8 cameras - written to disk as the reference's run folder
(camera files, TIFF image lists, config.txt, tracerConfig.txt).  Stores the
inputs (frames, cameras, configuration values) and the outputs: the four
track lists after the last frame, the estimated OTF, and unit vectors of
STB::predictNext, STB::checkLinearFit and PredField::calPredField/getDisp.

Usage: python tests/golden/make_golden_stb.py  (needs the reference checkout)
"""
from __future__ import annotations

import ctypes
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from pyopenlpt_amd import config as C  # noqa: E402
from pyopenlpt_amd import synth as S  # noqa: E402

from make_golden import cams_to_array  # noqa: E402

TRACER = dict(radius_search_obj_vox=3.0, n_initial_frames=4, radius_search_track_vox=30.0,
              shake_width_vox=0.25, pf_grid=(11, 11, 11), pf_r_vox=25.0, n_loop_ipr=2,
              n_shake_loop=4, thred_ghost=0.1, tol_2d_px=1.0, tol_3d_vox=1.0, n_cam_reduced=0,
              n_loop_ipr_reduced=2, min_obj_int=30, radius_obj=2.0)
N_PART, N_FRAMES, SIZE, SEED = 400, 10, 256, 61


def ref_lib():
    L = O.load_ref()
    L.ref_stb_open.restype = ctypes.c_void_p
    L.ref_stb_open.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]
    L.ref_stb_close.argtypes = [ctypes.c_void_p]
    L.ref_stb_process.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    for f in ("ref_stb_counts", "ref_stb_tracks", "ref_stb_predict", "ref_stb_linfit",
              "ref_pred_field"):
        getattr(L, f).argtypes = None
    return L


def make_case():
    vox = 40.0 / (0.95 * SIZE)
    cams = S.make_rig(SIZE, SIZE, max_intensity=255)
    seq = S.abc_sequence(N_PART, N_FRAMES, SEED, max_disp_mm=2.0 * vox)
    frames = S.render_sequence(cams, seq, seed=SEED + 1)
    return vox, cams, seq, frames


def run_reference(cfg_path, n_frames, n_cam):
    L = ref_lib()
    otf = np.zeros((n_cam, 4))
    h = L.ref_stb_open(cfg_path.encode(), otf.ctypes.data, 1)
    assert h, "reference STB construction failed"
    for f in range(n_frames):
        assert L.ref_stb_process(ctypes.c_void_p(h), f, 1) == 0, f
    out = {"otf_abca_est": otf}
    for which, name in enumerate(("short_active", "long_active", "long_inactive", "exit")):
        nt, npnt = ctypes.c_int32(), ctypes.c_int32()
        L.ref_stb_counts(ctypes.c_void_p(h), which, ctypes.byref(nt), ctypes.byref(npnt))
        tid = np.zeros(max(1, npnt.value), np.int32)
        fr = np.zeros(max(1, npnt.value), np.int32)
        xyz = np.zeros((max(1, npnt.value), 3))
        assert L.ref_stb_tracks(ctypes.c_void_p(h), which, npnt.value, O.iptr(tid), O.iptr(fr),
                                O.dptr(xyz)) == 0
        out[f"tr_{name}_track"] = tid[: npnt.value]
        out[f"tr_{name}_frame"] = fr[: npnt.value]
        out[f"tr_{name}_xyz"] = xyz[: npnt.value]
        out[f"tr_{name}_n"] = np.int32(nt.value)
    return L, h, out


def unit_vectors(L, h, vox):
    g = np.random.default_rng(71)
    out = {}
    # predictNext: lengths 2..9 (windows of the last 6 points), smooth and
    # noisy tracks, coordinates near zero (shifted branch) and all-zero axes
    n = 600
    lens = g.integers(2, 10, n)
    pos = np.zeros((n, 9, 3))
    for i in range(n):
        p0 = g.uniform(-20, 20, 3)
        if i % 7 == 0:
            p0[g.integers(0, 3)] = g.uniform(-0.9, 0.9)
        v = g.normal(0, 0.2, 3)
        a = g.normal(0, 0.02, 3)
        t = np.arange(9)[:, None]
        pos[i] = p0 + v * t + 0.5 * a * t * t + g.normal(0, 0.005 if i % 3 else 0.0, (9, 3))
        if i % 11 == 0:
            pos[i, :, g.integers(0, 3)] = 0.0
    pred = np.zeros((n, 3))
    ok = np.zeros(n, np.int32)
    for L_ in range(2, 10):
        sel = np.where(lens == L_)[0]
        if not len(sel):
            continue
        pp = np.ascontiguousarray(pos[sel, :L_])
        o = np.zeros((len(sel), 3))
        k = np.zeros(len(sel), np.int32)
        L.ref_stb_predict(ctypes.c_void_p(h), O.dptr(pp), len(sel), L_, O.dptr(o), O.iptr(k))
        pred[sel], ok[sel] = o, k
    out.update(pn_pos=pos, pn_len=lens.astype(np.int32), pn_pred=pred, pn_ok=ok)
    # checkLinearFit: linear tracks plus noise of several sizes around the
    # 5e-2 mm threshold, lengths 2..8
    n = 800
    lens = g.integers(2, 9, n)
    pos = np.zeros((n, 8, 3))
    sig = g.choice([0.0, 0.005, 0.02, 0.04, 0.08], n)
    for i in range(n):
        t = np.arange(8)[:, None]
        pos[i] = g.uniform(-20, 20, 3) + g.normal(0, 0.3, 3) * t + g.normal(0, 1, (8, 3)) * sig[i]
    okf = np.zeros(n, np.int32)
    for L_ in range(2, 9):
        sel = np.where(lens == L_)[0]
        if not len(sel):
            continue
        k = np.zeros(len(sel), np.int32)
        L.ref_stb_linfit(ctypes.c_void_p(h), O.dptr(np.ascontiguousarray(pos[sel, :L_])),
                         len(sel), L_, O.iptr(k))
        okf[sel] = k
    out.update(lf_pos=pos, lf_len=lens.astype(np.int32), lf_ok=okf)
    # PredField: a uniformly displaced cloud (plus outliers) on the STB's
    # 11^3 grid with r = 25 vox, queried at random points
    prev = g.uniform(-20, 20, (1500, 3))
    disp = np.array([0.3, -0.2, 0.15])
    curr = prev + disp + g.normal(0, 0.02, prev.shape)
    curr[::9] = g.uniform(-20, 20, (len(curr[::9]), 3))
    curr = curr[g.permutation(len(curr))]
    q = g.uniform(-22, 22, (300, 3))
    ng = 11 ** 3
    field = np.zeros((ng, 3))
    dq = np.zeros((len(q), 3))
    L.ref_pred_field(ctypes.c_void_p(h), O.dptr(prev), len(prev), O.dptr(curr), len(curr),
                     O.dptr(field), O.dptr(q), len(q), O.dptr(dq))
    out.update(pf_prev=prev, pf_curr=curr, pf_field=field, pf_q=q, pf_disp=dq)
    return out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libolpt_ref.so is missing: run `make -C oracle ref`")
    vox, cams, seq, frames = make_case()
    with tempfile.TemporaryDirectory() as tmp:
        cfg = C.write_run_folder(tmp, cams, frames, axis_limit=(-20, 20, -20, 20, -20, 20),
                                 voxel_to_mm=vox, tracer=TRACER)
        L, h, out = run_reference(cfg, N_FRAMES, len(cams))
        out.update(unit_vectors(L, h, vox))
        L.ref_stb_close(ctypes.c_void_p(h))
    truth_pos = np.stack([p for p, _ in seq])
    truth_ids = np.stack([i for _, i in seq])
    path = os.path.join(HERE, "stb_abc_256.npz")
    np.savez_compressed(path, frames=frames, cameras=cams_to_array(cams), voxel_mm=vox,
                        tracer=np.array(repr(TRACER)), truth_pos=truth_pos, truth_ids=truth_ids,
                        **out)
    counts = {k[3:-2]: int(v) for k, v in out.items() if k.startswith("tr_") and k.endswith("_n")}
    print(f"stb_abc_256: tracks {counts}; otf {out['otf_abca_est'][0].round(3).tolist()}; "
          f"predict ok {int(out['pn_ok'].sum())}/{len(out['pn_ok'])}; linfit ok "
          f"{int(out['lf_ok'].sum())}/{len(out['lf_ok'])} -> {os.path.getsize(path) // 1024} KB")


if __name__ == "__main__":
    main()
