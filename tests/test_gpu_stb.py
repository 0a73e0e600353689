"""Shake-the-box tracking on the GPU against the reference's own C++
(golden vectors written by tests/golden/make_golden_stb.py from the compiled
reference: STB.cpp, PredField.cpp, Track.cpp driven by its config reader).

Bars:
* predictNext: bit-exact positions (same FP64 operations in the same order).
* checkLinearFit: identical decisions.
* PredField: field and getDisp within 1e-9 mm; the only differences are
  the last bits of the device exp/log in the displacement histogram.
* The whole run (config files -> OTF estimate -> 10 frames of STB): the OTF
  estimate within 1e-12 relative; in FP64 mode the four track lists with the
  same tracks, frames and positions (<= 1e-9 mm); in FP32 mode (shake and
  residual in FP32) long-track counts within 2% and positions of matching
  points within one shake step.
"""
import ast
import os

import numpy as np
import pytest

from common import Case

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "stb_abc_256.npz")
LISTS = ("short_active", "long_active", "long_inactive", "exit")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLDEN))


@pytest.fixture(scope="module")
def ctx(gpu_available, gold):
    from pyopenlpt_amd.context import Context
    cams = [Case._cam(r) for r in gold["cameras"]]
    c = Context(cams, cams[0].n_row, cams[0].n_col, precision="f64")
    yield c
    c.close()


def test_predict_next_bit_exact(ctx, gold):
    from pyopenlpt_amd.stb import predict_next
    pos, lens = gold["pn_pos"], gold["pn_len"]
    win = np.zeros((len(lens), 6, 3))
    m = np.minimum(lens, 6)
    for i in range(len(lens)):
        win[i, : m[i]] = pos[i, lens[i] - m[i]: lens[i]]
    pred, ok = predict_next(ctx, win, m)
    assert np.array_equal(ok, gold["pn_ok"])
    sel = ok == 1
    assert np.array_equal(pred[sel], gold["pn_pred"][sel]), np.abs(pred[sel] - gold["pn_pred"][sel]).max()


def test_linear_fit_decisions(ctx, gold):
    from pyopenlpt_amd.stb import check_linear_fit
    pos, lens = gold["lf_pos"], gold["lf_len"]
    win = np.zeros((len(lens), 4, 3))
    for i in range(len(lens)):
        m = min(4, lens[i])
        win[i, :m] = pos[i, lens[i] - m: lens[i]]
    ok = check_linear_fit(ctx, 4, win, lens)
    assert np.array_equal(ok, gold["lf_ok"])
    assert 0 < ok.sum() < len(ok)


def test_pred_field(ctx, gold):
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.stb import PredField
    tr = ast.literal_eval(str(gold["tracer"]))
    vox = float(gold["voxel_mm"])
    sp = S.OlptStbParams()
    for i, v in enumerate((-20, 20, -20, 20, -20, 20)):
        sp.limit[i] = v
    sp.pf_nx, sp.pf_ny, sp.pf_nz = tr["pf_grid"]
    sp.pf_r_mm = tr["pf_r_vox"] * vox
    for i in range(3):
        sp.pf_nbin[i] = 11
        sp.pf_sigma[i] = 1.0
    sp.pf_smooth = 1
    field, disp = PredField(ctx, sp).calPredField(gold["pf_prev"], gold["pf_curr"], gold["pf_q"])
    assert np.abs(field - gold["pf_field"]).max() <= 1e-9
    assert np.abs(disp - gold["pf_disp"]).max() <= 1e-9
    assert np.abs(gold["pf_field"]).max() > 0.1   # a non-trivial field


def _run(tmp_path, gold, precision):
    from pyopenlpt_amd import config as C
    from pyopenlpt_amd.stb import STB
    cams = [Case._cam(r) for r in gold["cameras"]]
    tr = ast.literal_eval(str(gold["tracer"]))
    cfg = C.write_run_folder(str(tmp_path), cams, gold["frames"],
                             axis_limit=(-20, 20, -20, 20, -20, 20),
                             voxel_to_mm=float(gold["voxel_mm"]), tracer=tr)
    setting = C.BasicSetting()
    assert setting.readConfig(cfg)
    stb = STB(setting, "Tracer", setting.object_config_paths[0], precision=precision)
    ios = [C.ImageIO("", p) for p in setting.image_file_paths]
    for f in range(setting.frame_start, setting.frame_end + 1):
        stb.process_frame(f, np.stack([io.loadImg(f) for io in ios]))
    return stb


def test_stb_run_f64_matches_reference(tmp_path, gpu_available, gold):
    stb = _run(tmp_path, gold, "f64")
    est = np.stack([stb.cfg.otf.a[:, 0], stb.cfg.otf.b[:, 0], stb.cfg.otf.c[:, 0],
                    stb.cfg.otf.alpha[:, 0]], 1)
    assert np.allclose(est, gold["otf_abca_est"], rtol=1e-12, atol=0)
    for which, name in enumerate(LISTS):
        r = stb.rows(which)
        assert r["n_tracks"] == int(gold[f"tr_{name}_n"]), name
        assert np.array_equal(r["track"], gold[f"tr_{name}_track"]), name
        assert np.array_equal(r["frame"], gold[f"tr_{name}_frame"]), name
        if len(r["xyz"]):
            assert np.abs(r["xyz"] - gold[f"tr_{name}_xyz"]).max() <= 1e-9, name
    # files the reference writes at the last frame
    out = os.path.join(str(tmp_path), "out", "ConvergeTrack")
    assert os.path.exists(os.path.join(out, "LongTrackActive_9.csv"))
    from pyopenlpt_amd import config as C
    t = C.read_tracks(os.path.join(out, "LongTrackActive_9.csv"))
    assert np.array_equal(t["track"], gold["tr_long_active_track"])
    assert np.abs(t["xyz"] - gold["tr_long_active_xyz"]).max() <= 5e-7   # 6 decimals
    stb.close()


def test_stb_run_f32_close(tmp_path, gpu_available, gold):
    stb = _run(tmp_path, gold, "f32")
    r = stb.rows(1)
    want = int(gold["tr_long_active_n"])
    assert abs(r["n_tracks"] - want) <= max(2, int(0.02 * want))
    stb.close()


# ---------------------------------------------------------------------------
# larger inputs against the oracle's restatements (oracle/olpt_oracle.c,
# pinned to the same golden vectors by tests/test_oracle_stb.py)
# ---------------------------------------------------------------------------
def _tracks(n, max_len, seed):
    g = np.random.default_rng(seed)
    lens = g.integers(2, max_len + 1, n).astype(np.int32)
    t = np.arange(max_len)[None, :, None]
    pos = (g.uniform(-20, 20, (n, 1, 3)) + g.normal(0, 0.3, (n, 1, 3)) * t +
           g.normal(0, 0.02, (n, 1, 3)) * t * t + g.normal(0, 0.01, (n, max_len, 3)))
    pos[::13, :, 1] = g.uniform(-0.9, 0.9, (len(pos[::13]), 1))   # shifted branch
    return pos, lens


def test_predict_next_vs_oracle_100k(ctx):
    import oracle as O
    from pyopenlpt_amd.stb import predict_next
    pos, lens = _tracks(100_000, 12, 21)
    want = np.zeros((len(lens), 3))
    wok = np.zeros(len(lens), np.int32)
    O.load_port().orc_stb_predict(O.dptr(pos), O.iptr(lens), len(lens), pos.shape[1],
                                  O.dptr(want), O.iptr(wok))
    m = np.minimum(lens, 6)
    win = np.zeros((len(lens), 6, 3))
    for k in range(2, 13):
        sel = lens == k
        mm = min(k, 6)
        win[sel, :mm] = pos[sel, k - mm:k]
    got, ok = predict_next(ctx, win, m)
    assert np.array_equal(ok, wok)
    assert np.array_equal(got[ok == 1], want[wok == 1])


def test_linear_fit_vs_oracle_100k(ctx):
    import oracle as O
    from pyopenlpt_amd.stb import check_linear_fit
    pos, lens = _tracks(100_000, 10, 22)
    want = np.zeros(len(lens), np.int32)
    O.load_port().orc_stb_linear_fit(4, O.dptr(pos), O.iptr(lens), len(lens), pos.shape[1],
                                     O.iptr(want))
    win = np.zeros((len(lens), 4, 3))
    for k in range(2, 11):
        sel = lens == k
        mm = min(k, 4)
        win[sel, :mm] = pos[sel, k - mm:k]
    got = check_linear_fit(ctx, 4, win, lens)
    assert np.array_equal(got, want)
    assert 0.2 < want.mean() < 0.9


def test_pred_field_vs_oracle_31cube(ctx):
    import ctypes
    import oracle as O
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.stb import PredField
    g = np.random.default_rng(23)
    prev = g.uniform(-20, 20, (5000, 3))
    curr = prev + np.stack([0.3 * np.sin(prev[:, 2] / 6), 0.2 * np.cos(prev[:, 0] / 7),
                            0.1 * np.ones(len(prev))], 1) + g.normal(0, 0.01, prev.shape)
    curr[::10] = g.uniform(-20, 20, (len(curr[::10]), 3))
    q = g.uniform(-21, 21, (2000, 3))
    lim = np.array([-20.0, 20, -20, 20, -20, 20])
    r = 1.5
    sp = S.OlptStbParams()
    for i, v in enumerate(lim):
        sp.limit[i] = v
    sp.pf_nx = sp.pf_ny = sp.pf_nz = 31
    sp.pf_r_mm = r
    for i in range(3):
        sp.pf_nbin[i] = 11
        sp.pf_sigma[i] = 1.0
    sp.pf_smooth = 1
    field, disp = PredField(ctx, sp).calPredField(prev, curr, q)
    wf = np.zeros((31 ** 3, 3))
    wd = np.zeros((len(q), 3))
    O.load_port().orc_pred_field(O.dptr(lim), 31, 31, 31, ctypes.c_double(r),
                                 O.iptr(np.array([11, 11, 11], np.int32)), 1,
                                 O.dptr(np.ones(3)), O.dptr(prev), len(prev), O.dptr(curr),
                                 len(curr), O.dptr(wf), O.dptr(q), len(q), O.dptr(wd))
    assert np.abs(field - wf).max() <= 1e-9
    assert np.abs(disp - wd).max() <= 1e-9
    assert np.abs(wf).max() > 0.1


def test_tracks_save_load_resume(tmp_path, gpu_available, gold):
    """saveTracksAll after frame 6, loadTracksAll into a fresh STB (the
    reference's load_track resume; like STB::loadTracksAll, STB.cpp:1052-1076,
    it restores the two active lists): the loaded lists equal the saved ones
    to the files' 6 decimals, and frames 7-9 on both STBs end with the same
    active tracks and frames, and the same number of tracks leaving the
    active lists.  The loaded points carry the files' rounding (5e-7 mm) into
    later predictions, which can flip a branch of the shake search, so later
    positions agree within a tenth of the shake width."""
    from pyopenlpt_amd import config as C
    from pyopenlpt_amd.stb import STB
    cams = [Case._cam(r) for r in gold["cameras"]]
    tr = ast.literal_eval(str(gold["tracer"]))
    cfg = C.write_run_folder(str(tmp_path / "run"), cams, gold["frames"],
                             axis_limit=(-20, 20, -20, 20, -20, 20),
                             voxel_to_mm=float(gold["voxel_mm"]), tracer=tr)
    setting = C.BasicSetting()
    assert setting.readConfig(cfg)
    ios = [C.ImageIO("", p) for p in setting.image_file_paths]
    imgs = {f: np.stack([io.loadImg(f) for io in ios])
            for f in range(setting.frame_start, setting.frame_end + 1)}
    a = STB(setting, "Tracer", setting.object_config_paths[0], precision="f64",
            write_files=False)
    for f in range(0, 7):
        a.process_frame(f, imgs[f])
    save = str(tmp_path / "saved") + os.sep
    os.makedirs(save, exist_ok=True)
    a.saveTracksAll(save, 6)
    left_a = a.rows(2)["n_tracks"] + a.rows(3)["n_tracks"]
    b = STB(setting, "Tracer", setting.object_config_paths[0], precision="f64",
            write_files=False)
    b.loadTracksAll(save, 6)
    for which in (0, 1):
        ra, rb = a.rows(which), b.rows(which)
        assert ra["n_tracks"] == rb["n_tracks"], LISTS[which]
        assert np.array_equal(ra["track"], rb["track"]), LISTS[which]
        assert np.array_equal(ra["frame"], rb["frame"]), LISTS[which]
        if len(ra["xyz"]):
            assert np.abs(ra["xyz"] - rb["xyz"]).max() <= 5.1e-7, LISTS[which]
    assert a.rows(1)["n_tracks"] > 50
    for f in range(7, 10):
        a.process_frame(f, imgs[f])
        b.process_frame(f, imgs[f])
    tol = 0.1 * tr["shake_width_vox"] * float(gold["voxel_mm"])
    for which in (0, 1):
        ra, rb = a.rows(which), b.rows(which)
        assert ra["n_tracks"] == rb["n_tracks"], LISTS[which]
        assert np.array_equal(ra["frame"], rb["frame"]), LISTS[which]
        if len(ra["xyz"]):
            d = np.abs(ra["xyz"] - rb["xyz"]).max(1)
            assert d.max() <= tol, LISTS[which]
            assert np.median(d) <= 1e-6, LISTS[which]
    left_after = a.rows(2)["n_tracks"] + a.rows(3)["n_tracks"]
    assert left_after - left_a == b.rows(2)["n_tracks"] + b.rows(3)["n_tracks"]
    a.close()
    b.close()


def test_stb_learned_ipr_hook(tmp_path, gpu_available, gold):
    """olpt_stb_set_ipr_model: the tracker's IPR calls run with the learned
    proposal and the given IPR parameters.  Frames 0-3 (the initial phase:
    IPR per frame, then tracks from its objects through the predictive
    field) with an untrained network: every frame-0 point of the tracks is
    an object of olpt_ipr_run with the same network and parameters on the
    tracker's own context (same OTF estimate), and not every one is an
    object of the plain IPR; the model is not left on the context."""
    import torch
    from scipy.spatial import cKDTree
    from pyopenlpt_amd import config as C
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.context import to_device_image
    from pyopenlpt_amd.learned import ShakeNet
    from pyopenlpt_amd.stb import STB
    cams = [Case._cam(r) for r in gold["cameras"]]
    tr = ast.literal_eval(str(gold["tracer"]))
    vox = float(gold["voxel_mm"])
    cfg = C.write_run_folder(str(tmp_path), cams, gold["frames"],
                             axis_limit=(-20, 20, -20, 20, -20, 20), voxel_to_mm=vox, tracer=tr)
    setting = C.BasicSetting()
    assert setting.readConfig(cfg)
    ios = [C.ImageIO("", p) for p in setting.image_file_paths]
    imgs = [np.stack([io.loadImg(f) for io in ios]) for f in range(4)]
    net = ShakeNet(max_batch=4096, seed=3)
    stb = STB(setting, "Tracer", setting.object_config_paths[0], precision="f32",
              write_files=False)
    stb.set_ipr_model(net, vox, 0.6, n_shake_loop=1, shake_width_vox=0.03)
    for f in range(4):
        stb.process_frame(f, imgs[f])
    pts = []
    for which in range(4):
        r = stb.rows(which)
        pts.append(r["xyz"][r["frame"] == 0])
    pts = np.concatenate(pts)
    assert len(pts) > 100

    d, code = to_device_image(imgs[0])
    cap = 1 << 16

    def ipr(params):
        (ox, _u, n), _lt = stb.ctx.ipr_run(d, code, params, cap_obj=cap)
        torch.cuda.synchronize()
        return ox[0, : int(n[0].item())].cpu().numpy()

    plain = ipr(stb._ip)                    # no model left on the context
    lp = S.OlptIprParams.from_buffer_copy(stb._ipr_params)
    stb.ctx.set_shake_model(net, vox, 0.6)
    learned = ipr(lp)
    stb.ctx.set_shake_model(None)
    dl, _ = cKDTree(learned).query(pts)
    dp, _ = cKDTree(plain).query(pts)
    print(f"frame-0 track points {len(pts)}; learned IPR {len(learned)}, plain {len(plain)}; "
          f"max dist to learned {dl.max():.2e}, to plain {dp.max():.2e} mm")
    assert dl.max() <= 1e-9
    assert dp.max() > 1e-6
    stb.set_ipr_model(None)
    stb.close()
    net.close()
