"""On-disk formats (no GPU): run / tracer configuration parsing, OTF files,
TIFF image lists, object and track files, and the tracking entry points'
argument checks."""
import ctypes
import os

import numpy as np

from pyopenlpt_amd import config as C
from pyopenlpt_amd import synth as S
from pyopenlpt_amd.camfile import save_camera

TRACER = dict(radius_search_obj_vox=10, n_initial_frames=4, radius_search_track_vox=30,
              shake_width_vox=0.25, pf_grid=(51, 41, 31), pf_r_vox=25, n_loop_ipr=2,
              n_shake_loop=4, thred_ghost=0.1, tol_2d_px=0.8, tol_3d_vox=1.0, n_cam_reduced=1,
              n_loop_ipr_reduced=2, min_obj_int=30, radius_obj=2)


def test_stoi_stod_take_the_leading_number():
    assert C.stoi("12 # threads") == 12 and C.stoi("  -3,") == -3 and C.stoi("4.9") == 4
    assert C.stod(".25 # shake") == 0.25 and C.stod("1e-2x") == 0.01 and C.stod("-20") == -20.0


def test_config_lines_rules(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text("# comment\n\n  12 # trailing\n\t0.5\t\n#x\n   # indented comment\n7\n")
    # '#' at the start of the trimmed line skips it (Config.cpp:24-28)
    assert C.config_lines(str(p)) == ["12", "0.5", "7"]


def _write_cams(tmp_path, n=4):
    cams = S.make_rig(64, 80, max_intensity=255)[:n]
    files = []
    for k, c in enumerate(cams):
        f = str(tmp_path / f"cam{k}.txt")
        save_camera(c, f)
        files.append(f)
    return cams, files


def test_basic_and_tracer_config_round_trip(tmp_path):
    cams, files = _write_cams(tmp_path)
    tc = str(tmp_path / "tracer.txt")
    C.write_tracer_config(tc, **TRACER)
    cfg = str(tmp_path / "config.txt")
    C.write_basic_config(cfg, frame_range=(3, 49), fps=2, n_thread=-1, cam_files=files,
                         max_intensity=[255, 255, 4095, 65535], image_lists=["a", "b", "c", "d"],
                         axis_limit=(-20, 20, -15, 15, -10, 10), voxel_to_mm=0.04,
                         output_path="/tmp/out", object_types=["Tracer"], object_configs=[tc])
    s = C.BasicSetting()
    assert s.readConfig(cfg)
    assert (s.frame_start, s.frame_end, s.fps, s.n_thread, s.n_cam) == (3, 49, 2, 0, 4)
    assert [c.max_intensity for c in s.cam_list] == [255, 255, 4095, 65535]
    assert s.image_file_paths == ["a", "b", "c", "d"]
    assert s.axis_limit == (-20, 20, -15, 15, -10, 10)
    assert s.output_path == "/tmp/out/" and s.object_types == ["Tracer"]
    assert not s.load_track and s.load_track_frame == 3
    assert np.allclose(s.cam_list[2].R, cams[2].R) and s.cam_list[1].n_col == 80
    t = C.TracerConfig()
    assert t.readConfig(tc, s)
    v = 0.04
    assert t.radius_search_obj == 10 * v and t.radius_search_track == 30 * v
    assert t.shake_width == 0.25 * v and t.shakewidth_min == 0.25 * v / 20
    assert (t.pf_nx, t.pf_ny, t.pf_nz) == (51, 41, 31) and t.pf_r == 25 * v
    assert (t.n_loop_ipr, t.n_shake_loop, t.thred_ghost) == (2, 4, 0.1)
    assert t.tol_2d_px == 0.8 and t.tol_3d_mm == 1.0 * v
    assert (t.n_cam_reduced, t.n_loop_ipr_reduced, t.min_obj_int, t.radius_obj) == (1, 2, 30, 2.0)
    assert t.match_cam_count == 4
    # the default OTF before any estimate: 5x5x5, a = max_intensity / 2 (OTF.cpp:159-193)
    assert t.otf.shape == (5, 5, 5)
    assert np.allclose(t.otf.a[:, 0], [127.5, 127.5, 2047.5, 32767.5]) and np.all(t.otf.b == 1.5)
    p = t.ipr_params()
    assert p.n_shake_loop == 4 and p.limit[3] == 15
    q = t.stb_params(s)
    assert q.frame_end == 49 and q.pf_ny == 41 and q.pf_nbin[0] == 11 and q.pf_smooth == 1


def test_initial_frames_at_least_four(tmp_path):
    _, files = _write_cams(tmp_path)
    tc = str(tmp_path / "tracer.txt")
    C.write_tracer_config(tc, **{**TRACER, "n_initial_frames": 2})
    s = C.BasicSetting(n_cam=4, voxel_to_mm=1.0, cam_list=S.make_rig(8, 8))
    t = C.TracerConfig()
    t.readCommonConfig(tc, s)
    assert t.n_initial_frames == 4   # Config.cpp:197-199


def test_load_track_flag(tmp_path):
    _, files = _write_cams(tmp_path)
    cfg = str(tmp_path / "config.txt")
    C.write_basic_config(cfg, frame_range=(0, 9), fps=1, n_thread=0, cam_files=files,
                         max_intensity=[255] * 4, image_lists=list("abcd"),
                         axis_limit=(-1, 1, -1, 1, -1, 1), voxel_to_mm=0.1, output_path="o/",
                         object_types=["Tracer"], object_configs=["t.txt"], load_track=(1, 5),
                         load_track_path="tracks/")
    s = C.BasicSetting()
    assert s.readConfig(cfg)
    assert s.load_track and s.load_track_frame == 5 and s.load_track_path == "tracks/"


def test_otf_file_round_trip_and_reference_layout(tmp_path):
    g = np.random.default_rng(0)
    o = C.Otf(2, (3, 2, 2), (-1, 1, -2, 2, -3, 3), g.uniform(100, 200, (2, 12)),
              g.uniform(1, 2, (2, 12)), g.uniform(1, 2, (2, 12)), g.uniform(-0.3, 0.3, (2, 12)))
    p = str(tmp_path / "otf.txt")
    C.save_otf(o, p)
    r = C.load_otf(p)
    assert r.shape == (3, 2, 2) and r.boundary == (-1, 1, -2, 2, -3, 3)
    for k in ("a", "b", "c", "alpha"):
        assert np.allclose(getattr(r, k), getattr(o, k), rtol=1e-8)
    # the layout of the reference's own sample (header comments, one row per
    # camera, trailing comments)
    q = tmp_path / "otf2.txt"
    q.write_text("# Size: (n_cam,nx,ny,nz,n_grid)\n2,2,2,2,8\n# Boundary\n0,1,0,1,0,1 # box\n"
                 "# a: (n_cam*n_grid)\n" + "\n".join([",".join(["125.0"] * 8)] * 2) + "\n"
                 "# b\n" + "\n".join([",".join(["1.5"] * 8)] * 2) + "\n"
                 "# c\n" + "\n".join([",".join(["1.25"] * 8)] * 2) + "\n"
                 "# alpha\n" + "\n".join([",".join(["0"] * 8)] * 2) + "\n")
    r = C.load_otf(str(q))
    assert r.n_cam == 2 and np.all(r.a == 125.0) and np.all(r.c == 1.25) and np.all(r.alpha == 0)


def test_tiff_round_trip_8_and_16_bit(tmp_path):
    g = np.random.default_rng(1)
    for dt in (np.uint8, np.uint16):
        img = g.integers(0, np.iinfo(dt).max, (37, 53)).astype(dt)
        p = str(tmp_path / f"i{dt.__name__}.tif")
        C.write_tiff(p, img)
        got = C.read_tiff(p)
        assert got.dtype == dt and np.array_equal(got, img)
    lst = tmp_path / "names.txt"
    lst.write_text(f"{tmp_path}/iuint8.tif\n{tmp_path}/iuint16.tif\n")
    io = C.ImageIO("", str(lst))
    assert io.loadImg(1).dtype == np.uint16
    try:
        io.loadImg(2)
        raise AssertionError("out-of-range frame accepted")
    except IndexError:
        pass


def test_save_obj_info_layout(tmp_path):
    p = str(tmp_path / "o.csv")
    xyz = np.array([[1.0, -2.5, 3.25], [0.125, 0.0, 1e-7]])
    uv = np.arange(2 * 2 * 2, dtype=float).reshape(2, 2, 2) + 0.5
    C.save_obj_info(p, xyz, uv)
    lines = open(p).read().splitlines()
    assert lines[0] == "WorldX,WorldY,WorldZ,Error,Ncam,cam0_x(col),cam0_y(row),cam1_x(col),cam1_y(row)"
    assert lines[1] == "1,-2.5,3.25,0.5,1.5,2.5,3.5"
    assert lines[2] == "0.125,0,1e-07,4.5,5.5,6.5,7.5"


def test_read_tracks(tmp_path):
    p = tmp_path / "t.csv"
    p.write_text("TrackID,FrameID,WorldX,WorldY,WorldZ,cam0_x(col),cam0_y(row)\n"
                 "0,3,1.000000,2.000000,3.000000,10.500000,11.500000\n"
                 "0,4,1.100000,2.000000,3.000000,10.600000,11.500000\n"
                 "1,4,-1.000000,0.000000,0.500000,1.000000,2.000000\n")
    t = C.read_tracks(str(p))
    assert t["track"].tolist() == [0, 0, 1] and t["frame"].tolist() == [3, 4, 4]
    assert t["uv"].shape == (3, 1, 2) and t["xyz"][1, 0] == 1.1


def test_stb_entry_points_check_arguments():
    from pyopenlpt_amd import _lib
    L = _lib.lib()
    out = ctypes.c_void_p()
    assert L.olpt_stb_create(None, None, None, None, ctypes.byref(out)) == 1
    assert L.olpt_stb_process_frame(None, 0, None, 0) == 1
    assert L.olpt_pred_field(None, None, None, 0, None, 0, None, None, 0, None) == 1
    assert L.olpt_estimate_otf(None, None, 0, 1, None, None, None) == 1


def test_abc_sequence_bounds_and_steps():
    seq = S.abc_sequence(300, 6, 3, max_disp_mm=0.3)
    for (p0, i0), (p1, i1) in zip(seq[:-1], seq[1:]):
        assert np.all(np.abs(p1) <= 20.0)
        same = i0 == i1
        assert np.linalg.norm(p1[same] - p0[same], axis=1).max() <= 0.3 + 1e-12


def test_save_obj_info_matches_reference_bytes(tmp_path):
    """save_obj_info writes the file the reference's IPR::saveObjInfo wrote
    for the same objects (tests/golden/make_golden_objinfo.py): header,
    default-format doubles, empty columns for a camera without an
    observation - byte for byte."""
    import os
    import numpy as np
    from pyopenlpt_amd import config as C
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    z = np.load(os.path.join(here, "objinfo_u8_256.npz"))
    out = tmp_path / "objs.csv"
    C.save_obj_info(str(out), z["xyz"], z["uv"])
    want = open(os.path.join(here, "objinfo_u8_256.csv"), "rb").read()
    assert out.read_bytes() == want
