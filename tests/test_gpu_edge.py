"""GPU edge cases and error behaviour of the C-ABI."""
import numpy as np
import pytest
import torch

from common import GOLDEN_CASES, Case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gpu_available):
    return Case("golden_u8_256")


def make_ctx(case, precision="f32", stream=None, **kw):
    from pyopenlpt_amd.context import Context
    ctx = Context(case.cams, case.H, case.W, precision=precision, stream=stream,
                  max_2d=kw.get("max_2d", 1 << 15), max_cand=1 << 18, max_obj=1 << 15)
    otf, keep = case.otf()
    ctx.set_otf(otf)
    ctx._keep = keep
    return ctx


def test_empty_images(case):
    from pyopenlpt_amd.context import to_device_image
    ctx = make_ctx(case)
    d, code = to_device_image(np.zeros_like(case.images))
    assert len(ctx.find_peaks(d[0].contiguous(), code, case.params)) == 0
    (ox, _u, n_obj), lt = ctx.ipr_run(d, code, case.params, cap_obj=1 << 15)
    assert int(n_obj[0].item()) == 0
    assert (lt[0][:4] == 0).all()
    ctx.close()


def test_stereo_needs_two_active_cameras(case):
    from pyopenlpt_amd._lib import FatalError
    ctx = make_ctx(case)
    with pytest.raises(FatalError) as e:
        ctx.stereo_match(case.peaks(), case.params, active=[1, 0, 0, 0])
    assert e.value.status == 1   # ErrorCode::InvalidArgument (StereoMatch.cpp:39-40)
    ctx.close()


def test_check_cameras_configuration_is_rejected(case):
    """match_cam_count below the active count: the reference's checkMatch
    leaves e_check NaN and match() raises InvalidArgument (REQUIRE,
    StereoMatch.cpp:126-127); the library returns that status up front."""
    from pyopenlpt_amd._lib import FatalError
    ctx = make_ctx(case)
    p = case.params
    old = p.match_cam_count
    p.match_cam_count = 3
    try:
        with pytest.raises(FatalError) as e:
            ctx.stereo_match(case.peaks(), p)
        assert e.value.status == 1
    finally:
        p.match_cam_count = old
    ctx.close()


def test_peak_capacity_error(case):
    from pyopenlpt_amd._lib import FatalError
    from pyopenlpt_amd.context import to_device_image
    ctx = make_ctx(case)
    d, code = to_device_image(case.images)
    with pytest.raises(FatalError) as e:
        ctx.find_peaks(d[0].contiguous(), code, case.params, cap=10)
    assert e.value.status == 102
    ctx.close()


def test_reduced_subsets_match_reference():
    """n_cam_reduced = 1 (4 subsets of 3 active cameras)."""
    c = Case("golden_reduced_256")
    from pyopenlpt_amd.context import to_device_image
    ctx = make_ctx(c, "f64")
    d, code = to_device_image(c.images)
    (ox, _u, n_obj), lt = ctx.ipr_run(d, code, c.params, cap_obj=1 << 15)
    loops = c.z["ipr_loops"]
    assert len(loops) == 2 + 4 * 2
    assert np.array_equal(lt[0][: len(loops)], loops)
    ctx.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("streams,batch", [(1, 1), (3, 1), (1, 5), (1, 3), (2, 2)])
def test_batched_frames_equal_single_frames(case, streams, batch, precision):
    """Frames of one call - on one stream, spread over several workspaces
    and streams, or several frames per kernel launch (frames_per_batch,
    chunks of 5, 3 + 2 and 2 + 2 + 1 frames) - give the single-frame results
    bit for bit."""
    from pyopenlpt_amd.context import to_device_image
    cases = [Case(n) for n in ("golden_u8_256", "golden_reduced_256")]
    cams_ok = all(np.array_equal(cases[0].z["cameras"], c.z["cameras"]) for c in cases)
    if not cams_ok:
        cases = [case, case]
    cases = cases + cases[::-1] + cases[:1]          # 5 frames: uneven over 3 streams
    imgs = np.stack([c.images for c in cases])
    ctx = make_ctx(case, precision)
    ctx.set_tuning(n_streams=streams, frames_per_batch=batch)
    d, code = to_device_image(imgs)
    (ox, ou, n_obj), lt = ctx.ipr_run(d, code, case.params, n_frames=len(cases),
                                      cap_obj=1 << 14)
    single = make_ctx(case, precision)
    for f, c in enumerate(cases):
        d1, _ = to_device_image(c.images)
        (o1, u1, n1), lt1 = single.ipr_run(d1, code, case.params, cap_obj=1 << 14)
        k = int(n1[0].item())
        assert int(n_obj[f].item()) == k, f
        assert torch.equal(ox[f, :k], o1[0, :k]), f
        assert torch.equal(ou[f, :k], u1[0, :k]), f
        assert np.array_equal(lt[f], lt1[0]), f
    assert ctx.diagnostics().tolist() == single.diagnostics().tolist() == [0] * 8
    single.close()
    ctx.close()


def test_graph_replays_match_eager(case):
    from pyopenlpt_amd.context import to_device_image
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        d, code = to_device_image(case.images)
        ctx = make_ctx(case, "f32", stream=s)
        outs_e = ctx.alloc_outputs(1, 1 << 14, d.device)
        ctx.ipr_run(d, code, case.params, cap_obj=1 << 14, outputs=outs_e)
        s.synchronize()
        ref_n = int(outs_e[2][0].item())
        ref_x = outs_e[0][0, :ref_n].clone()
        outs_g = ctx.alloc_outputs(1, 1 << 14, d.device)
        ctx.graph_capture(d, code, case.params, 1, outs_g, 1 << 14)
        for rep in range(3):
            outs_g[0].zero_()
            outs_g[2].zero_()
            ctx.graph_launch()
            s.synchronize()
            n = int(outs_g[2][0].item())
            assert n == ref_n, rep
            assert torch.equal(outs_g[0][0, :n], ref_x), rep
            # eager work between replays
            ctx.ipr_run(d, code, case.params, cap_obj=1 << 14, outputs=outs_e)
            s.synchronize()
    ctx.close()


@pytest.mark.parametrize("batch", [1, 2])
def test_graph_with_streams_matches_eager(case, batch):
    """A captured multi-frame run that forks over two streams (one or two
    frames per launch) replays to the eager result, with eager work between
    replays."""
    from pyopenlpt_amd.context import to_device_image
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        imgs = np.stack([case.images, case.images[:, ::-1, :].copy(), case.images])
        d, code = to_device_image(imgs)
        ctx = make_ctx(case, "f32", stream=s)
        ctx.set_tuning(n_streams=2, frames_per_batch=batch)
        outs_e = ctx.alloc_outputs(3, 1 << 14, d.device)
        ctx.ipr_run(d, code, case.params, n_frames=3, cap_obj=1 << 14, outputs=outs_e)
        s.synchronize()
        ref_n = outs_e[2].clone()
        ref_x = outs_e[0].clone()
        outs_g = ctx.alloc_outputs(3, 1 << 14, d.device)
        ctx.graph_capture(d, code, case.params, 3, outs_g, 1 << 14)
        for rep in range(3):
            outs_g[0].zero_()
            outs_g[2].zero_()
            ctx.graph_launch()
            s.synchronize()
            assert torch.equal(outs_g[2], ref_n), rep
            for f in range(3):
                n = int(ref_n[f].item())
                assert torch.equal(outs_g[0][f, :n], ref_x[f, :n]), (rep, f)
            ctx.ipr_run(d, code, case.params, n_frames=3, cap_obj=1 << 14, outputs=outs_e)
            s.synchronize()
    ctx.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_tuning_leaves_results_unchanged(case, precision):
    """Launch shapes (olpt_set_tuning) change the schedule only: the IPR
    objects, their order and the loop counts are identical under each."""
    from pyopenlpt_amd.context import to_device_image
    imgs = np.stack([case.images, case.images[:, ::-1, :].copy(), case.images])
    d, code = to_device_image(imgs)
    ctx = make_ctx(case, precision)
    ref = None
    for t in ({}, dict(shake_blocks=64, build_blocks=64, prune_blocks=64),
              dict(n_streams=3, shake_blocks=16384, build_blocks=65536, prune_blocks=8192),
              dict(n_streams=1, frames_per_batch=3),
              dict(n_streams=2, shake_blocks=1000, build_blocks=777, prune_blocks=129,
                   frames_per_batch=2)):
        ctx.set_tuning(**t)
        (ox, _u, n_obj), lt = ctx.ipr_run(d, code, case.params, n_frames=3, cap_obj=1 << 14)
        torch.cuda.synchronize()
        got = (n_obj.cpu(), [ox[f, :int(n_obj[f].item())].cpu() for f in range(3)], lt)
        if ref is None:
            ref = got
            assert int(ref[0].sum()) > 0
            continue
        assert torch.equal(got[0], ref[0]), t
        for f in range(3):
            assert torch.equal(got[1][f], ref[1][f]), (t, f)
        assert np.array_equal(got[2], ref[2]), t
    assert ctx.get_tuning() == dict(n_streams=2, shake_blocks=1000, build_blocks=777,
                                    prune_blocks=129, frames_per_batch=2)
    ctx.close()


def test_tuning_rejects_out_of_range_and_autotune_keeps_a_candidate(case):
    from pyopenlpt_amd._lib import FatalError
    from pyopenlpt_amd.context import to_device_image
    ctx = make_ctx(case, "f32")
    before = ctx.get_tuning()
    for bad in (dict(shake_blocks=0), dict(build_blocks=1 << 20), dict(n_streams=9),
                dict(frames_per_batch=0)):
        with pytest.raises(FatalError) as e:
            ctx.set_tuning(**bad)
        assert e.value.status == 1
    assert ctx.get_tuning() == before
    d, code = to_device_image(np.stack([case.images, case.images]))
    space = {"n_streams": (1, 2), "shake_blocks": (1024, 4096), "prune_blocks": (512, 2048)}
    best, log = ctx.autotune(d, code, case.params, 2, space=space, reps=1)
    assert ctx.get_tuning() == best
    # the start, then each knob's other value, each shape timed once
    assert len(log) == 4 and len({tuple(sorted(t.items())) for t, _ in log}) == 4
    assert all(best[k] in v for k, v in space.items())
    ctx.close()


@pytest.mark.parametrize("shape", [(40, 2100), (37, 1090), (9, 1027), (24, 4200), (6, 67),
                                   (3, 3)])
def test_peaks_match_oracle_on_odd_widths(shape):
    """Rows of several 64-pixel chunks per 16-chunk group, partial last groups
    and chunks, and tiny images: the GPU peaks are the oracle's, in order
    (dense bright pixels with plateaus and equal neighbours)."""
    from common import oracle_peaks
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.context import Context, to_device_image
    H, W = shape
    rng = np.random.default_rng(H * 7919 + W)
    img = rng.integers(0, 60, (H, W)).astype(np.uint16)
    n = max(1, H * W // 12)
    img[rng.integers(0, H, n), rng.integers(0, W, n)] = rng.integers(100, 3000, n).astype(np.uint16)
    img[rng.integers(0, H, n // 3), rng.integers(0, W, n // 3)] = 500
    p = S.default_params(0.04, min_obj_int=50)
    want = oracle_peaks(img.astype(np.float64), p)
    cap = max(64, H * W // 2)
    ctx = Context(S.make_rig(H, W, n_cam=2), H, W, precision="f64", max_2d=cap,
                  max_cand=1 << 12, max_obj=1 << 12)
    d, code = to_device_image(img)
    got = ctx.find_peaks(d, code, p, cap=cap)
    assert got.shape == want.shape
    if len(want):
        assert np.abs(got - want).max() <= 1e-9
    ctx.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_residual_box_width_rounds_in_double(case, precision):
    """calROIBound's box edge round(2 r) is taken in double on the host for
    every kernel: at r = 1.2499999999 it is 2 (in FP32, 2 * (float)r = 2.5
    rounds to 3).  The residual then equals the oracle's (FP64 to 1e-9, FP32
    to 2e-3 of the 255 scale)."""
    import ctypes
    import oracle as O
    from pyopenlpt_amd.context import to_device_image
    ctx = make_ctx(case, precision)
    d, code = to_device_image(case.images)
    p = case.params
    old = p.radius_obj
    p.radius_obj = 1.2499999999
    try:
        xyz = case.z["ipr_xyz"][:300]
        got = ctx.residual(d, code, p, xyz)
        otf, _k = case.otf()
        want = np.zeros((case.n_cam, case.H, case.W))
        O.load_port().orc_residual(O.cams_array(case.cams), case.n_cam,
                                   O.dptr(case.images.astype(np.float64)), case.H, case.W,
                                   ctypes.byref(p), ctypes.byref(otf),
                                   O.dptr(np.ascontiguousarray(xyz)), len(xyz), O.dptr(want))
    finally:
        p.radius_obj = old
    err = np.abs(got - want).max()
    print(f"r = 1.2499999999 residual ({precision}): max |diff| {err:.3g}")
    assert err <= (1e-9 if precision == "f64" else 2e-3)
    ctx.close()
