"""Learned shake proposal entry points: rendering and network inputs against
independent numpy statements, and the IPR loop with a model set.

Rendering and patches are checked in FP64 mode (<= 1e-9 before the BF16
rounding of the patches; the patches themselves to BF16 resolution, relative
2^-8).  The numpy statements follow the reference's definitions: calROIBound
(Shake.cpp:246-277, round half away from zero, empty at or beyond the image
border), the Gaussian intensity (Shake.cpp:901-922) and the augmented image
(Shake.cpp:279-328)."""
import math

import numpy as np
import pytest
import torch

from common import Case

pytestmark = pytest.mark.gpu


def _round(v):
    return math.copysign(math.floor(abs(v) + 0.5), v)


def roi_bound(H, W, cx, cy, d):
    if cx <= 0 or cy <= 0 or cx >= W or cy >= H:
        return None
    w = int(_round(2 * d))
    cmin = int(_round(cx - (w - 1) / 2.0))
    rmin = int(_round(cy - (w - 1) / 2.0))
    cmax, rmax = cmin + w, rmin + w
    return max(rmin, 0), min(rmax, H), max(cmin, 0), min(cmax, W)


def proj_box(cam, X, amp, b, c, r, H, W):
    """[(rows, cols, values)] of one particle's projection in one camera."""
    u, v = cam.project(X[None])[0]
    box = roi_bound(H, W, u, v, r)
    if box is None:
        return None
    r0, r1, c0, c1 = box
    rr, cc = np.mgrid[r0:r1, c0:c1]
    dx, dy = cc - u, rr - v
    val = np.maximum(amp * np.exp(-(b * dx * dx + c * dy * dy)), 0.0)
    return rr, cc, val, (u, v)


@pytest.fixture(scope="module")
def setup(gpu_available):
    from pyopenlpt_amd.context import Context
    case = Case("golden_u8_256")
    ctx = Context(case.cams, case.H, case.W, precision="f64", max_2d=1 << 15,
                  max_cand=1 << 18, max_obj=1 << 15)
    otf, keep = case.otf()
    ctx.set_otf(otf)
    ctx._keep = keep
    yield case, ctx
    ctx.close()


def test_render_matches_numpy(setup):
    case, ctx = setup
    rng = np.random.default_rng(4)
    xyz = rng.uniform(-18, 18, (150, 3))
    amp = rng.uniform(80, 160, 150)
    got = ctx.render(torch.from_numpy(xyz).cuda(), torch.from_numpy(amp).cuda(), r_px=2.5)
    got = got.cpu().numpy()
    _, b, c, _ = case.abca
    want = np.zeros((case.n_cam, case.H, case.W))
    for k, cam in enumerate(case.cams):
        for i in range(len(xyz)):
            pb = proj_box(cam, xyz[i], amp[i], b, c, 2.5, case.H, case.W)
            if pb is None:
                continue
            rr, cc, val, _ = pb
            want[k, rr, cc] = np.maximum(want[k, rr, cc], val)
    assert np.abs(got - want).max() <= 1e-9


@pytest.mark.parametrize("r_px", [2.5, 12.0])
def test_render_partial_tiles_dense_cluster(gpu_available, r_px):
    """The residual/render kernel on a 250 x 270 frame (partial 32-pixel
    tiles at the right and bottom edges), boxes up to the largest supported
    (24 px at r = 12), and a cluster dense enough that the objects reaching
    one tile exceed one LDS batch of 256."""
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.context import Context
    H, W = 250, 270
    cams = S.make_rig(H, W)
    ctx = Context(cams, H, W, precision="f64", max_2d=1 << 14, max_cand=1 << 16,
                  max_obj=1 << 12)
    otf, keep = S.make_otf(len(cams), (125.0, 1.5, 1.5, 0.0))
    ctx.set_otf(otf)
    ctx._keep = keep
    rng = np.random.default_rng(11)
    xyz = np.concatenate([rng.uniform(-2.0, 2.0, (400, 3)) + np.array([5.0, -3.0, 2.0]),
                          rng.uniform(-21.0, 21.0, (300, 3))])
    amp = rng.uniform(80, 160, len(xyz))
    try:
        got = ctx.render(torch.from_numpy(xyz).cuda(), torch.from_numpy(amp).cuda(), r_px=r_px)
        got = got.cpu().numpy()
    finally:
        ctx.close()
    want = np.zeros((len(cams), H, W))
    for k, cam in enumerate(cams):
        for i in range(len(xyz)):
            pb = proj_box(cam, xyz[i], amp[i], 1.5, 1.5, r_px, H, W)
            if pb is None:
                continue
            rr, cc, val, _ = pb
            want[k, rr, cc] = np.maximum(want[k, rr, cc], val)
    assert (want > 0).sum() > 0.05 * want.size
    assert np.abs(got - want).max() <= 1e-9


def test_patches_match_numpy(setup):
    case, ctx = setup
    from pyopenlpt_amd import _lib
    a, b, c, _ = case.abca
    img = case.images.astype(np.float64)
    xyz = case.z["sm_obj_xyz"][:60]
    H, W, r = case.H, case.W, case.params.radius_obj
    # residual of all objects (active cameras, no clamp), then aug windows
    boxes = [[proj_box(cam, xyz[i], a, b, c, r, H, W) for i in range(len(xyz))]
             for cam in case.cams]
    res = img.copy()
    for k in range(case.n_cam):
        mx = np.zeros((H, W))
        for pb in boxes[k]:
            if pb is not None:
                rr, cc, val, _ = pb
                mx[rr, cc] = np.maximum(mx[rr, cc], val)
        res[k] = img[k] - mx
    want = np.zeros((len(xyz), 512))
    for i in range(len(xyz)):
        for k, cam in enumerate(case.cams):
            u, v = cam.project(xyz[i][None])[0]
            cu, cv = int(_round(u)), int(_round(v))
            for p in range(121):
                row, col = cv + p // 11 - 5, cu + p % 11 - 5
                if not (0 <= row < H and 0 <= col < W):
                    continue
                val = res[k, row, col]
                pb = boxes[k][i]
                if pb is not None:
                    rr, cc, pv, _ = pb
                    hit = (rr == row) & (cc == col)
                    if hit.any():
                        val += pv[hit][0]
                want[i, k * 121 + p] = max(val, 0.0) / a
    got = ctx.shake_patches(torch.from_numpy(img).cuda(), _lib.OLPT_F64, case.params,
                            torch.from_numpy(np.ascontiguousarray(xyz)).cuda())
    got = got.float().cpu().numpy()
    assert got.shape == want.shape
    assert np.abs(got[:, 484:]).max() == 0.0
    tol = 2.0 ** -8 * np.maximum(np.abs(want), 1e-3) + 1e-6
    assert (np.abs(got - want) <= tol).all(), np.abs(got - want).max()


def test_learned_ipr_runs_and_plain_is_restored(gpu_available):
    from pyopenlpt_amd.context import Context, to_device_image
    from pyopenlpt_amd.learned import ShakeNet
    case = Case("golden_u8_256")
    ctx = Context(case.cams, case.H, case.W, precision="f32", max_2d=1 << 15,
                  max_cand=1 << 18, max_obj=1 << 15)
    otf, keep = case.otf()
    ctx.set_otf(otf)
    d, code = to_device_image(case.images)
    (ox, _u, n0), _ = ctx.ipr_run(d, code, case.params, cap_obj=1 << 14)
    n_plain = int(n0[0].item())
    ref = ox[0, :n_plain].clone()
    net = ShakeNet(max_batch=1024, seed=2)
    vox = case.params.tol_3d_mm   # tol_3d is one voxel in these cases
    ctx.set_shake_model(net, vox, 0.5)
    (ox, _u, n1), _ = ctx.ipr_run(d, code, case.params, cap_obj=1 << 14)
    n_learned = int(n1[0].item())
    assert n_learned > 0.8 * n_plain
    assert torch.isfinite(ox[0, :n_learned]).all()
    ctx.set_shake_model(None)
    (ox, _u, n2), _ = ctx.ipr_run(d, code, case.params, cap_obj=1 << 14)
    assert int(n2[0].item()) == n_plain
    assert torch.equal(ox[0, :n_plain], ref)
    ctx.close()
    net.close()


def test_ghost_logit_threshold(gpu_available):
    """olpt_set_ghost_logit: an infinite threshold changes nothing; a
    threshold below every logit flags every object of every loop a ghost,
    so the IPR output is empty; a finite threshold leaves a subset."""
    from pyopenlpt_amd.context import Context, to_device_image
    from pyopenlpt_amd.learned import ShakeNet
    case = Case("golden_u8_256")
    ctx = Context(case.cams, case.H, case.W, precision="f32", max_2d=1 << 15,
                  max_cand=1 << 18, max_obj=1 << 15)
    otf, keep = case.otf()
    ctx.set_otf(otf)
    d, code = to_device_image(case.images)
    net = ShakeNet(max_batch=1024, seed=2)
    vox = case.params.tol_3d_mm

    def run(thr):
        ctx.set_shake_model(net, vox, 0.5, ghost_logit=thr)
        (ox, _u, n), _ = ctx.ipr_run(d, code, case.params, cap_obj=1 << 14)
        k = int(n[0].item())
        return ox[0, :k].clone()

    base = run(None)
    assert len(base) > 0
    assert torch.equal(run(float("inf")), base)
    assert len(run(-1e30)) == 0
    assert len(run(0.0)) <= len(base) + len(base) // 2
    ctx.set_shake_model(None)
    ctx.close()
    net.close()
