#!/usr/bin/env python3
"""Write golden vectors from the reference's own C++ IPR path.

Runs oracle/_ref/libolpt_ref.so (the reference sources compiled by
oracle/Makefile, driven through ref_harness.cpp) on small seeded synthetic
frames and stores inputs and stage outputs as compressed .npz fixtures:

  images, cameras, params, OTF   inputs (self-contained, no RNG needed)
  peaks_cam{k}                   ObjectFinder2D::findObject2D per camera
  sm_cand_ids / sm_cand_e /      StereoMatch stages on those peaks
  sm_sel / sm_obj_xyz / sm_obj_ids
  sh_xyz / sh_flags / sh_score / Shake::runShake on the matched objects +
  sh_res_sum / sh_res_sample     IPR-mode residual (per-camera sums and a
                                 fixed sample of pixels)
  ipr_xyz / ipr_uv / ipr_loops   IPR::runIPR

Usage: python tests/golden/make_golden.py [case ...]  (needs the reference checkout;
no names = every case)
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from pyopenlpt_amd import synth as S  # noqa: E402

CAM_FIELDS = [f for f, _ in S.OlptCamera._fields_]
PARAM_FIELDS = [f for f, _ in S.OlptIprParams._fields_]


def cams_to_array(cams) -> np.ndarray:
    rows = []
    for c in cams:
        cc = c.to_c()
        v = [cc.n_row, cc.n_col, cc.n_dist, cc.is_distorted, cc.fx, cc.fy, cc.cx, cc.cy]
        v += list(cc.dist) + list(cc.R) + list(cc.T) + list(cc.Rinv) + list(cc.Tinv)
        v += [cc.max_intensity]
        rows.append(v)
    return np.array(rows, dtype=np.float64)


def params_to_array(p) -> np.ndarray:
    v = []
    for f in PARAM_FIELDS:
        x = getattr(p, f)
        if f == "limit":
            v += list(x)
        else:
            v.append(float(x))
    return np.array(v)


WANT = set(sys.argv[1:])   # case names to write; empty = all


def run_case(name: str, w, otf_abca, *, reduced: int = 0, sample: int = 4096, grid=None):
    """`grid` = dict(shape, boundary, a, b, c, alpha) gives a full OTF grid
    (arrays [n_cam][nx*ny*nz]); otherwise the uniform 2x2x2 grid of otf_abca."""
    if WANT and name not in WANT:
        return
    lib = O.load_ref()
    p = w.params
    p.n_cam_reduced = reduced
    H, W = w.images.shape[1:]
    n_cam = len(w.cams)
    cams = O.cams_array(w.cams)
    if grid is None:
        otf, _keep = S.make_otf(n_cam, otf_abca)
    else:
        otf, _keep = S.make_otf_grid(n_cam, grid["shape"], grid["boundary"], grid["a"],
                                     grid["b"], grid["c"], grid["alpha"])
    imgs = w.images.astype(np.float64).copy()
    out = dict(images=w.images, cameras=cams_to_array(w.cams), params=params_to_array(p),
               otf_abca=np.array(otf_abca, dtype=np.float64), truth=w.truth)
    if grid is not None:
        out.update(otf_shape=np.array(grid["shape"], np.int32),
                   otf_boundary=np.array(grid["boundary"], np.float64),
                   **{f"otf_{k}": np.asarray(grid[k], np.float64).reshape(n_cam, -1)
                      for k in ("a", "b", "c", "alpha")})
    # peaks
    pts = []
    for c in range(n_cam):
        xy = np.zeros((1 << 18, 2))
        n = ctypes.c_int32()
        rc = lib.ref_find_peaks(O.dptr(imgs[c].copy()), H, W, ctypes.byref(p), 1 << 18,
                                O.dptr(xy), ctypes.byref(n))
        assert rc == 0
        pts.append(xy[: n.value].copy())
        out[f"peaks_cam{c}"] = pts[-1]
    # stereo on the peaks
    cap = 1 << 20
    cid = np.zeros((cap, n_cam), np.int32)
    ce = np.zeros(cap)
    si = np.zeros(cap, np.int32)
    ox = np.zeros((cap, 3))
    oi = np.zeros((cap, n_cam), np.int32)
    cnt = np.zeros(4, np.int32)
    npts = np.array([len(a) for a in pts], np.int32)
    rc = lib.ref_stereo_match(cams, n_cam, None, O.dptr(np.concatenate(pts)), O.iptr(npts),
                              ctypes.byref(p), cap, O.iptr(cid), O.dptr(ce), O.iptr(si), cap,
                              O.dptr(ox), O.iptr(oi), O.iptr(cnt))
    assert rc == 0 and cnt[3] == 1, (rc, cnt)
    M, ns, no = (int(x) for x in cnt[:3])
    out.update(sm_cand_ids=cid[:M], sm_cand_e=ce[:M], sm_sel=si[:ns], sm_obj_xyz=ox[:no],
               sm_obj_ids=oi[:no])
    # shake on the matched objects + IPR-mode residual of the kept ones
    xo = np.zeros((no, 3))
    fl = np.zeros(no, np.uint8)
    sc = np.zeros(no)
    res = np.zeros((n_cam, H, W))
    rc = lib.ref_shake(cams, n_cam, None, O.dptr(imgs), H, W, ctypes.byref(p),
                       ctypes.byref(otf), O.dptr(ox[:no].copy()), no, O.dptr(xo), O.u8ptr(fl),
                       O.dptr(sc), O.dptr(res))
    assert rc == 0
    rng = np.random.default_rng(7)
    idx = rng.integers(0, H * W, size=sample)
    out.update(sh_xyz=xo, sh_flags=fl, sh_score=sc, sh_res_sum=res.reshape(n_cam, -1).sum(1),
               sh_res_idx=idx, sh_res_sample=res.reshape(n_cam, -1)[:, idx])
    # shake with planted flags: duplicates within tol_3d (Repeated) and
    # positions with no particle (Ghost), interleaved with the matches
    if no > 40:
        g = np.random.default_rng(99)
        dup = ox[:20] + g.normal(0, 0.1 * p.tol_3d_mm, (20, 3))
        phantom = g.uniform(-15, 15, (30, 3))
        # out of every camera's view: fewer than 2 usable cameras -> score
        # 1e-12 (Shake.cpp:742-744) -> Ghost; in-view phantoms cannot fall
        # below 0.1 x mean because each factor is clamped to [2/3, 3/2]
        far = g.uniform(300, 400, (10, 3))
        xin = np.concatenate([ox[:no], dup, phantom, far])
        xin = xin[g.permutation(len(xin))]
        n2 = len(xin)
        xo2 = np.zeros((n2, 3))
        fl2 = np.zeros(n2, np.uint8)
        sc2 = np.zeros(n2)
        rc = lib.ref_shake(cams, n_cam, None, O.dptr(imgs), H, W, ctypes.byref(p),
                           ctypes.byref(otf), O.dptr(xin.copy()), n2, O.dptr(xo2),
                           O.u8ptr(fl2), O.dptr(sc2), None)
        assert rc == 0
        out.update(shf_xyz_in=xin, shf_xyz=xo2, shf_flags=fl2, shf_score=sc2)
    # IPR::runIPR
    capo = 1 << 18
    ix = np.zeros((capo, 3))
    iu = np.zeros((capo, n_cam, 2))
    lt = np.zeros(64, np.int32)
    c2 = np.zeros(2, np.int32)
    rc = lib.ref_run_ipr(cams, n_cam, O.dptr(imgs), H, W, ctypes.byref(p), ctypes.byref(otf),
                         capo, O.dptr(ix), O.dptr(iu), 64, O.iptr(lt), O.iptr(c2))
    assert rc == 0
    out.update(ipr_xyz=ix[: c2[0]], ipr_uv=iu[: c2[0]], ipr_loops=lt[: c2[1]])
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    extra = (f" planted-flags {np.bincount(out['shf_flags'], minlength=4).tolist()}"
             if "shf_flags" in out else "")
    print(f"{name}: peaks {npts.tolist()} cand {M} sel {ns} obj {no} "
          f"flags {np.bincount(fl, minlength=4).tolist()}{extra} ipr {int(c2[0])} loops "
          f"{lt[: c2[1]].tolist()} -> {os.path.getsize(path) // 1024} KB")


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libolpt_ref.so is missing: run `make -C oracle ref`")
    # 256^2 u8, uniform, as the 512^2 CPU-parity case at a quarter of the area
    w = S.workload_c1(seed=11, n=600, size=256)
    run_case("golden_u8_256", w, (125.0, 1.5, 1.5, 0.0))
    # u16 with a ~ U(800,1200), b = c = 1.2, noise 10, 6 shakes
    cams = S.make_rig(256, 256, max_intensity=65535)
    pts = S.uniform_particles(700, 21)
    amp = np.random.default_rng(22).uniform(800, 1200, 700)
    img = S.render(cams, pts, amp, 1.2, 1.2, 0.0, 10.0, 23, np.uint16)
    vox = 40.0 / (0.95 * 256)
    p = S.default_params(vox, n_shake_loop=6, min_obj_int=240)
    run_case("golden_u16_256", S.Workload("u16", cams, img, pts, p, (1000.0, 1.2, 1.2, 0.0),
                                          (-20, 20) * 3, vox), (1000.0, 1.2, 1.2, 0.0))
    # distorted cameras, radial + tangential (distort + fixed-point undistort);
    # larger coefficients bend the far-field epipolar points the reference
    # projects (StereoMatch.cpp:1053-1066) beyond the 2D tolerance
    cams = S.make_rig(256, 256, dist_coeff=(5e-3, 0.0, 1e-5, -1e-5, 0.0),
                      max_intensity=255)
    pts = S.uniform_particles(500, 31)
    img = S.render(cams, pts, 125.0, 1.5, 1.5, 0.0, 2.0, 32, np.uint8)
    p = S.default_params(vox, min_obj_int=30)
    run_case("golden_distorted_256", S.Workload("dist", cams, img, pts, p,
                                                (125.0, 1.5, 1.5, 0.0), (-20, 20) * 3, vox),
             (125.0, 1.5, 1.5, 0.0))
    # reduced-camera passes (n_cam_reduced = 1: 4 subsets of 3 cameras)
    w = S.workload_c1(seed=41, n=800, size=256)
    w.params.n_loop_ipr = 2
    w.params.n_loop_ipr_reduced = 2
    run_case("golden_reduced_256", w, (125.0, 1.5, 1.5, 0.0), reduced=1)
    # limit2DObjectNumber (IPR.cpp:19-39): ~740 peaks per camera cut to 500
    # by the reference's std::shuffle (the two-positions-per-draw branch)
    w = S.workload_c1(seed=61, n=800, size=256)
    w.params.n_obj2d_process_max = 500
    run_case("golden_limit_256", w, (125.0, 1.5, 1.5, 0.0))
    # non-uniform OTF: a 3 x 4 x 2 grid (unequal axes, so a swapped index
    # order changes the interpolated values) whose box [-18,18] x [-20,20] x
    # [-15,15] is smaller than the particle volume (clamping at the box);
    # a, b, c and alpha vary per camera and node, alpha up to +-0.3 rad
    shape = (3, 4, 2)
    ng = int(np.prod(shape))
    g = np.random.default_rng(51)
    grid = dict(shape=shape, boundary=(-18.0, 18.0, -20.0, 20.0, -15.0, 15.0),
                a=g.uniform(105.0, 145.0, (4, ng)), b=g.uniform(1.2, 1.8, (4, ng)),
                c=g.uniform(1.2, 1.8, (4, ng)), alpha=g.uniform(-0.3, 0.3, (4, ng)))
    cams = S.make_rig(256, 256, max_intensity=255)
    pts = S.uniform_particles(600, 52)
    img = S.render(cams, pts, g.uniform(105.0, 145.0, 600), 1.5, 1.5, 0.0, 2.0, 53, np.uint8)
    p = S.default_params(vox, min_obj_int=30)
    run_case("golden_otfgrid_256", S.Workload("otfgrid", cams, img, pts, p,
                                              (125.0, 1.5, 1.5, 0.0), (-20, 20) * 3, vox),
             (125.0, 1.5, 1.5, 0.0), grid=grid)


if __name__ == "__main__":
    main()
