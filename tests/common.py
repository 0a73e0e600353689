"""Helpers shared by the tests: golden fixtures and oracle calls."""
from __future__ import annotations

import ctypes
import glob
import os

import numpy as np

import oracle as O
from pyopenlpt_amd import synth as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
GOLDEN_CASES = sorted(os.path.basename(p)[:-4]
                      for p in glob.glob(os.path.join(GOLDEN, "golden_*.npz")))


class Case:
    """A golden fixture: inputs (images, cameras, params, OTF) and the
    reference's stage outputs."""

    def __init__(self, name: str):
        self.name = name
        self.z = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
        self.images = self.z["images"]
        self.n_cam, self.H, self.W = self.images.shape
        self.cams = [self._cam(r) for r in self.z["cameras"]]
        self.params = self._params(self.z["params"])
        self.abca = tuple(float(v) for v in self.z["otf_abca"])

    @staticmethod
    def _cam(r) -> S.PinholeCamera:
        n_row, n_col, n_dist = int(r[0]), int(r[1]), int(r[2])
        fx, fy, cx, cy = r[4:8]
        dist = r[8:20].copy()
        R = r[20:29].reshape(3, 3)
        T = r[29:32]
        Rinv = r[32:41].reshape(3, 3)
        Tinv = r[41:44]
        return S.PinholeCamera(n_row=n_row, n_col=n_col, fx=fx, fy=fy, cx=cx, cy=cy,
                               dist=dist, n_dist=n_dist, R=R, T=T, Rinv=Rinv, Tinv=Tinv,
                               max_intensity=r[44])

    @staticmethod
    def _params(v) -> S.OlptIprParams:
        p = S.OlptIprParams()
        i = 0
        for f, _t in S.OlptIprParams._fields_:
            if f == "limit":
                for k in range(6):
                    p.limit[k] = v[i + k]
                i += 6
            else:
                cur = getattr(p, f)
                setattr(p, f, type(cur)(v[i]) if not isinstance(cur, float) else float(v[i]))
                i += 1
        return p

    def otf(self):
        """The case's OTF: a full grid when the fixture stores one, else the
        uniform 2x2x2 grid of otf_abca."""
        if "otf_a" in self.z:
            shape = tuple(int(v) for v in self.z["otf_shape"])
            return S.make_otf_grid(self.n_cam, shape, self.z["otf_boundary"], self.z["otf_a"],
                                   self.z["otf_b"], self.z["otf_c"], self.z["otf_alpha"])
        return S.make_otf(self.n_cam, self.abca)

    def peaks(self):
        return [self.z[f"peaks_cam{c}"] for c in range(self.n_cam)]


# --------------------------------------------------------------------------
# oracle wrappers
# --------------------------------------------------------------------------
def oracle_peaks(img: np.ndarray, params) -> np.ndarray:
    L = O.load_port()
    H, W = img.shape
    xy = np.zeros((H * W // 2 + 16, 2))
    n = ctypes.c_int32()
    rc = L.orc_find_peaks(O.dptr(np.ascontiguousarray(img, dtype=np.float64)), H, W,
                          ctypes.byref(params), len(xy), O.dptr(xy), ctypes.byref(n))
    assert rc == 0
    return xy[: n.value].copy()


def oracle_stereo(cams, pts, params, active=None):
    L = O.load_port()
    n_cam = len(cams)
    cap = 1 << 21
    cid = np.zeros((cap, n_cam), np.int32)
    ce = np.zeros(cap)
    si = np.zeros(cap, np.int32)
    ox = np.zeros((cap, 3))
    oi = np.zeros((cap, n_cam), np.int32)
    cnt = np.zeros(4, np.int32)
    npts = np.array([len(a) for a in pts], np.int32)
    act = None if active is None else O.iptr(np.asarray(active, np.int32))
    rc = L.orc_stereo_match(O.cams_array(cams), n_cam, act, O.dptr(np.concatenate(pts)),
                            O.iptr(npts), ctypes.byref(params), cap, O.iptr(cid), O.dptr(ce),
                            O.iptr(si), cap, O.dptr(ox), O.iptr(oi), O.iptr(cnt))
    M, ns, no = (int(x) for x in cnt[:3])
    return rc, dict(cand_ids=cid[:M], cand_e=ce[:M], sel=si[:ns], obj_xyz=ox[:no],
                    obj_ids=oi[:no])


def _otf(n_cam, otf):
    """An OlptOtf (or (struct, keepalive) pair) as given, or the uniform grid
    of an (a, b, c, alpha) tuple."""
    if isinstance(otf, tuple) and len(otf) == 2 and isinstance(otf[0], S.OlptOtf):
        return otf
    if isinstance(otf, S.OlptOtf):
        return otf, None
    return S.make_otf(n_cam, otf)


def oracle_shake(cams, images, params, abca, xyz, active=None, residual=True):
    L = O.load_port()
    n_cam, H, W = images.shape
    otf, _k = _otf(n_cam, abca)
    n = len(xyz)
    xo = np.zeros((max(n, 1), 3))
    fl = np.zeros(max(n, 1), np.uint8)
    sc = np.zeros(max(n, 1))
    res = np.zeros((n_cam, H, W)) if residual else None
    act = None if active is None else O.iptr(np.asarray(active, np.int32))
    rc = L.orc_shake(O.cams_array(cams), n_cam, act,
                     O.dptr(np.ascontiguousarray(images, dtype=np.float64)), H, W,
                     ctypes.byref(params), ctypes.byref(otf),
                     O.dptr(np.ascontiguousarray(xyz, dtype=np.float64)), n, O.dptr(xo),
                     O.u8ptr(fl), O.dptr(sc), O.dptr(res) if residual else None)
    assert rc == 0
    return xo[:n], fl[:n], sc[:n], res


def oracle_ipr(cams, images, params, abca):
    L = O.load_port()
    n_cam, H, W = images.shape
    otf, _k = _otf(n_cam, abca)
    cap = 1 << 19
    ox = np.zeros((cap, 3))
    ou = np.zeros((cap, n_cam, 2))
    lt = np.zeros(256, np.int32)
    cnt = np.zeros(2, np.int32)
    rc = L.orc_run_ipr(O.cams_array(cams), n_cam,
                       O.dptr(np.ascontiguousarray(images, dtype=np.float64)), H, W,
                       ctypes.byref(params), ctypes.byref(otf), cap, O.dptr(ox), O.dptr(ou),
                       256, O.iptr(lt), O.iptr(cnt))
    assert rc == 0
    return ox[: cnt[0]], ou[: cnt[0]], lt[: cnt[1]]


# --------------------------------------------------------------------------
# FP32-mode position comparison
# --------------------------------------------------------------------------
def f32_position_stats(got: np.ndarray, got_loops, want: np.ndarray, want_loops,
                       tol_mm: float = 1e-4, near_mm: float = 0.0) -> dict:
    """Object-by-object comparison of an FP32-mode IPR output with the FP64
    reference.  Loop 1's objects come from the same (FP64) stereo candidates in
    both, so when loop 1 kept the same number of objects they correspond by
    index: `loop1_share` is the share beyond tol_mm, `loop1_q99` the 99th
    percentile and `loop1_max` the largest deviation.  Later loops run on
    residual images that differ in the last bits, so their objects are
    compared by nearest neighbour: `nn_share` is the share of all output
    objects with no reference object within tol_mm, `far_share` the share
    with none within near_mm."""
    from scipy.spatial import cKDTree
    out = dict(n_got=len(got), n_want=len(want))
    n1g, n1w = int(got_loops[0]), int(want_loops[0])
    out["loop1_equal"] = n1g == n1w
    if n1g == n1w and n1g:
        d = np.abs(got[:n1g] - want[:n1w]).max(axis=1)
        out["loop1_share"] = float((d > tol_mm).mean())
        out["loop1_q99"] = float(np.quantile(d, 0.99))
        out["loop1_max"] = float(d.max())
    if len(got) and len(want):
        d, _ = cKDTree(want).query(got)
        out["nn_share"] = float((d > tol_mm).mean())
        out["far_share"] = float((d > near_mm).mean())
    return out
