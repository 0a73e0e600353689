"""The reference's Python classes for the IPR path, over the C-ABI.

Class, method and argument names follow the reference's pybind module
(src/pybind_OpenLPT): ``ObjectFinder2D().findObject2D(image, obj_cfg)``
(pyObjectFinder.cpp:12-22), ``StereoMatch(cams, obj2d_by_cam, obj_cfg)``
with ``match()`` and ``pruneMatch(match_candidates, e_checks)``
(pyStereoMatch.cpp:119-180), ``Shake(cams, obj_cfg)`` with
``runShake(objs, img_orig) -> (flags, objs)``, ``run_Shake`` and
``calResidualImage`` (pyShake.cpp:131-215), ``OTF.getOTFParam(cam_id,
pt_world) -> (a, b, c, cos alpha)`` (pyOTF.cpp:49-53) and ``IPR(cams)`` with
``runIPR(cfg, images)`` and ``saveObjInfo(filename, objs, cfg)``
(pyIPR.cpp:14-51).  Every computation runs in libolpt_hip.so on the GPU.

Inputs: cameras are ``synth.PinholeCamera`` (``camfile.load_camera`` reads
the reference's camera files); ``obj_cfg`` / ``cfg`` is ``config.TracerConfig``
(its ``otf`` an ``config.Otf``); images are numpy arrays (uint8, uint16,
float32 or float64, row-major, x = column) - ``Image.from_numpy`` keeps the
dtype, where the reference's accepts only uint8.  As in the reference, the
constructors copy their inputs and ``runShake`` returns new objects; the
copies keep ``_r2d_px`` and ``_r_px`` (the reference's copy drops them,
pybind_utils.h:23,50).  ``precision`` picks the projection/shake arithmetic:
"f32" (default) or "f64" (the reference's, to rounding).
"""
from __future__ import annotations

import copy
import ctypes
import dataclasses
import enum

import numpy as np

from . import config as C
from . import synth as S

__all__ = ["Image", "Tracer2D", "Tracer3D", "ObjFlag", "ObjectFinder2D", "StereoMatch", "Shake",
           "OTF", "IPR"]


class ObjFlag(enum.IntFlag):
    """Shake.h:19 (``None`` is spelled NONE here)."""
    NONE = 0
    Ghost = 1
    Repeated = 2


@dataclasses.dataclass
class Tracer2D:
    """ObjectInfo.h Tracer2D: centre (x = col, y = row) and radius in px."""
    _pt_center: np.ndarray
    _r_px: float = 2.0

    def clone(self) -> "Tracer2D":
        return Tracer2D(np.array(self._pt_center, np.float64), self._r_px)


@dataclasses.dataclass
class Tracer3D:
    """ObjectInfo.h Tracer3D: world centre (mm), per-camera 2D observations
    (None where a camera has none) and the 2D radius."""
    _pt_center: np.ndarray
    _obj2d_list: list
    _r2d_px: float = 2.0
    _is_tracked: bool = False


class Image:
    """Image.from_numpy (pyMatrix.cpp:147-166) for uint8, uint16, float32
    and float64 arrays: the array itself, contiguous, dtype kept."""

    @staticmethod
    def from_numpy(arr) -> np.ndarray:
        a = np.ascontiguousarray(arr)
        if a.dtype not in (np.uint8, np.uint16, np.float32, np.float64) or a.ndim != 2:
            raise TypeError("Image.from_numpy takes a 2D uint8/uint16/float32/float64 array")
        return a


# --------------------------------------------------------------------------
def _cams_copy(cams) -> list:
    return [copy.deepcopy(c) for c in cams]


def _context(cams, precision: str, otf: "C.Otf | None" = None, **kw):
    from .context import Context
    H, W = int(cams[0].n_row), int(cams[0].n_col)
    ctx = Context(cams, H, W, precision=precision, **kw)
    if otf is not None:
        o, keep = otf.to_c()
        ctx.set_otf(o)
        ctx._otf_keep = keep
    return ctx


def _stack(images) -> np.ndarray:
    imgs = [Image.from_numpy(i) for i in images]
    if len({i.dtype for i in imgs}) != 1:
        raise TypeError("all images must share one dtype")
    return np.ascontiguousarray(np.stack(imgs))


def _xyz(objs) -> np.ndarray:
    return np.array([np.asarray(o._pt_center, np.float64) for o in objs], np.float64).reshape(-1, 3)


def _otf_of(cfg) -> "C.Otf":
    if getattr(cfg, "otf", None) is None:
        raise ValueError("obj_cfg.otf is not set (TracerConfig.readConfig or an Otf)")
    return cfg.otf


# --------------------------------------------------------------------------
class ObjectFinder2D:
    """ObjectFinder2D::findObject2D, tracer branch (ObjectFinder.cpp:37-111)."""

    def findObject2D(self, image, obj_cfg) -> list:  # noqa: N802 - reference name
        import torch
        from . import _lib
        from .context import ptr, to_device_image
        img = Image.from_numpy(image)
        H, W = img.shape
        d, code = to_device_image(img)
        p = obj_cfg.ipr_params()
        cap = max(16, H * W // 2)
        if H <= 4096 and W <= 4096:
            xy = torch.empty((cap, 2), dtype=torch.float64, device=d.device)
            n = torch.zeros(1, dtype=torch.int32, device=d.device)
            st = torch.cuda.current_stream(d.device)
            _lib.check(_lib.lib().olpt_find_peaks_static(
                ctypes.c_void_p(st.cuda_stream), ptr(d), code, H, W, ctypes.byref(p), ptr(xy),
                cap, ptr(n)))
            st.synchronize()
            pts = xy[: int(n.item())].cpu().numpy()
        else:
            cam = S.make_rig(H, W, n_cam=1)[0]
            ctx = _context([cam], "f64")
            pts = ctx.find_peaks(d, code, p, cap=cap)
            ctx.close()
        r = float(obj_cfg.radius_obj)
        return [Tracer2D(np.array(q), r) for q in pts]


class StereoMatch:
    """StereoMatch(cams, obj2d_by_cam, obj_cfg) (StereoMatch.cpp:20-64); the
    constructor copies the cameras and observations."""

    def __init__(self, cams, obj2d_by_cam, obj_cfg, precision: str = "f32"):
        self._cams = _cams_copy(cams)
        self._obj2d = [[o.clone() for o in lst] for lst in obj2d_by_cam]
        self._cfg = obj_cfg
        self._precision = precision
        self._active = [1] * len(self._cams)

    def _pts(self):
        return [np.array([o._pt_center for o in lst], np.float64).reshape(-1, 2)
                for lst in self._obj2d]

    def _ctx(self):
        n = max(1, max(len(lst) for lst in self._obj2d))
        return _context(self._cams, self._precision, max_2d=max(1024, n))

    def match(self) -> list:
        """match() (StereoMatch.cpp:65-178): the triangulated selection, each
        object carrying copies of its matched observations."""
        ctx = self._ctx()
        try:
            r = ctx.stereo_match(self._pts(), self._cfg.ipr_params(), candidates=False)
        finally:
            ctx.close()
        out = []
        r2d = float(self._cfg.radius_obj)
        for x, ids in zip(r["obj_xyz"], r["obj_ids"]):
            o2d = [self._obj2d[c][i].clone() if i >= 0 else None for c, i in enumerate(ids)]
            out.append(Tracer3D(np.array(x), o2d, r2d))
        return out

    def pruneMatch(self, match_candidates, e_checks) -> list:  # noqa: N802 - reference name
        """pruneMatch (StereoMatch.cpp:631-980): the selected id vectors in
        candidate order."""
        ids = np.asarray(match_candidates, np.int32).reshape(len(e_checks), len(self._cams))
        ctx = self._ctx()
        try:
            sel = ctx.prune_match(ids, np.asarray(e_checks, np.float64),
                                  [len(lst) for lst in self._obj2d], self._cfg.ipr_params())
        finally:
            ctx.close()
        return [list(map(int, ids[j])) for j in np.flatnonzero(sel)]


class Shake:
    """Shake(cams, obj_cfg) (Shake.cpp:36-244); the constructor copies the
    cameras and keeps the configuration."""

    def __init__(self, cams, obj_cfg, precision: str = "f32"):
        self._cams = _cams_copy(cams)
        self._cfg = obj_cfg
        self._precision = precision

    def _ctx(self, n_obj: int):
        return _context(self._cams, self._precision, _otf_of(self._cfg),
                        max_obj=max(1024, int(n_obj)))

    def _run(self, objs, img_orig):
        from .context import to_device_image
        d, code = to_device_image(_stack(img_orig))
        xyz = _xyz(objs)
        ctx = self._ctx(len(xyz))
        try:
            xo, fl, _sc, _res = ctx.shake(d, code, self._cfg.ipr_params(), xyz,
                                          want_residual=False)
        finally:
            ctx.close()
        return xo, fl

    def runShake(self, objs, img_orig):  # noqa: N802 - reference name
        """(flags, new objects): the shaken copies, their observations
        re-projected from the new positions (runShake reprojects,
        Shake.cpp:61-65, 441-660)."""
        xo, fl = self._run(objs, img_orig)
        out = []
        proj = [cam.project(xo) for cam in self._cams] if len(xo) else []
        for i, o in enumerate(objs):
            o2d = [Tracer2D(np.array(proj[k][i]), getattr(o, "_r2d_px", 2.0))
                   for k in range(len(self._cams))]
            out.append(Tracer3D(np.array(xo[i]), o2d, getattr(o, "_r2d_px", 2.0)))
        return [ObjFlag(int(f)) for f in fl], out

    def run_Shake(self, objs, img_orig):  # noqa: N802 - reference name
        return [ObjFlag(int(f)) for f in self._run(objs, img_orig)[1]]

    def calResidualImage(self, objs, img_orig, flags=None) -> list:  # noqa: N802
        """IPR-mode residual (all cameras, clamped at 0) of the objects not
        flagged (Shake.h:40-48)."""
        from .context import to_device_image
        keep = [o for i, o in enumerate(objs) if flags is None or int(flags[i]) == 0]
        d, code = to_device_image(_stack(img_orig))
        ctx = self._ctx(len(keep))
        try:
            res = ctx.residual(d, code, self._cfg.ipr_params(), _xyz(keep))
        finally:
            ctx.close()
        return list(res)


class OTF:
    """OTF (OTF.h:33-78): an OTF file path, a config.Otf, or nothing (then
    loadParam)."""

    def __init__(self, otf=None):
        self._param = C.load_otf(otf) if isinstance(otf, str) else otf

    def loadParam(self, *args):  # noqa: N802 - reference name
        """loadParam(otf_file) or loadParam(n_cam, nx, ny, nz, boundary,
        cam_list) (OTF.cpp:159-271)."""
        if len(args) == 1:
            self._param = C.load_otf(args[0])
        else:
            n_cam, nx, ny, nz, boundary, cams = args
            self._param = C.Otf.default(n_cam, (nx, ny, nz), tuple(boundary),
                                        [c.max_intensity for c in cams])

    def saveParam(self, otf_file: str) -> None:  # noqa: N802 - reference name
        C.save_otf(self._param, otf_file)

    def getOTFParam(self, cam_id: int, pt_world):  # noqa: N802 - reference name
        """(a, b, c, cos alpha) at pt_world - the fourth value, labelled alpha,
        is cos alpha as in pyOTF.cpp:49-53 (OTF::getOTFParam, OTF.cpp:
        273-338, FP64)."""
        return tuple(float(v) for v in self.getOTFParams(cam_id, [pt_world])[0][:4])

    def getOTFParams(self, cam_id: int, pts) -> np.ndarray:  # noqa: N802
        """Many points at once: [n][5] = (a, b, c, cos alpha, sin alpha)."""
        o = self._param
        cam = S.make_rig(64, 64, n_cam=o.n_cam)
        ctx = _context(cam, "f64", o, max_2d=1024, max_cand=1024, max_obj=1024)
        try:
            return ctx.otf_param(int(cam_id), np.asarray(pts, np.float64).reshape(-1, 3))
        finally:
            ctx.close()


class IPR:
    """IPR(cams) (IPR.h:30-57); keeps a copy of the cameras."""

    def __init__(self, cams, precision: str = "f32"):
        self._cam_list = _cams_copy(cams)
        self._precision = precision

    def runIPR(self, cfg, images) -> list:  # noqa: N802 - reference name
        """runIPR (IPR.cpp:195-306): every loop's objects in loop order, each
        with its final projections."""
        from .context import to_device_image
        d, code = to_device_image(_stack(images))
        ctx = _context(self._cam_list, self._precision, _otf_of(cfg))
        try:
            (ox, ouv, n_obj), _lt = ctx.ipr_run(d, code, cfg.ipr_params())
            n = int(n_obj[0].item())
            xyz = ox[0, :n].cpu().numpy()
            uv = ouv[0, :n].cpu().numpy()
        finally:
            ctx.close()
        r = float(cfg.radius_obj)
        return [Tracer3D(xyz[i].copy(), [Tracer2D(uv[i, k].copy(), r) for k in range(uv.shape[1])], r)
                for i in range(n)]

    def saveObjInfo(self, filename: str, objs, cfg) -> None:  # noqa: N802 - reference name
        """saveObjInfo (IPR.cpp:307-343)."""
        xyz = _xyz(objs)
        n_cam = len(self._cam_list)
        uv = np.array([[o._obj2d_list[k]._pt_center if o._obj2d_list[k] is not None
                        else (np.nan, np.nan)
                        for k in range(n_cam)] for o in objs], np.float64).reshape(-1, n_cam, 2)
        C.save_obj_info(filename, xyz, uv)
