"""Device context over the C-ABI, using torch tensors for device memory.

torch is plumbing here: it allocates HBM and provides the stream; every
computation runs in libolpt_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .synth import OlptCamera, OlptIprParams, OlptOtf, OlptTuning

_DTYPE_CODE = {torch.uint8: _lib.OLPT_U8, torch.int16: _lib.OLPT_U16,
               torch.float32: _lib.OLPT_F32, torch.float64: _lib.OLPT_F64}


def dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.uint16 if hasattr(torch, "uint16") else False:
        return _lib.OLPT_U16
    try:
        return _DTYPE_CODE[t.dtype]
    except KeyError as e:
        raise _lib.FatalError(3, f"unsupported image dtype {t.dtype}") from e


def to_device_image(img: np.ndarray, device="cuda") -> tuple[torch.Tensor, int]:
    """numpy uint8/uint16/float32/float64 -> device tensor + dtype code.
    uint16 travels as int16 storage (same bits)."""
    if img.dtype == np.uint8:
        return torch.from_numpy(np.ascontiguousarray(img)).to(device), _lib.OLPT_U8
    if img.dtype == np.uint16:
        t = torch.from_numpy(np.ascontiguousarray(img).view(np.int16)).to(device)
        return t, _lib.OLPT_U16
    if img.dtype == np.float32:
        return torch.from_numpy(np.ascontiguousarray(img)).to(device), _lib.OLPT_F32
    if img.dtype == np.float64:
        return torch.from_numpy(np.ascontiguousarray(img)).to(device), _lib.OLPT_F64
    raise _lib.FatalError(3, f"unsupported image dtype {img.dtype}")


def ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """One olpt_ctx bound to torch's current stream on `device`."""

    def __init__(self, cams, n_row: int, n_col: int, *, device: int = 0,
                 precision: str = "f32", max_2d: int = 0, max_cand: int = 0,
                 max_obj: int = 0, stream: torch.cuda.Stream | None = None):
        self.L = _lib.lib()
        self.device = device
        self.stream = stream if stream is not None else torch.cuda.current_stream(device)
        h = ctypes.c_void_p()
        check(self.L.olpt_ctx_create(device, ctypes.c_void_p(self.stream.cuda_stream),
                                     ctypes.byref(h)))
        self.h = h
        self.n_cam = len(cams)
        self.H, self.W = n_row, n_col
        check(self.L.olpt_ctx_reserve(self.h, self.n_cam, n_row, n_col, max_2d,
                                      max_cand, max_obj))
        self.set_precision(precision)
        self.set_cameras(cams)

    def close(self):
        if getattr(self, "h", None):
            self.L.olpt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration ---------------------------------------------------
    def set_precision(self, precision: str):
        self.precision = precision
        check(self.L.olpt_set_precision(
            self.h, _lib.OLPT_PREC_F64 if precision == "f64" else _lib.OLPT_PREC_F32))

    def set_concurrency(self, n_streams: int):
        """Spread the frames of one ipr_run over n_streams workspaces, each
        on its own HIP stream (1 = every frame on the context stream)."""
        check(self.L.olpt_set_concurrency(self.h, int(n_streams)))

    def set_tuning(self, **kw):
        """Change launch shapes (olpt_set_tuning): any of n_streams,
        shake_blocks, build_blocks, prune_blocks; the rest stay."""
        t = self.get_tuning()
        for k, v in kw.items():
            if k not in t:
                raise KeyError(k)
            t[k] = int(v)
        check(self.L.olpt_set_tuning(self.h, ctypes.byref(OlptTuning(**t))))

    def get_tuning(self) -> dict:
        t = OlptTuning()
        check(self.L.olpt_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    TUNING_SPACE = {
        "n_streams": (1, 2, 4, 6, 8),
        "shake_blocks": (1024, 2048, 4096, 8192, 16384),
        "build_blocks": (1024, 2048, 4096, 8192),
        "prune_blocks": (128, 256, 512, 1024, 2048, 4096),
    }

    def autotune(self, imgs: torch.Tensor, dtype: int, params: OlptIprParams,
                 n_frames: int, space: dict | None = None, reps: int = 2,
                 passes: int = 1) -> tuple[dict, list]:
        """Time olpt_ipr_run on the caller's frames over launch shapes and
        keep the fastest (coordinate descent: one knob at a time, the others
        held at the best so far; each shape is timed once).  The shapes do not change results, only
        speed.  Returns (best tuning, [(tuning, seconds per call)])."""
        import time
        space = space or self.TUNING_SPACE
        outputs = self.alloc_outputs(n_frames, 1 << 18, imgs.device)
        log = []
        seen = {}

        def measure(t):
            key = tuple(sorted(t.items()))
            if key in seen:          # a shape is timed once
                return seen[key]
            self.set_tuning(**t)
            self.ipr_run(imgs, dtype, params, n_frames, outputs=outputs)
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                self.ipr_run(imgs, dtype, params, n_frames, outputs=outputs)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            log.append((dict(t), best))
            seen[key] = best
            return best

        best_t = self.get_tuning()
        best_s = measure(best_t)
        for _ in range(passes):
            for knob, values in space.items():
                for v in values:
                    if v == best_t[knob]:
                        continue
                    cand = dict(best_t, **{knob: v})
                    sec = measure(cand)
                    if sec < best_s:
                        best_t, best_s = cand, sec
        self.set_tuning(**best_t)
        del outputs
        return best_t, log

    @property
    def real_dtype(self):
        return torch.float64 if self.precision == "f64" else torch.float32

    def set_cameras(self, cams):
        arr = (OlptCamera * len(cams))()
        for i, c in enumerate(cams):
            arr[i] = c.to_c() if hasattr(c, "to_c") else c
        self._cams = arr
        check(self.L.olpt_set_cameras(self.h, arr, len(cams)))

    def set_otf(self, otf: OlptOtf):
        check(self.L.olpt_set_otf(self.h, ctypes.byref(otf)))

    def sync(self):
        check(self.L.olpt_sync(self.h))

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(8, np.int32)
        check(self.L.olpt_diagnostics(self.h, out.ctypes.data_as(_lib._I32P)))
        return out

    STAGES = ("peaks", "stereo_build", "prune", "shake", "residual", "enum2", "host_issue")

    def build_counts(self) -> dict:
        """Work-list sizes of the last stereo build (frame 0 of the first
        workspace): pairs, survivors and items per level, candidates,
        selected candidates, objects."""
        v = np.zeros(24, np.int32)
        check(self.L.olpt_build_counts(self.h, v.ctypes.data_as(_lib._I32P)))
        return dict(pairs=int(v[0]), surv=v[1:9].tolist(), items=v[9:17].tolist(),
                    cand=int(v[17]), sel=int(v[18]), obj=int(v[19]))

    def stage_timing(self, enable: bool = True):
        check(self.L.olpt_stage_timing(self.h, 1 if enable else 0))

    def stage_times(self) -> dict:
        ms = np.zeros(8)
        calls = np.zeros(8, np.int32)
        check(self.L.olpt_stage_times(self.h, ms.ctypes.data_as(_lib._DP),
                                      calls.ctypes.data_as(_lib._I32P)))
        return {name: (float(ms[i]), int(calls[i])) for i, name in enumerate(self.STAGES)}

    def stage_stats(self, enable: bool) -> dict:
        b = np.zeros(8)
        n = np.zeros(8, np.int32)
        check(self.L.olpt_stage_stats(self.h, 1 if enable else 0,
                                      b.ctypes.data_as(_lib._DP),
                                      n.ctypes.data_as(_lib._I32P)))
        return {name: (float(b[i]), int(n[i])) for i, name in enumerate(self.STAGES)}

    def graph_capture(self, imgs: torch.Tensor, dtype: int, params: OlptIprParams,
                      n_frames: int, outputs, cap_obj: int):
        ox, ouv, n_obj = outputs
        check(self.L.olpt_ipr_graph_capture(self.h, ptr(imgs), dtype, n_frames,
                                            ctypes.byref(params), ptr(ox), ptr(ouv),
                                            cap_obj, ptr(n_obj)))

    def graph_launch(self):
        check(self.L.olpt_ipr_graph_launch(self.h))

    # -- stages ----------------------------------------------------------
    def find_peaks(self, img: torch.Tensor, dtype: int, params: OlptIprParams,
                   cap: int = 1 << 18) -> np.ndarray:
        xy = torch.empty((cap, 2), dtype=torch.float64, device=img.device)
        n = ctypes.c_int32()
        check(self.L.olpt_find_peaks(self.h, ptr(img), dtype, self.H, self.W,
                                     ctypes.byref(params), ptr(xy), cap,
                                     ctypes.byref(n)))
        return xy[: n.value].cpu().numpy()

    def stereo_match(self, pts: list[np.ndarray], params: OlptIprParams,
                     active=None, cap_cand: int = 1 << 21, cap_obj: int = 1 << 18,
                     candidates: bool = True):
        cap2d = max(1, max(len(p) for p in pts))
        xy = np.zeros((self.n_cam, cap2d, 2))
        for k, p in enumerate(pts):
            xy[k, : len(p)] = p
        d_xy = torch.from_numpy(xy).cuda(self.device)
        npts = np.array([len(p) for p in pts], np.int32)
        act = None if active is None else np.asarray(active, np.int32)
        dev = d_xy.device
        if candidates:
            cid = torch.empty((cap_cand, self.n_cam), dtype=torch.int32, device=dev)
            ce = torch.empty(cap_cand, dtype=torch.float64, device=dev)
            sel = torch.empty(cap_cand, dtype=torch.uint8, device=dev)
        else:
            cid = ce = sel = None
        ox = torch.empty((cap_obj, 3), dtype=torch.float64, device=dev)
        oi = torch.empty((cap_obj, self.n_cam), dtype=torch.int32, device=dev)
        counts = np.zeros(3, np.int32)
        check(self.L.olpt_stereo_match(
            self.h, None if act is None else act.ctypes.data_as(_lib._I32P),
            ptr(d_xy), cap2d, npts.ctypes.data_as(_lib._I32P), ctypes.byref(params),
            ptr(cid), ptr(ce), ptr(sel), cap_cand, ptr(ox), ptr(oi), cap_obj,
            counts.ctypes.data_as(_lib._I32P)))
        M, _, no = (int(x) for x in counts)
        if not candidates:
            return dict(obj_xyz=ox[:no].cpu().numpy(), obj_ids=oi[:no].cpu().numpy(),
                        counts=counts)
        return dict(cand_ids=cid[:M].cpu().numpy(), cand_e=ce[:M].cpu().numpy(),
                    sel=sel[:M].cpu().numpy().astype(bool), obj_xyz=ox[:no].cpu().numpy(),
                    obj_ids=oi[:no].cpu().numpy(), counts=counts)

    def otf_param(self, cam: int, xyz: np.ndarray) -> np.ndarray:
        """OTF::getOTFParam at points xyz [n][3]: [n][5] = (a, b, c, cos
        alpha, sin alpha), FP64."""
        dev = torch.device("cuda", self.device)
        x = torch.from_numpy(np.ascontiguousarray(np.atleast_2d(xyz), dtype=np.float64)).to(dev)
        out = torch.empty((max(len(x), 1), 5), dtype=torch.float64, device=dev)
        check(self.L.olpt_otf_param(self.h, int(cam), ptr(x), len(x), ptr(out)))
        return out[: len(x)].cpu().numpy()

    def prune_match(self, cand_ids: np.ndarray, cand_e: np.ndarray, n_pts, params: OlptIprParams,
                    active=None) -> np.ndarray:
        """StereoMatch::pruneMatch on given candidates ([n][n_cam] ids, -1 =
        unused; e_check [n]); returns the selection mask [n]."""
        n = len(cand_e)
        dev = torch.device("cuda", self.device)
        ids = torch.from_numpy(np.ascontiguousarray(cand_ids, dtype=np.int32)).to(dev)
        e = torch.from_numpy(np.ascontiguousarray(cand_e, dtype=np.float64)).to(dev)
        sel = torch.zeros(max(n, 1), dtype=torch.uint8, device=dev)
        npts = np.ascontiguousarray(n_pts, dtype=np.int32)
        act = None if active is None else np.asarray(active, np.int32)
        check(self.L.olpt_prune_match(self.h, None if act is None else act.ctypes.data_as(_lib._I32P),
                                      npts.ctypes.data_as(_lib._I32P), ctypes.byref(params),
                                      ptr(ids), ptr(e), n, ptr(sel)))
        return sel[:n].cpu().numpy().astype(bool)

    def shake(self, imgs: torch.Tensor, dtype: int, params: OlptIprParams,
              xyz: np.ndarray, active=None, want_residual: bool = True):
        n = len(xyz)
        dev = imgs.device
        d_in = torch.from_numpy(np.ascontiguousarray(xyz, dtype=np.float64)).to(dev)
        d_out = torch.empty((max(n, 1), 3), dtype=torch.float64, device=dev)
        d_fl = torch.empty(max(n, 1), dtype=torch.uint8, device=dev)
        d_sc = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        d_res = (torch.empty((self.n_cam, self.H, self.W), dtype=self.real_dtype, device=dev)
                 if want_residual else None)
        act = None if active is None else np.asarray(active, np.int32)
        check(self.L.olpt_shake(self.h, None if act is None else act.ctypes.data_as(_lib._I32P),
                                ptr(imgs), dtype, ctypes.byref(params), ptr(d_in), n,
                                ptr(d_out), ptr(d_fl), ptr(d_sc), ptr(d_res)))
        return (d_out[:n].cpu().numpy(), d_fl[:n].cpu().numpy(), d_sc[:n].cpu().numpy(),
                None if d_res is None else d_res.cpu().numpy())

    def residual(self, imgs: torch.Tensor, dtype: int, params: OlptIprParams,
                 xyz: np.ndarray) -> np.ndarray:
        dev = imgs.device
        d_in = torch.from_numpy(np.ascontiguousarray(xyz, dtype=np.float64)).to(dev)
        d_res = torch.empty((self.n_cam, self.H, self.W), dtype=self.real_dtype, device=dev)
        check(self.L.olpt_residual(self.h, ptr(imgs), dtype, ctypes.byref(params),
                                   ptr(d_in), len(xyz), ptr(d_res)))
        return d_res.cpu().numpy()

    # -- learned shake proposal ----------------------------------------------
    def set_shake_model(self, net, voxel_mm: float = 0.0, max_step_vox: float = 0.5,
                        ghost_logit: float | None = None):
        """Use a learned.ShakeNet before the plain shake iterations (None:
        plain shake only).  ghost_logit: objects whose ghost logit exceeds
        it are flagged ghosts after the shake (None: off)."""
        self._net = net
        check(self.L.olpt_set_shake_model(self.h, net.h if net is not None else None,
                                          float(voxel_mm), float(max_step_vox)))
        check(self.L.olpt_set_ghost_logit(
            self.h, float("inf") if ghost_logit is None else float(ghost_logit)))

    def render(self, xyz: torch.Tensor, amp: torch.Tensor | None = None,
               r_px: float = 2.5) -> torch.Tensor:
        """Max-fusion rendering of particles xyz [n][3] (float64, device)."""
        out = torch.empty((self.n_cam, self.H, self.W), dtype=self.real_dtype, device=xyz.device)
        check(self.L.olpt_render(self.h, ptr(xyz), ptr(amp), xyz.shape[0], float(r_px), ptr(out)))
        return out

    def shake_patches(self, imgs: torch.Tensor, dtype: int, params: OlptIprParams,
                      xyz: torch.Tensor, active=None) -> torch.Tensor:
        """Network inputs [n][512] (bfloat16) of objects xyz [n][3] (device)."""
        n = xyz.shape[0]
        out = torch.zeros((max(n, 1), 512), dtype=torch.bfloat16, device=xyz.device)
        act = None if active is None else np.asarray(active, np.int32)
        check(self.L.olpt_shake_patches(self.h, None if act is None else act.ctypes.data_as(_lib._I32P),
                                        ptr(imgs), dtype, ctypes.byref(params), ptr(xyz), n,
                                        ptr(out)))
        return out[:n]

    def ipr_run(self, imgs: torch.Tensor, dtype: int, params: OlptIprParams,
                n_frames: int = 1, cap_obj: int = 1 << 18, cap_loop: int = 64,
                outputs=None):
        dev = imgs.device
        if outputs is None:
            outputs = self.alloc_outputs(n_frames, cap_obj, dev)
        ox, ouv, n_obj = outputs
        lt = np.zeros((n_frames, cap_loop), np.int32)
        check(self.L.olpt_ipr_run(self.h, ptr(imgs), dtype, n_frames, ctypes.byref(params),
                                  ptr(ox), ptr(ouv), cap_obj, ptr(n_obj),
                                  lt.ctypes.data_as(_lib._I32P), cap_loop))
        return outputs, lt

    def alloc_outputs(self, n_frames, cap_obj, dev):
        ox = torch.empty((n_frames, cap_obj, 3), dtype=torch.float64, device=dev)
        ouv = torch.empty((n_frames, cap_obj, self.n_cam, 2), dtype=torch.float64, device=dev)
        n_obj = torch.zeros(n_frames, dtype=torch.int32, device=dev)
        return ox, ouv, n_obj
