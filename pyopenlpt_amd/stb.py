"""Shake-the-box tracking and the command-line run, over the C-ABI.

Mirrors the reference's Python surface for tracking (pySTB.cpp:86-230,
pyPredField.cpp, pyOpenLPT.cpp:83-93): ``STB(setting, obj_type,
obj_cfg_path)`` with ``process_frame``, ``saveTracksAll`` / ``loadTracksAll``
and the four track lists, ``PredField``, and ``run(config_path)``, the
frame loop of the command-line program (main.cpp:39-139).  Every frame's
computation runs in libolpt_hip.so (olpt_stb_*); only configuration
parsing and image decoding happen here.
"""
from __future__ import annotations

import ctypes
import dataclasses
import sys

import numpy as np
import torch

from . import _lib
from . import config as C
from . import synth as S
from ._lib import check
from .context import Context, ptr, to_device_image

SHORT_ACTIVE, LONG_ACTIVE, LONG_INACTIVE, EXIT = 0, 1, 2, 3


@dataclasses.dataclass
class Track:
    """Track (Track.h:19-47): frame ids and world positions."""
    _t_list: list
    xyz: np.ndarray

    @property
    def _obj3d_list(self):
        return [tuple(p) for p in self.xyz]


class STB:
    """STB(setting, obj_type, obj_cfg_path) (STB.cpp:6-38).  Only the
    "Tracer" object type is supported; the tracer configuration is read
    with its OTF estimated from the first five frames on the GPU.  Volume
    self-calibration is not part of this build (the cameras stay fixed)."""

    def __init__(self, setting: C.BasicSetting, obj_type: str, obj_cfg_path: str, *,
                 precision: str = "f32", device: int = 0, write_files: bool = True,
                 ctx: Context | None = None):
        if obj_type != "Tracer":
            raise _lib.FatalError(3, f"Unsupported object type:{obj_type}")
        self.setting = setting
        cams = setting.cam_list
        H, W = cams[0].n_row, cams[0].n_col
        self.ctx = ctx or Context(cams, H, W, device=device, precision=precision)
        self.cfg = C.TracerConfig()
        self.cfg.readConfig(obj_cfg_path, setting, ctx=self.ctx)
        otf, self._otf_keep = self.cfg.otf.to_c()
        self.ctx.set_otf(otf)
        self._ip = self.cfg.ipr_params()
        self._sp = self.cfg.stb_params(setting)
        h = ctypes.c_void_p()
        out = setting.output_path.encode() if write_files else b""
        check(_lib.lib().olpt_stb_create(self.ctx.h, ctypes.byref(self._sp), ctypes.byref(self._ip),
                                         out, ctypes.byref(h)))
        self.h = h

    def set_ipr_model(self, net, voxel_mm: float = 0.0, max_step_vox: float = 0.6,
                      n_shake_loop: int = 1, shake_width_vox: float = 0.03):
        """No counterpart in the reference: run every IPR call of the
        tracker with the learned shake proposal `net` (a learned.ShakeNet)
        followed by n_shake_loop plain iterations from shake_width_vox; the
        shake of the predicted particles keeps the configuration's plain
        schedule.  net None restores the plain IPR."""
        if net is None:
            self._ipr_net = None
            check(_lib.lib().olpt_stb_set_ipr_model(self.h, None, 0.0, 0.0, None))
            return
        vox = voxel_mm or self.setting.voxel_to_mm
        ip = S.OlptIprParams.from_buffer_copy(self._ip)
        ip.n_shake_loop = int(n_shake_loop)
        ip.shake_width = float(shake_width_vox) * vox
        self._ipr_net, self._ipr_params = net, ip
        check(_lib.lib().olpt_stb_set_ipr_model(self.h, net.h, float(vox), float(max_step_vox),
                                                ctypes.byref(ip)))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().olpt_stb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def process_frame(self, frame_id: int, img_list) -> None:
        """STB::processFrame (STB.cpp:40-89).  img_list: [n_cam][H][W]
        numpy (uint8/uint16/float) or a device tensor with its dtype code
        given as (tensor, code)."""
        if isinstance(img_list, tuple):
            d, code = img_list
        else:
            d, code = to_device_image(np.ascontiguousarray(np.stack(list(img_list))
                                                           if isinstance(img_list, list)
                                                           else img_list))
        torch.cuda.synchronize()
        check(_lib.lib().olpt_stb_process_frame(self.h, int(frame_id), ptr(d), code))

    def saveTracksAll(self, folder: str, t: int) -> None:  # noqa: N802 - reference name
        check(_lib.lib().olpt_stb_save_tracks(self.h, folder.encode(), int(t)))

    def loadTracksAll(self, folder: str, t: int) -> None:  # noqa: N802
        check(_lib.lib().olpt_stb_load_tracks(self.h, folder.encode(), int(t)))

    def rows(self, which: int) -> dict:
        """One list as arrays: track index, frame id, xyz per point."""
        L = _lib.lib()
        nt, npnt = ctypes.c_int32(), ctypes.c_int32()
        check(L.olpt_stb_track_count(self.h, which, ctypes.byref(nt), ctypes.byref(npnt)))
        n = npnt.value
        tid = np.zeros(max(1, n), np.int32)
        fr = np.zeros(max(1, n), np.int32)
        xyz = np.zeros((max(1, n), 3))
        check(L.olpt_stb_get_tracks(self.h, which, n, tid.ctypes.data_as(_lib._I32P),
                                    fr.ctypes.data_as(_lib._I32P), xyz.ctypes.data_as(_lib._DP)))
        return dict(n_tracks=nt.value, track=tid[:n], frame=fr[:n], xyz=xyz[:n])

    def _tracks(self, which: int) -> list[Track]:
        """Rows come track by track in list order (every track non-empty)."""
        r = self.rows(which)
        tid = r["track"]
        bounds = np.r_[0, np.flatnonzero(np.diff(tid)) + 1, len(tid)]
        return [Track(r["frame"][a:b].tolist(), r["xyz"][a:b])
                for a, b in zip(bounds[:-1], bounds[1:])]

    @property
    def _short_track_active(self):
        return self._tracks(SHORT_ACTIVE)

    @property
    def _long_track_active(self):
        return self._tracks(LONG_ACTIVE)

    @property
    def _long_track_inactive(self):
        return self._tracks(LONG_INACTIVE)

    @property
    def _exit_track(self):
        return self._tracks(EXIT)


def predict_next(ctx: Context, windows: np.ndarray, lens: np.ndarray):
    """STB::predictNext for tracks given by their last min(len, 6) points
    (windows [n][6][3], oldest first): (pred [n][3], ok [n])."""
    n = len(lens)
    d_pos = torch.from_numpy(np.ascontiguousarray(windows, np.float64)).cuda()
    d_len = torch.from_numpy(np.ascontiguousarray(lens, np.int32)).cuda()
    out = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    ok = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    check(_lib.lib().olpt_stb_predict(ctx.h, ptr(d_pos), ptr(d_len), n, ptr(out), ptr(ok)))
    return out.cpu().numpy(), ok.cpu().numpy()


def check_linear_fit(ctx: Context, n_initial_frames: int, windows: np.ndarray, lens: np.ndarray):
    """STB::checkLinearFit for tracks given by their last min(len, 4) points
    (windows [n][4][3]) and lengths: ok [n]."""
    n = len(lens)
    d_pos = torch.from_numpy(np.ascontiguousarray(windows, np.float64)).cuda()
    d_len = torch.from_numpy(np.ascontiguousarray(lens, np.int32)).cuda()
    ok = torch.zeros(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    check(_lib.lib().olpt_stb_linear_fit(ctx.h, int(n_initial_frames), ptr(d_pos), ptr(d_len), n,
                                         ptr(ok)))
    return ok.cpu().numpy()


class PredField:
    """PredField(cfg) + calPredField(prev, curr) + getDisp (PredField.cpp)."""

    def __init__(self, ctx: Context, stb_params: S.OlptStbParams):
        self.ctx = ctx
        self.sp = stb_params

    def calPredField(self, prev: np.ndarray, curr: np.ndarray, queries=None):  # noqa: N802
        sp = self.sp
        ng = sp.pf_nx * sp.pf_ny * sp.pf_nz
        dp = torch.from_numpy(np.ascontiguousarray(prev, np.float64).reshape(-1, 3)).cuda()
        dc = torch.from_numpy(np.ascontiguousarray(curr, np.float64).reshape(-1, 3)).cuda()
        q = np.zeros((0, 3)) if queries is None else np.ascontiguousarray(queries, np.float64)
        dq = torch.from_numpy(q.reshape(-1, 3)).cuda()
        field = torch.zeros((ng, 3), dtype=torch.float64, device="cuda")
        disp = torch.zeros((max(1, len(q)), 3), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        check(_lib.lib().olpt_pred_field(self.ctx.h, ctypes.byref(sp), ptr(dp), len(dp), ptr(dc),
                                         len(dc), ptr(field), ptr(dq), len(q), ptr(disp)))
        self.field = field.cpu().numpy()
        return self.field, disp[: len(q)].cpu().numpy()


def run(config_path: str, *, precision: str = "f32", device: int = 0, verbose: bool = True) -> int:
    """The command-line run (main.cpp:39-139, pyOpenLPT.cpp:83-93): read the
    configuration, build one STB per object type, load previous tracks if
    asked, then feed every frame's images to every STB.  Returns 0 on
    success, 1 on failure (EXIT_SUCCESS / EXIT_FAILURE)."""
    setting = C.BasicSetting()
    try:
        if not setting.readConfig(config_path):
            print(f"Error: Failed to read basic configuration from file: {config_path}",
                  file=sys.stderr)
            return 1
        stbs = [STB(setting, t, p, precision=precision, device=device)
                for t, p in zip(setting.object_types, setting.object_config_paths)]
        if setting.load_track:
            for s in stbs:
                s.loadTracksAll(setting.load_track_path, setting.load_track_frame)
        ios = [C.ImageIO("", p) for p in setting.image_file_paths]
        start = setting.load_track_frame + 1 if setting.load_track else setting.frame_start
        for f in range(start, setting.frame_end + 1):
            imgs = np.stack([io.loadImg(f) for io in ios])
            d = to_device_image(imgs)
            for s in stbs:
                s.process_frame(f, d)
            if verbose:
                c = [s.rows(LONG_ACTIVE)["n_tracks"] for s in stbs]
                print(f"frame {f}: active long tracks {c}", flush=True)
        for s in stbs:
            s.close()
    except (_lib.FatalError, ValueError, OSError) as e:
        print(f"Program aborted due to error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(run(sys.argv[1]))
