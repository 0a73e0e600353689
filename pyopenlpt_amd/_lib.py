"""ctypes binding of libolpt_hip.so (include/olpt_c.h).

The shared library is the product: if it is missing this module raises at
import time - there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

from .synth import OlptCamera, OlptIprParams, OlptOtf, OlptStbParams, OlptTuning

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libolpt_hip.so")

OLPT_OK = 0
OLPT_U8, OLPT_U16, OLPT_F32, OLPT_F64 = 0, 1, 2, 3
OLPT_PREC_F32, OLPT_PREC_F64 = 0, 1

# ErrorCode names of the reference (inc/error.hpp:41-48) + GPU codes
STATUS_NAMES = {
    1: "InvalidArgument", 2: "OutOfRange", 3: "UnsupportedType",
    4: "IOfailure", 5: "NoEnoughData", 100: "GpuError",
    101: "AllocationFailed", 102: "CapacityExceeded", 103: "PruneTiesNotReproduced",
}


class FatalError(RuntimeError):
    """Raised for a non-zero olpt_status (the reference raises FatalError
    carrying an ErrorCode, inc/error.hpp:98-126)."""

    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"[{STATUS_NAMES.get(status, status)}] {message}")


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_I32P = ctypes.POINTER(ctypes.c_int32)
_DP = ctypes.POINTER(ctypes.c_double)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    L.olpt_last_error.restype = ctypes.c_char_p
    L.olpt_ctx_create.argtypes = [_I, _P, ctypes.POINTER(_P)]
    L.olpt_ctx_destroy.argtypes = [_P]
    L.olpt_ctx_reserve.argtypes = [_P, _I, _I, _I, _I, _I, _I]
    L.olpt_set_precision.argtypes = [_P, _I]
    L.olpt_set_concurrency.argtypes = [_P, _I]
    L.olpt_set_tuning.argtypes = [_P, ctypes.POINTER(OlptTuning)]
    L.olpt_get_tuning.argtypes = [_P, ctypes.POINTER(OlptTuning)]
    L.olpt_set_cameras.argtypes = [_P, ctypes.POINTER(OlptCamera), _I]
    L.olpt_set_otf.argtypes = [_P, ctypes.POINTER(OlptOtf)]
    L.olpt_sync.argtypes = [_P]
    L.olpt_stream.argtypes = [_P]
    L.olpt_stream.restype = _P
    L.olpt_diagnostics.argtypes = [_P, _I32P]
    L.olpt_build_counts.argtypes = [_P, _I32P]
    L.olpt_set_ghost_logit.argtypes = [_P, ctypes.c_double]
    L.olpt_find_peaks.argtypes = [_P, _P, _I, _I, _I, ctypes.POINTER(OlptIprParams),
                                  _P, _I, _I32P]
    L.olpt_stereo_match.argtypes = [_P, _I32P, _P, _I, _I32P,
                                    ctypes.POINTER(OlptIprParams), _P, _P, _P, _I,
                                    _P, _P, _I, _I32P]
    L.olpt_otf_param.argtypes = [_P, _I, _P, _I, _P]
    L.olpt_find_peaks_static.argtypes = [_P, _P, _I, _I, _I, ctypes.POINTER(OlptIprParams),
                                         _P, _I, _P]
    L.olpt_prune_match.argtypes = [_P, _I32P, _I32P, ctypes.POINTER(OlptIprParams), _P, _P,
                                   _I, _P]
    L.olpt_shake.argtypes = [_P, _I32P, _P, _I, ctypes.POINTER(OlptIprParams), _P,
                             _I, _P, _P, _P, _P]
    L.olpt_residual.argtypes = [_P, _P, _I, ctypes.POINTER(OlptIprParams), _P, _I,
                                _P]
    L.olpt_ipr_run.argtypes = [_P, _P, _I, _I, ctypes.POINTER(OlptIprParams), _P,
                               _P, _I, _P, _I32P, _I]
    L.olpt_ipr_graph_capture.argtypes = [_P, _P, _I, _I,
                                         ctypes.POINTER(OlptIprParams), _P, _P, _I,
                                         _P]
    L.olpt_ipr_graph_launch.argtypes = [_P]
    L.olpt_stage_timing.argtypes = [_P, _I]
    L.olpt_stage_times.argtypes = [_P, _DP, _I32P]
    L.olpt_stage_stats.argtypes = [_P, _I, _DP, _I32P]
    _F = ctypes.c_float
    _FP = ctypes.POINTER(ctypes.c_float)
    L.olpt_mlp_num_params.restype = ctypes.c_int32
    L.olpt_mlp_last_error.restype = ctypes.c_char_p
    L.olpt_mlp_create.argtypes = [_I, _P, _I, ctypes.POINTER(_P)]
    L.olpt_mlp_destroy.argtypes = [_P]
    L.olpt_mlp_set_params.argtypes = [_P, _FP, ctypes.c_int32]
    L.olpt_mlp_get_params.argtypes = [_P, _FP, ctypes.c_int32]
    L.olpt_mlp_get_grads.argtypes = [_P, _FP, ctypes.c_int32]
    L.olpt_mlp_infer.argtypes = [_P, _P, ctypes.c_int32, _P]
    L.olpt_mlp_train_step.argtypes = [_P, _P, _P, ctypes.c_int32, _F, _F, _FP]
    L.olpt_set_shake_model.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double]
    L.olpt_render.argtypes = [_P, _P, _P, _I, ctypes.c_double, _P]
    L.olpt_shake_patches.argtypes = [_P, _I32P, _P, _I, ctypes.POINTER(OlptIprParams), _P, _I, _P]
    for name in ("olpt_set_shake_model", "olpt_render", "olpt_shake_patches"):
        getattr(L, name).restype = ctypes.c_int32
    _SP = ctypes.POINTER(OlptStbParams)
    _IP = ctypes.POINTER(OlptIprParams)
    L.olpt_stb_create.argtypes = [_P, _SP, _IP, ctypes.c_char_p, ctypes.POINTER(_P)]
    L.olpt_stb_destroy.argtypes = [_P]
    L.olpt_stb_set_ipr_model.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double, _IP]
    L.olpt_stb_process_frame.argtypes = [_P, _I, _P, _I]
    L.olpt_stb_track_count.argtypes = [_P, _I, _I32P, _I32P]
    L.olpt_stb_get_tracks.argtypes = [_P, _I, ctypes.c_int32, _I32P, _I32P, _DP]
    L.olpt_stb_set_tracks.argtypes = [_P, _I, ctypes.c_int32, _I32P, _I32P, _DP]
    L.olpt_stb_save_tracks.argtypes = [_P, ctypes.c_char_p, _I]
    L.olpt_stb_load_tracks.argtypes = [_P, ctypes.c_char_p, _I]
    L.olpt_stb_predict.argtypes = [_P, _P, _P, _I, _P, _P]
    L.olpt_stb_linear_fit.argtypes = [_P, _I, _P, _P, _I, _P]
    L.olpt_pred_field.argtypes = [_P, _SP, _P, _I, _P, _I, _P, _P, _I, _P]
    L.olpt_estimate_otf.argtypes = [_P, _P, _I, _I, _IP, _DP, _I32P]
    for name in ("olpt_stb_create", "olpt_stb_destroy", "olpt_stb_set_ipr_model",
                 "olpt_stb_process_frame",
                 "olpt_stb_track_count", "olpt_stb_get_tracks", "olpt_stb_set_tracks",
                 "olpt_stb_save_tracks", "olpt_stb_load_tracks", "olpt_stb_predict",
                 "olpt_stb_linear_fit", "olpt_pred_field", "olpt_estimate_otf"):
        getattr(L, name).restype = ctypes.c_int32
    for name in ("olpt_mlp_create", "olpt_mlp_destroy", "olpt_mlp_set_params",
                 "olpt_mlp_get_params", "olpt_mlp_get_grads", "olpt_mlp_infer",
                 "olpt_mlp_train_step"):
        getattr(L, name).restype = ctypes.c_int32
    for name in ("olpt_ctx_create", "olpt_ctx_destroy", "olpt_ctx_reserve",
                 "olpt_set_precision", "olpt_set_concurrency", "olpt_set_tuning",
                 "olpt_get_tuning", "olpt_set_cameras",
                 "olpt_set_otf",
                 "olpt_sync", "olpt_diagnostics", "olpt_build_counts", "olpt_set_ghost_logit", "olpt_find_peaks",
                 "olpt_stereo_match", "olpt_prune_match", "olpt_find_peaks_static", "olpt_otf_param", "olpt_shake", "olpt_residual",
                 "olpt_ipr_run", "olpt_ipr_graph_capture",
                 "olpt_ipr_graph_launch", "olpt_stage_timing",
                 "olpt_stage_times", "olpt_stage_stats"):
        getattr(L, name).restype = ctypes.c_int32
    _lib = L
    return L


def check(status: int) -> None:
    if status != OLPT_OK:
        msg = lib().olpt_last_error().decode(errors="replace")
        raise FatalError(status, msg)


def check_mlp(status: int) -> None:
    if status != OLPT_OK:
        msg = lib().olpt_mlp_last_error().decode(errors="replace")
        raise FatalError(status, msg)
