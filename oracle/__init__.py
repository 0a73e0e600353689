"""CPU checkers for the IPR hot path - TEST INFRASTRUCTURE ONLY.

Two libraries, both loaded through ctypes:

* ``oracle/liboracle.so`` ("port"): the from-scratch C restatement of the
  reference algorithm in ``olpt_oracle.c`` (every function cites the
  reference file:line it restates).
* ``oracle/_ref/libolpt_ref.so`` ("reference"): the reference's own sources
  compiled by ``oracle/Makefile`` together with ``ref_harness.cpp``.  It is
  built where the reference checkout exists and is used to write golden
  vectors and to time the reference CPU path.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker: the product path
(``pyopenlpt_amd``) never calls into it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libolpt_ref.so")

_c_int_p = ctypes.POINTER(ctypes.c_int32)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_u8_p = ctypes.POINTER(ctypes.c_uint8)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_dbl_p)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_int_p)


def u8ptr(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_c_u8_p)


def cams_array(cams):
    from pyopenlpt_amd.synth import OlptCamera
    arr = (OlptCamera * len(cams))()
    for i, c in enumerate(cams):
        arr[i] = c.to_c() if hasattr(c, "to_c") else c
    return arr


_ref = None
_port = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def load_ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_SO)
    return _ref


def load_port():
    global _port
    if _port is None:
        _port = ctypes.CDLL(PORT_SO)
    return _port
