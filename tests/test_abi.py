"""The C-ABI library loads and exports every entry point include/olpt_c.h
declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "olpt_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(olpt_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("olpt_ctx_create", "olpt_find_peaks", "olpt_stereo_match", "olpt_shake",
              "olpt_residual", "olpt_ipr_run", "olpt_ipr_graph_capture"):
        assert s in syms


def test_library_exports_all_declared_symbols():
    from pyopenlpt_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libolpt_hip.so is not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_status_codes_mirror_reference_error_codes():
    src = open(os.path.join(ROOT, "include", "olpt_c.h")).read()
    for name, val in (("INVALID_ARGUMENT", 1), ("OUT_OF_RANGE", 2), ("UNSUPPORTED", 3),
                      ("IO", 4), ("NO_ENOUGH_DATA", 5)):
        assert re.search(rf"OLPT_E_{name}\s*=\s*{val}\b", src)


def test_last_error_is_a_string():
    from pyopenlpt_amd import _lib
    L = _lib.lib()
    assert isinstance(L.olpt_last_error(), bytes)
    # invalid arguments are reported, not thrown, and need no GPU
    assert L.olpt_ctx_reserve(None, 4, 10, 10, 0, 0, 0) == 1
    assert b"null" in L.olpt_last_error()
