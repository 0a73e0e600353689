#!/usr/bin/env python3
"""Write the reference's own IPR::saveObjInfo output (IPR.cpp:307-343,
Tracer3D::saveObject3D ObjectInfo.cpp:76-94) for a golden object list.

Input: golden_u8_256's runIPR objects (positions and projections), with a
few cameras' observations removed (NaN) and a few values chosen for the
stream format: large, tiny, negative and exactly integral numbers.  Stored:
objinfo_u8_256.npz (xyz, uv) and objinfo_u8_256.csv (the file the reference
wrote).

Usage: python tests/golden/make_golden_objinfo.py   (needs `make -C oracle ref`)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from common import Case  # noqa: E402


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libolpt_ref.so is missing: run `make -C oracle ref`")
    case = Case("golden_u8_256")
    xyz = np.ascontiguousarray(case.z["ipr_xyz"][:200])
    uv = np.ascontiguousarray(case.z["ipr_uv"][:200])
    uv[3, 1] = np.nan
    uv[7, 0] = np.nan
    uv[7, 3] = np.nan
    xyz[10] = (123456789.0, -1e-5, 0.0)
    xyz[11] = (-0.0, 1e300, -7.0)
    uv[12, 2] = (1.5e-7, 999999.5)
    path = os.path.join(HERE, "objinfo_u8_256.csv")
    rc = O.load_ref().ref_save_obj_info(O.cams_array(case.cams), case.n_cam, O.dptr(xyz),
                                        O.dptr(uv), len(xyz), path.encode())
    assert rc == 0
    np.savez_compressed(os.path.join(HERE, "objinfo_u8_256.npz"), xyz=xyz, uv=uv)
    print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
