#!/usr/bin/env python3
"""Write golden vectors for StereoMatch::pruneMatch alone, from the
reference's own code (oracle/_ref, ref_prune in ref_harness.cpp).

Cases, each a candidate set as match() hands it over (distinct id vectors in
lexicographic order, StereoMatch.cpp:153-156) with e_check values:

  random   3,000 candidates over 4 cameras of ~300 points, about 8 per point;
           e_check uniform, with 10 % exact repeats and 5 % one-ulp neighbours
           (eq_eps-equal and transitive: the tie blocks of :742-790)
  dense    1,500 candidates over 60 points per camera (~25 per point), e
           drawn from 40 values, so percentile and e ties are everywhere
  chain    the random set plus three candidates sharing one point whose
           e_check values form an eq_eps chain (a ~ b, b ~ c, a !~ c): the
           comparator is not transitive there, so the reference's order is
           whatever its std::stable_sort produces

Stored: ids, e, n_pts and the reference's selection mask per case.

Usage: python tests/golden/make_golden_prune.py   (needs `make -C oracle ref`)
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


def candidates(rng, n_pts, m, p_use=0.9):
    n_cam = len(n_pts)
    seen = set()
    rows = []
    while len(rows) < m:
        use = rng.random(n_cam) < p_use
        if use.sum() < 2:
            continue
        ids = tuple(int(rng.integers(0, n_pts[c])) if use[c] else -1 for c in range(n_cam))
        if ids in seen:
            continue
        seen.add(ids)
        rows.append(ids)
    rows.sort()
    return np.array(rows, np.int32)


def random_e(rng, m, choices=None):
    if choices is not None:
        return rng.choice(choices, size=m)
    e = rng.uniform(1e-3, 4e-2, m)
    rep = rng.random(m) < 0.10
    e[rep] = e[rng.integers(0, m, rep.sum())]
    nb = rng.random(m) < 0.05
    src = rng.integers(0, m, nb.sum())
    e[nb] = np.nextafter(e[src], np.inf)
    return e


def ref_select(cams, ids, e, n_pts, params):
    lib = O.load_ref()
    m = len(e)
    sel = np.zeros(max(m, 1), np.uint8)
    rc = lib.ref_prune(O.cams_array(cams), len(cams), None, O.iptr(np.asarray(n_pts, np.int32)),
                       ctypes.byref(params), O.iptr(np.ascontiguousarray(ids)),
                       O.dptr(np.ascontiguousarray(e)), m, O.u8ptr(sel))
    assert rc == 0
    return sel[:m].copy()


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libolpt_ref.so is missing: run `make -C oracle ref`")
    cams = S.make_rig(256, 256, max_intensity=255)
    params = S.default_params(40.0 / (0.95 * 256))
    out = {}
    rng = np.random.default_rng(2024)
    n_pts = [300, 320, 310, 305]
    ids = candidates(rng, n_pts, 3000)
    e = random_e(rng, len(ids))
    out["random"] = (ids, e, n_pts)
    n_pts_d = [60, 60, 60, 60]
    ids_d = candidates(rng, n_pts_d, 1500)
    e_d = random_e(rng, len(ids_d), choices=rng.uniform(1e-3, 4e-2, 40))
    out["dense"] = (ids_d, e_d, n_pts_d)
    # eq_eps chain on a point of its own (camera 0, point 300): e_a = 0.01,
    # e_b = e_a + 0.8e-11, e_c = e_a + 1.6e-11 with the tolerance
    # 1e-9 * 0.01 + 1e-12 = 1.1e-11; the candidate with the largest e comes
    # first in index order, so (index, e) ordering is cyclic
    n_pts_c = [301, 323, 313, 308]
    extra = np.array([[300, 320, 310, 305], [300, 321, 311, 306], [300, 322, 312, 307]], np.int32)
    e_extra = np.array([0.01 + 1.6e-11, 0.01 + 0.8e-11, 0.01])
    ids_c = np.concatenate([ids, extra])
    e_c = np.concatenate([e, e_extra])
    order = np.lexsort(ids_c.T[::-1])
    out["chain"] = (ids_c[order], e_c[order], n_pts_c)
    z = {}
    for name, (ci, ce, npts) in out.items():
        sel = ref_select(cams, ci, ce, npts, params)
        z[f"{name}_ids"] = ci
        z[f"{name}_e"] = ce
        z[f"{name}_npts"] = np.array(npts, np.int32)
        z[f"{name}_sel"] = sel
        print(f"{name}: {len(ce)} candidates, {int(sel.sum())} selected")
    path = os.path.join(HERE, "prune_cases.npz")
    np.savez_compressed(path, **z)
    print(f"-> {path} ({os.path.getsize(path) // 1024} KB)")


if __name__ == "__main__":
    main()
