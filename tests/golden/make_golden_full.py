#!/usr/bin/env python3
"""Write full-size golden vectors from the reference's own IPR::runIPR.

Runs oracle/_ref/libolpt_ref.so (the reference sources compiled by
oracle/Makefile, driven through ref_harness.cpp) in FP64 on the full-size
seeded workloads and stores what the GPU tests compare against:

  ref_c3_f0.npz   ppp 0.05 clustered frame 0 (1024^2 u16, 52,429 particles)
  ref_c3_f1.npz   the same workload, frame 1
  ref_c2.npz      1024^2 u16, 20,000 uniform particles, all cameras
  ref_c2_red1.npz the same with n_cam_reduced = 1 (12 IPR loops)
  ref_noise_limit.npz  3,000 particles on 4 x 512^2 u8 under uniform noise
                  (~50,000 peaks per camera) with n_obj2d_process_max =
                  3,000: limit2DObjectNumber's shuffle on its one-position-
                  per-draw branch

Each file holds the workload's name and seed, a SHA-256 of the rendered
images (the GPU test regenerates them from the seed and checks the hash
first), the per-loop cumulative object counts, the objects' positions
(float64 where `exact` is set, so that FP64-mode runs can be held to 1e-9 mm;
float32 otherwise) and the reference's reconstruction/ghost rates and RMSE
against ground truth with find_ghosts semantics at a 1-voxel tolerance.

Usage: python tests/golden/make_golden_full.py [name ...]   (needs the
reference checkout built by `make -C oracle ref`; no names = all)
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from pyopenlpt_amd import synth as S  # noqa: E402

WANT = set(sys.argv[1:])


def image_hash(images: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(images).tobytes()).hexdigest()


def run(name: str, w: S.Workload, *, exact: bool, seed: int, frame: int = -1):
    if WANT and name not in WANT:
        return
    lib = O.load_ref()
    n_cam = len(w.cams)
    H, W = w.images.shape[1:]
    otf, _keep = S.make_otf(n_cam, w.otf_abca)
    imgs = w.images.astype(np.float64).copy()
    cap = 1 << 19
    ox = np.zeros((cap, 3))
    ou = np.zeros((cap, n_cam, 2))
    lt = np.zeros(256, np.int32)
    cnt = np.zeros(2, np.int32)
    t = time.perf_counter()
    rc = lib.ref_run_ipr(O.cams_array(w.cams), n_cam, O.dptr(imgs), H, W,
                         ctypes.byref(w.params), ctypes.byref(otf), cap, O.dptr(ox),
                         O.dptr(ou), 256, O.iptr(lt), O.iptr(cnt))
    el = time.perf_counter() - t
    assert rc == 0, rc
    n = int(cnt[0])
    xyz = ox[:n].copy()
    q = S.ghost_eval(xyz, w.truth, w.voxel_mm)
    q = {k: (np.nan if v is None else v) for k, v in q.items()}
    out = dict(workload=np.array(w.name), seed=np.int64(seed), frame=np.int64(frame),
               image_sha256=np.array(image_hash(w.images)),
               n_cam_reduced=np.int32(w.params.n_cam_reduced),
               n_loop_ipr_reduced=np.int32(w.params.n_loop_ipr_reduced),
               loops=lt[: cnt[1]].copy(),
               xyz=xyz if exact else xyz.astype(np.float32),
               recon_rate=np.float64(q["recon_rate"]), ghost_rate=np.float64(q["ghost_rate"]),
               rmse_mm=np.float64(q["rmse_mm"]), voxel_mm=np.float64(w.voxel_mm))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {n} objects, loops {out['loops'].tolist()}, recon {q['recon_rate']:.4f} "
          f"ghost {q['ghost_rate']:.4f} rmse {q['rmse_mm']:.5f} mm ({el:.1f} s) -> "
          f"{os.path.getsize(path) // 1024} KB", flush=True)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libolpt_ref.so is missing: run `make -C oracle ref`")
    for f in (0, 1):
        run(f"ref_c3_f{f}", S.workload_c3_frame(f), exact=(f == 0), seed=1234, frame=f)
    w = S.workload_c2()
    run("ref_c2", w, exact=True, seed=1234)
    w = S.workload_c2()
    w.params.n_cam_reduced = 1
    w.params.n_loop_ipr_reduced = 2
    run("ref_c2_red1", w, exact=False, seed=1234)
    run("ref_noise_limit", S.workload_noise_limit(), exact=True, seed=4242)


if __name__ == "__main__":
    main()
