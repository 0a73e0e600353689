#!/usr/bin/env python3
"""IPR throughput benchmark on one MI355X.

Workload: the ppp = 0.05 configuration - 4 pinhole cameras, 1024 x 1024
uint16 images, 52,429 particles per frame drawn from 8 Gaussian clusters
(sigma 5 mm) in a 40 mm cube, OTF a ~ U(800, 1200), b = c = 1.2, noise
sigma 10; 4 IPR loops.  A "step" runs IPR on a batch of 64 frames resident
in HBM: every kernel launch covers 16 frames (the frame is blockIdx.z) and
the four 16-frame batches run on 4 HIP streams; the step is replayed from a
captured hipGraph.  Legs: plain shake (the reference schedule, 6 iterations
from 0.25 voxel) in FP32 and in FP64, the plain step issued one frame per
call on one stream (how a drop-in runIPR replacement is driven), and the
learned proposal (the BF16 MLP trained here on separately seeded frames,
then 1 plain iteration from 0.03 voxel).  value = reconstructed particles
per second of the learned leg when its 3D RMSE against ground truth is not
above the plain leg's, else of the plain leg; every leg is reported in
config.

One JSON line is printed on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "IPR particles/s at ppp=0.05, learned vs plain shake at equal 3D RMSE"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _frame(args):
    from pyopenlpt_amd import synth as S
    f, ppp, size = args
    return S.workload_c3_frame(f, ppp=ppp, size=size)


def make_frames(n_frames: int, ppp: float, size: int):
    """Seeded frames (seed 1234 + f), rendered in a pool of spawned workers
    (joined before returning)."""
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(16, os.cpu_count() or 1, n_frames))
    if workers == 1:
        return [_frame((f, ppp, size)) for f in range(n_frames)]
    ex = cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))
    try:
        return list(ex.map(_frame, [(f, ppp, size) for f in range(n_frames)]))
    finally:
        ex.shutdown(wait=True)


def quality(xyz: np.ndarray, truth: np.ndarray, vox: float) -> dict:
    """find_ghosts semantics (demo/openLPT_utils/ghost_eval.py:72-92) at a
    1-voxel tolerance."""
    from pyopenlpt_amd import synth as S
    return S.ghost_eval(xyz, truth, vox)


CPU_CHILD = r"""
import ctypes, json, os, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
import oracle as O
from pyopenlpt_amd import synth as S
ppp, size, kind, reps = float(sys.argv[2]), int(sys.argv[3]), sys.argv[4], int(sys.argv[5])
w = S.workload_c3_frame(0, ppp=ppp, size=size)
H, W = w.images.shape[1:]
cams = O.cams_array(w.cams)
otf, keep = S.make_otf(4, w.otf_abca)
imgs = w.images.astype(np.float64).copy()
cap = 1 << 19
ox = np.zeros((cap, 3)); ou = np.zeros((cap, 4, 2)); lt = np.zeros(64, np.int32)
cnt = np.zeros(2, np.int32)
if kind == "reference":
    lib = O.load_ref(); fn = lib.ref_run_ipr
else:
    lib = O.load_port(); fn = lib.orc_run_ipr
times = []
for r in range(1 + reps):                      # 1 warm-up, then reps timed runs
    t = time.perf_counter()
    rc = fn(cams, 4, O.dptr(imgs), H, W, ctypes.byref(w.params), ctypes.byref(otf), cap,
            O.dptr(ox), O.dptr(ou), 64, O.iptr(lt), O.iptr(cnt))
    if r:
        times.append(time.perf_counter() - t)
    assert rc == 0
print(json.dumps(dict(n=int(cnt[0]), times=times)))
"""


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def cpu_baseline(ppp: float, size: int, threads: int, reps: int = 3) -> dict | None:
    """The reference's IPR::runIPR (or the C restatement when the reference
    build is absent) on frame 0 of the workload with OpenMP on `threads`
    host threads: 1 warm-up, then the median of `reps` runs (wall time
    around runIPR only)."""
    import oracle as O
    import statistics
    kind = "reference" if O.ref_available() else "port"
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    log(f"cpu baseline: {kind} IPR, frame 0, {threads} threads, 1 warm-up + {reps} runs")
    r = subprocess.run([sys.executable, "-c", CPU_CHILD, ROOT, str(ppp), str(size), kind, str(reps)],
                       capture_output=True, text=True, env=env, timeout=1800)
    if r.returncode != 0:
        log("cpu baseline failed:", r.stderr[-2000:])
        return None
    res = json.loads(r.stdout.strip().splitlines()[-1])
    med = statistics.median(res["times"])
    return {"value": round(res["n"] / med, 1), "unit": "particles/s", "cores": threads,
            "kind": kind,
            "sample": (f"frame 0 of the same workload ({res['n']} objects), IPR::runIPR wall time, "
                       f"median of {reps} after 1 warm-up: {med:.2f} s (runs "
                       f"{', '.join(f'{t:.2f}' for t in res['times'])} s)"),
            "host": {"nproc": os.cpu_count(), "cpu_model": cpu_model(),
                     "omp_threads": threads}}


def traffic_per_launch(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from profiles/<kernel>__<workload>.pmc.md
    (written by `python -m pyopenlpt_amd.prof pmc`), or None."""
    name = kernel.replace("<", "").replace(">", "")
    path = os.path.join(ROOT, "profiles", f"{name}__{workload}.pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_dispatch:"):
            return float(line.split(":", 1)[1])
    return None


# ---------------------------------------------------------------------------
# tracking workload (--workload stb)
# ---------------------------------------------------------------------------
STB_TRACER = dict(radius_search_obj_vox=10.0, n_initial_frames=4, radius_search_track_vox=30.0,
                  shake_width_vox=0.25, pf_grid=(51, 51, 51), pf_r_vox=25.0, n_loop_ipr=4,
                  n_shake_loop=4, thred_ghost=0.1, tol_2d_px=1.0, tol_3d_vox=1.0,
                  n_cam_reduced=0, n_loop_ipr_reduced=2, min_obj_int=240, radius_obj=2.0)


def train_proposal(cams, H: int, W: int, otf, params, vox: float, n_particles: int,
                   steps: int):
    """The learned shake proposal, trained here on GPU-rendered frames
    (seeds 9001..; ppp 0.01-0.08, a ~ U(800, 1200), noise 10) matched by
    this library's own peaks and stereo.  Returns (net, stats)."""
    import torch
    from pyopenlpt_amd.context import Context
    from pyopenlpt_amd.learned import ShakeNet, make_training_set, train
    t0 = time.perf_counter()
    tctx = Context(cams, H, W, precision="f32")
    tctx.set_otf(otf)
    X, T, tstats = make_training_set(tctx, params, vox, n_particles=n_particles, seed=9001)
    net = ShakeNet(max_batch=8192, seed=0)
    losses = train(net, X, T, steps=steps, batch=8192, lr=3e-3, ghost_weight=0.5)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del X, T
    tctx.close()
    log(f"learned proposal: {tstats['samples']} samples from {tstats['frames']} frames, "
        f"{steps} steps, loss {np.mean(losses[:10]):.3f} -> {np.mean(losses[-50:]):.3f}, "
        f"{t_train:.1f} s")
    return net, dict(samples=tstats["samples"], frames=tstats["frames"],
                     skipped_frames=tstats["skipped_frames"], particles=tstats["particles"],
                     steps=steps, seconds=round(t_train, 2), seeds="9001..")


def _stb_frame(args):
    from pyopenlpt_amd import synth as S
    cams, p, amp, seed = args
    return S.render(cams, p, amp, 1.2, 1.2, 0.0, 10.0, seed, np.uint16)


def make_stb_frames(n_part: int, n_frames: int, size: int, seed: int = 1234):
    """ABC-flow sequence (A = B = C = 1, period 40 mm, RK4, at most 3 px per
    frame, particles leaving the cube re-seeded), rendered as the C2 frames
    (u16, a ~ U(800, 1200) per particle, b = c = 1.2, noise 10)."""
    import concurrent.futures as cf
    from pyopenlpt_amd import synth as S
    vox = 40.0 / (0.95 * size)
    cams = S.make_rig(size, size, max_intensity=65535)
    seq = S.abc_sequence(n_part, n_frames, seed, max_disp_mm=3.0 * vox)
    rng = np.random.default_rng(seed + 5)
    amp = rng.uniform(800.0, 1200.0, max(int(ids.max()) for _, ids in seq) + 1)
    jobs = [(cams, p, amp[ids], seed + 100 + f) for f, (p, ids) in enumerate(seq)]
    workers = max(1, min(16, os.cpu_count() or 1, n_frames))
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        frames = np.stack(list(ex.map(_stb_frame, jobs)))
    return vox, cams, seq, frames


STB_CPU_CHILD = r"""
import ctypes, json, sys, time
sys.path.insert(0, sys.argv[1])
import oracle as O
L = O.load_ref()
L.ref_stb_open.restype = ctypes.c_void_p
L.ref_stb_open.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]
L.ref_stb_process.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
h = L.ref_stb_open(sys.argv[2].encode(), None, 1)
assert h
k = int(sys.argv[3])
t = time.perf_counter()
for f in range(k):
    assert L.ref_stb_process(ctypes.c_void_p(h), f, 1) == 0
print(json.dumps(dict(frames=k, seconds=time.perf_counter() - t)))
"""


def track_quality(stb, seq, vox: float) -> dict:
    """Tracks of at least 4 points (all four lists) against the ground
    truth: a point is matched to the nearest true particle of its frame
    within 1 voxel; a track is correct when all its points match the same
    particle.  Also the share of the last frame's particles carried by an
    active long track."""
    from scipy.spatial import cKDTree
    trees = [cKDTree(p) for p, _ in seq]
    n_tr = n_ok = n_pts = 0
    for which in range(4):
        r = stb.rows(which)
        if not len(r["track"]):
            continue
        order = np.argsort(r["track"], kind="stable")
        tid, fr, xyz = r["track"][order], r["frame"][order], r["xyz"][order]
        cuts = np.flatnonzero(np.diff(tid)) + 1
        for a, b in zip(np.r_[0, cuts], np.r_[cuts, len(tid)]):
            if b - a < 4:
                continue
            ids = []
            for f, q in zip(fr[a:b], xyz[a:b]):
                d, i = trees[f].query(q)
                ids.append(seq[f][1][i] if d < vox else -1)
            n_tr += 1
            n_pts += b - a
            n_ok += int(ids[0] >= 0 and all(v == ids[0] for v in ids))
    last = stb.rows(1)
    lf = len(seq) - 1
    sel = last["frame"] == lf
    d, i = trees[lf].query(last["xyz"][sel]) if sel.any() else (np.zeros(0), np.zeros(0, int))
    covered = len(np.unique(i[d < vox])) / len(seq[lf][0])
    return dict(tracks_min4=n_tr, correct_frac=round(n_ok / max(1, n_tr), 4),
                mean_len=round(n_pts / max(1, n_tr), 2), last_frame_coverage=round(covered, 4))


def stb_main(args) -> int:
    """Shake-the-box tracking on the ABC-flow sequence: the plain leg (every
    IPR call on the tracer configuration's schedule, 4 shake iterations from
    0.25 voxel) and the learned leg (every IPR call with the learned shake
    proposal trained here, then 1 plain iteration from 0.03 voxel; the shake
    of the predictions stays plain).  value = frames/s of the learned leg
    when its track purity (correct_frac) is not below the plain leg's, else
    of the plain leg."""
    import shutil
    import tempfile
    import torch
    from pyopenlpt_amd import config as C
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.context import to_device_image
    from pyopenlpt_amd.stb import STB
    n_frames = args.frames if args.frames_set else 100
    log(f"tracking workload: {args.stb_particles} particles, {n_frames} frames, {args.size}^2")
    vox, cams, seq, frames = make_stb_frames(args.stb_particles, n_frames, args.size)
    tmp = tempfile.mkdtemp(prefix="olpt_stb_")
    k_cpu = max(5, args.stb_cpu_frames)
    try:
        cfg = C.write_run_folder(tmp, cams, frames, axis_limit=(-20, 20, -20, 20, -20, 20),
                                 voxel_to_mm=vox, tracer=STB_TRACER,
                                 n_tiff=max(5, k_cpu if not args.no_cpu_baseline else 5))
        d_frames = [to_device_image(frames[f]) for f in range(n_frames)]
        torch.cuda.synchronize()
        setting = C.BasicSetting()
        setting.readConfig(cfg)

        def run_leg(net, label):
            stb = STB(setting, "Tracer", setting.object_config_paths[0],
                      precision=args.precision, write_files=False)
            stb.ctx.set_concurrency(1)
            if net is not None:
                stb.set_ipr_model(net, vox)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            t_k = None
            for f in range(n_frames):
                stb.process_frame(f, d_frames[f])
                if f + 1 == k_cpu:
                    torch.cuda.synchronize()
                    t_k = time.perf_counter() - t0
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            q = track_quality(stb, seq, vox)
            counts = {n: stb.rows(w)["n_tracks"] for w, n in enumerate(
                ("short_active", "long_active", "long_inactive", "exit"))}
            ip = stb._ip if net is None else stb._ipr_params
            cfg_params = S.OlptIprParams.from_buffer_copy(stb._ip)
            stb.close()
            log(f"{label}: {n_frames} frames in {wall:.2f} s -> {n_frames / wall:.1f} frames/s; "
                f"{q}")
            return dict(frames_per_s=round(n_frames / wall, 3),
                        ms_per_frame=round(1000 * wall / n_frames, 3),
                        first_frames_per_s=round(k_cpu / t_k, 3) if t_k else None,
                        ipr_shake_loops=int(ip.n_shake_loop),
                        ipr_shake_width_vox=round(ip.shake_width / vox, 4),
                        tracks=counts, quality=q, _params=cfg_params)

        plain = run_leg(None, "plain")
        params = plain.pop("_params")
        learned = None
        if args.shake == "both":
            otf, _keep = S.make_otf(4, (1000.0, 1.2, 1.2, 0.0))
            net, tstats = train_proposal(cams, args.size, args.size, otf, params, vox,
                                         args.train_particles, args.train_steps)
            learned = run_leg(net, "learned")
            learned.pop("_params")
            learned["train"] = tstats
            net.close()
        cpu = None
        if not args.no_cpu_baseline:
            import oracle as O
            if O.ref_available():
                env = dict(os.environ, OMP_NUM_THREADS=str(args.cpu_threads), OMP_PROC_BIND="close")
                log(f"cpu baseline: reference STB, frames 0..{k_cpu - 1}, {args.cpu_threads} threads")
                r = subprocess.run([sys.executable, "-c", STB_CPU_CHILD, ROOT, cfg, str(k_cpu)],
                                   capture_output=True, text=True, env=env, timeout=1800)
                if r.returncode == 0:
                    res = json.loads(r.stdout.strip().splitlines()[-1])
                    cpu = {"value": round(res["frames"] / res["seconds"], 4), "unit": "frames/s",
                           "cores": args.cpu_threads, "kind": "reference",
                           "sample": (f"frames 0-{k_cpu - 1} of the same sequence (4 initial-phase + "
                                      f"{k_cpu - 4} convergence frames), STB::processFrame wall "
                                      f"{res['seconds']:.2f} s"),
                           "gpu_same_frames_per_s": plain["first_frames_per_s"],
                           "host": {"nproc": os.cpu_count(), "cpu_model": cpu_model(),
                                    "omp_threads": args.cpu_threads}}
                else:
                    log("cpu baseline failed:", r.stderr[-2000:])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    use_learned = (learned is not None and
                   learned["quality"]["correct_frac"] >= plain["quality"]["correct_frac"])
    best = learned if use_learned else plain
    out = {"metric": "STB frames/s (tracking, C5 ABC flow)", "value": best["frames_per_s"],
           "unit": "frames/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": best["ms_per_frame"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
           "data": "synthetic",
           "config": {"workload": f"stb_abc_{args.stb_particles}p_{n_frames}f_{args.size}sq_u16_4cam",
                      "frames": n_frames, "particles": args.stb_particles,
                      "ipr_shake": "learned" if use_learned else "plain",
                      "tracks": best["tracks"], "quality": best["quality"],
                      "plain": plain, "learned": learned,
                      "learned_over_plain": (round(learned["frames_per_s"] / plain["frames_per_s"], 3)
                                             if learned else None),
                      "step": "one frame of STB::processFrame, images in HBM"},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ipr", choices=["ipr", "stb"],
                    help="ipr: the headline IPR throughput; stb: shake-the-box tracking")
    ap.add_argument("--stb-particles", type=int, default=30000)
    ap.add_argument("--stb-cpu-frames", type=int, default=6)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=64, help="frames per step")
    ap.add_argument("--ppp", type=float, default=0.05)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--eager", action="store_true",
                    help="launch the kernels each step instead of replaying a captured hipGraph")
    ap.add_argument("--streams", type=int, default=4,
                    help="batches of a step run concurrently on this many HIP streams")
    ap.add_argument("--batch", type=int, default=16,
                    help="frames covered by every kernel launch (frames_per_batch)")
    ap.add_argument("--no-fp64-leg", dest="fp64_leg", action="store_false",
                    help="skip the FP64 timing of the plain step")
    ap.add_argument("--no-per-frame-leg", dest="per_frame_leg", action="store_false",
                    help="skip the one-frame-per-call timing of the plain step")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of timing candidates "
                         "on the first frames (untimed) and keeping the fastest")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-steps", type=int, default=1,
                    help="single-stream steps timed per k_shake launch for the roofline")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the run, an elementwise kernel that reads and writes this many "
                         "MiB (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1),
                    help="OpenMP threads of the CPU baseline (default: OMP_NUM_THREADS, the "
                         "host's CPU allotment, else every core)")
    ap.add_argument("--shake", default="both", choices=["plain", "both"],
                    help="both: also train the learned proposal and time its leg")
    ap.add_argument("--train-particles", type=int, default=6_000_000)
    ap.add_argument("--train-steps", type=int, default=6000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); frames are "
              "independent, so N GPUs would each run their own batch.", file=sys.stderr)
        return 2
    args.frames_set = "--frames" in sys.argv
    if args.workload == "stb":
        return stb_main(args)

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.ppp, args.size, args.cpu_threads)

    import torch
    from pyopenlpt_amd import synth as S
    from pyopenlpt_amd.context import Context, to_device_image

    if not args.eager:
        # graph capture needs a created stream; make it torch's current one
        torch.cuda.set_stream(torch.cuda.Stream())

    workload = f"ipr_ppp{args.ppp}_{args.size}sq_u16_4cam_clustered"
    log(f"generating {args.frames} frames (ppp {args.ppp}, {args.size}^2)")
    frames = make_frames(args.frames, args.ppp, args.size)
    w0 = frames[0]
    H, W = w0.images.shape[1:]
    batch_imgs = np.stack([f.images for f in frames])     # [F][4][H][W] uint16
    dimg, code = to_device_image(batch_imgs)
    otf, _keep = S.make_otf(4, w0.otf_abca)
    # capacities for this workload: ~35k peaks per camera, ~490k stereo
    # candidates and ~21k objects per loop
    caps = dict(max_2d=1 << 16, max_cand=1 << 20, max_obj=1 << 17)
    ctx = Context(w0.cams, H, W, precision=args.precision, **caps)
    ctx.set_otf(otf)
    ctx.set_tuning(n_streams=args.streams, frames_per_batch=args.batch)
    cap_obj = 1 << 17
    outs = ctx.alloc_outputs(args.frames, cap_obj, dimg.device)
    stream = torch.cuda.current_stream()
    tune_log = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        space = {k: v for k, v in Context.TUNING_SPACE.items() if k != "n_streams"}
        best, tune_log = ctx.autotune(dimg, code, w0.params, args.frames, space=space, reps=1)
        for t, sec in tune_log:
            log(f"  {t}: {1000 * sec:.1f} ms")
        log(f"autotune over {len(tune_log)} shapes on {args.frames} frames "
            f"({time.perf_counter() - t0:.1f} s): {best}")
    tuning = ctx.get_tuning()

    def timed_leg(c, params, label, steps, graph):
        """Warm-up, then `steps` timed steps with events on the context
        stream; returns ms/step, particles/step and quality over the batch."""
        def step():
            if graph:
                c.graph_launch()
            else:
                c.ipr_run(dimg, code, params, n_frames=args.frames, cap_obj=cap_obj, outputs=outs)
        if graph:
            c.graph_capture(dimg, code, params, args.frames, outs, cap_obj)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        c.stage_times()      # reset the host-issue counter
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        h_ms, h_calls = c.stage_times()["host_issue"]
        ms = e0.elapsed_time(e1) / steps
        n_obj = outs[2].cpu().numpy()
        per_step = int(n_obj.sum())
        rate = per_step / (ms / 1000.0)
        log(f"{label}: {steps} steps: {ms:.2f} ms/step (wall {1000 * wall / steps:.2f}, "
            f"host issue {h_ms / max(1, h_calls):.1f}), {per_step} particles/step -> "
            f"{rate:.0f} particles/s")
        q = [quality(outs[0][f, : n_obj[f]].cpu().numpy(), frames[f].truth, frames[f].voxel_mm)
             for f in range(args.frames)]
        rmse = [x["rmse_mm"] for x in q if x["rmse_mm"] is not None]
        recon = float(np.mean([x["recon_rate"] for x in q]))
        qual = dict(rmse_mm=float(np.mean(rmse)) if rmse else None,
                    rmse_px=float(np.mean(rmse) / w0.voxel_mm) if rmse else None,
                    recon_rate=recon,
                    ghost_rate=float(np.mean([x["ghost_rate"] for x in q])))
        true_rate = recon * sum(len(f.truth) for f in frames) / (ms / 1000.0)
        return dict(particles_per_s=round(rate, 1), ms_per_step=round(ms, 3),
                    true_particles_per_s=round(true_rate, 1), graph=bool(graph),
                    host_issue_ms_per_step=round(h_ms / h_calls, 3) if h_calls else None,
                    particles_per_step=per_step, shake_loops=int(params.n_shake_loop),
                    quality=qual)

    def rooflines(c, params, label):
        """Live rooflines of k_shake and of k_b_enum at level 2 (the build's
        costliest kernel) on the leg `value` reports: one eager
        single-stream step with HIP events around every launch of each
        (concurrent streams would add waiting time to the intervals), and
        their algorithmic bytes per launch from an instrumented step.
        Returns (dominant, other), the dominant being the one with more
        time per step."""
        c.stage_stats(True)
        c.ipr_run(dimg, code, params, n_frames=args.frames, cap_obj=cap_obj, outputs=outs)
        stats = c.stage_stats(False)
        c.set_tuning(n_streams=1)
        c.stage_timing(True)
        c.stage_times()
        torch.cuda.synchronize()
        for _ in range(args.roofline_steps):
            c.ipr_run(dimg, code, params, n_frames=args.frames, cap_obj=cap_obj, outputs=outs)
        torch.cuda.synchronize()
        times = c.stage_times()
        c.stage_timing(False)
        c.set_tuning(n_streams=tuning["n_streams"])
        stage_ms = {k: round(v[0] / args.roofline_steps, 3) for k, v in times.items()
                    if v[1] and k != "host_issue"}
        out = []
        for stage, kernel, what in (
                ("shake", "k_shake",
                 "N x C x P^2 x s patch gathers (P = 8, s = 4 bytes) + 48 bytes per object, "
                 "summed over the launch's frames"),
                ("enum2", "k_b_enum<2>",
                 "per level-1 survivor its 16-byte path record read and per level-2 item its "
                 "16-byte record written, plus the line table once (each 2D point's 32-byte "
                 "lines on the other cameras), summed over the launch's frames")):
            if not (times.get(stage, (0, 0))[1] and stats.get(stage, (0, 0))[1]):
                continue
            ms_dom, calls = times[stage]
            bytes_dom, launches = stats[stage]
            per_launch = bytes_dom / launches
            avg_ms = ms_dom / calls
            achieved = per_launch / (avg_ms / 1000.0) / 1e9
            out.append({"bound": "hbm", "kernel": kernel, "leg": label,
                        "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 6),
                        "traffic": traffic_per_launch(kernel, workload),
                        "bytes_per_launch": int(per_launch), "algorithmic_bytes": what,
                        "avg_launch_ms": round(avg_ms, 4),
                        "ms_per_step_1stream": round(ms_dom / args.roofline_steps, 3),
                        "frames_per_launch": int(tuning["frames_per_batch"]),
                        "measured": f"{args.roofline_steps} eager single-stream step(s) of the "
                                    f"{label} leg, HIP events around each {kernel} launch",
                        "stage_ms_per_step_1stream": stage_ms})
        out.sort(key=lambda r: -r["ms_per_step_1stream"])
        return (out[0] if out else None), (out[1] if len(out) > 1 else None)

    graph = not args.eager
    # ---- plain shake: the reference schedule (6 iterations from 0.25 vox)
    plain = timed_leg(ctx, w0.params, "plain", args.steps, graph)

    # ---- the same step in FP64 (the reference's arithmetic): a
    # same-precision figure beside the CPU baseline
    plain_f64 = None
    if args.fp64_leg and args.precision == "f32":
        ctx.set_precision("f64")
        plain_f64 = timed_leg(ctx, w0.params, "plain fp64", max(1, args.steps // 2), graph)
        ctx.set_precision(args.precision)

    # ---- the drop-in pattern: one frame per call, one stream (what a
    # runIPR replacement does frame by frame), same kernels
    per_frame = None
    if args.per_frame_leg:
        ctx.set_tuning(n_streams=1, frames_per_batch=1)
        t0 = time.perf_counter()
        for f in range(args.frames):
            ctx.ipr_run(dimg[f:f + 1], code, w0.params, cap_obj=cap_obj)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        n_pf = 0
        for f in range(args.frames):
            (_x, _u, n1), _ = ctx.ipr_run(dimg[f:f + 1], code, w0.params, cap_obj=cap_obj)
            n_pf += int(n1[0].item())
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1)
        per_frame = dict(particles_per_s=round(n_pf / (ms / 1000.0), 1), ms_per_step=round(ms, 3),
                         particles_per_step=n_pf,
                         how="olpt_ipr_run once per frame, one stream, one frame per launch")
        log(f"plain per-frame: {ms:.1f} ms for {args.frames} frames -> "
            f"{per_frame['particles_per_s']:.0f} particles/s")
        ctx.set_tuning(n_streams=tuning["n_streams"], frames_per_batch=tuning["frames_per_batch"])

    # ---- learned proposal + 1 plain iteration (from 0.03 vox)
    learned = None
    lp = None
    if args.shake == "both":
        net, tstats = train_proposal(w0.cams, H, W, otf, w0.params, w0.voxel_mm,
                                     args.train_particles, args.train_steps)
        lp = S.OlptIprParams.from_buffer_copy(w0.params)
        lp.n_shake_loop = 1
        lp.shake_width = 0.03 * w0.voxel_mm
        ctx.set_shake_model(net, w0.voxel_mm, 0.6)
        learned = timed_leg(ctx, lp, "learned", args.steps, graph)
        learned["train"] = dict(tstats, seeds="9001.. (eval frames 1234..)")

    # the metric: learned vs plain at equal 3D RMSE (learned counts when its
    # RMSE is not above plain's)
    use_learned = (learned is not None and learned["quality"]["rmse_mm"] is not None and
                   plain["quality"]["rmse_mm"] is not None and
                   learned["quality"]["rmse_mm"] <= plain["quality"]["rmse_mm"])
    best = learned if use_learned else plain
    value = best["particles_per_s"]
    ms_per_step = best["ms_per_step"]

    # Roofline of k_shake on the leg `value` reports.  Algorithmic bytes per
    # launch: the per-object ROI patch gathers N x C x P^2 x s (P = 8, s = 4
    # bytes in FP32) plus positions, over the launch's frames; HBM traffic per
    # launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    # this workload, when present.
    roofline = roofline_other = None
    if args.roofline_steps > 0:
        if use_learned:
            roofline, roofline_other = rooflines(ctx, lp, "learned")
        else:
            ctx.set_shake_model(None)
            roofline, roofline_other = rooflines(ctx, w0.params, "plain")
    ctx.set_shake_model(None)

    if args.copy_calib_mib > 0:
        nel = args.copy_calib_mib * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device=dimg.device)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log(f"calibration: 3 x neg over {args.copy_calib_mib} MiB")

    ratio = (lambda a, b: round(a["particles_per_s"] / b["particles_per_s"], 3)
             if a and b else None)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particles/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": workload,
                   "frames_per_step": args.frames, "particles_per_frame": len(w0.truth),
                   "ipr_loops": int(w0.params.n_loop_ipr),
                   "shake": "learned" if use_learned else "plain",
                   "graph": graph, "streams": tuning["n_streams"],
                   "frames_per_launch": tuning["frames_per_batch"],
                   "tuning": tuning, "autotuned": not args.no_autotune, "capacities": caps,
                   "plain": plain, "plain_fp64": plain_f64, "plain_per_frame": per_frame,
                   "learned": learned,
                   "learned_over_plain": ratio(learned, plain),
                   "learned_batched_over_plain_per_frame": ratio(learned, per_frame),
                   "plain_batched_over_plain_per_frame": ratio(plain, per_frame)},
        "roofline": roofline,
        "roofline_next": roofline_other,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
