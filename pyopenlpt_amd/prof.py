"""Summaries of rocprofv3 CSV output for the profiles/ folder.

  python -m pyopenlpt_amd.prof stats <run_kernel_stats.csv> [--top 25]
      per-kernel calls / total / average duration (from --kernel-trace --stats)

  python -m pyopenlpt_amd.prof trace <run_kernel_trace.csv> [--top 25] [--after PATTERN]
      the same from a kernel trace, optionally only the dispatches after the
      last one matching PATTERN (the timed steps after a training phase)

  python -m pyopenlpt_amd.prof pmc --fetch <fetch_counter_collection.csv>
      --write <write_counter_collection.csv> --kernel k_shake
      --calib-kernel neg --calib-bytes N
      HBM traffic per dispatch of one kernel from FETCH_SIZE / WRITE_SIZE
      (KiB per dispatch), each scaled by what the same counter reads for a
      plain elementwise kernel that reads N bytes and writes N bytes.

  python -m pyopenlpt_amd.prof table --stats <run_kernel_stats.csv>
      --fetch <csv> --write <csv> [--sq <csv>] --calib-kernel neg --calib-bytes N
      per-kernel table: time share, average duration, HBM bytes per dispatch
      (calibrated FETCH_SIZE + WRITE_SIZE) and the achieved GB/s they give;
      with --sq, VALU instructions per wave and the share of wave cycles spent
      waiting (SQ counters of a third run)

  python -m pyopenlpt_amd.prof step [--frames 16] [--batch 16] [--steps 1]
      [--shake-loops 6] [--calib-mib 0] [--learned]
      the program the profiles are taken on: the ppp 0.05 clustered frames
      (1024^2 u16, 4 cameras) resident in HBM, one eager olpt_ipr_run over
      all of them per step on one stream, `batch` frames per launch; then,
      with --calib-mib, three launches of an elementwise kernel that reads
      and writes that many MiB (the counters' calibration)

Both print text for profiles/ and, for pmc, the key lines bench.py reads:
  fetch_bytes_per_dispatch: <float>
  write_bytes_per_dispatch: <float>
"""
from __future__ import annotations

import argparse
import csv
import re
import statistics
import sys


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"\(.*", "", name)
    return name[:110]


def stats(path: str, top: int) -> str:
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    out = [f"total kernel time {tot / 1e6:.3f} ms over {sum(int(r['Calls']) for r in rows)} "
           "dispatches", "",
           "| kernel | calls | total ms | % | avg us |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{_short(r['Name'])}` | {int(r['Calls'])} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} |")
    return "\n".join(out)


def trace(path: str, top: int, after: str | None) -> str:
    """Per-kernel calls / total / average from a kernel-trace CSV; with
    `after`, only the dispatches that follow the last one whose name matches
    that pattern (e.g. the timed steps after a training phase)."""
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    if after:
        last = max((i for i, r in enumerate(rows) if re.search(after, r["Kernel_Name"])),
                   default=-1)
        rows = rows[last + 1:]
    tot: dict[str, float] = {}
    calls: dict[str, int] = {}
    for r in rows:
        n = _short(r["Kernel_Name"])
        tot[n] = tot.get(n, 0.0) + int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        calls[n] = calls.get(n, 0) + 1
    T = sum(tot.values())
    out = [f"total kernel time {T / 1e6:.3f} ms over {len(rows)} dispatches", "",
           "| kernel | calls | total ms | % | avg us |", "|---|---:|---:|---:|---:|"]
    for n in sorted(tot, key=lambda k: -tot[k])[:top]:
        out.append(f"| `{n}` | {calls[n]} | {tot[n] / 1e6:.3f} | {100 * tot[n] / T:.1f} | "
                   f"{tot[n] / calls[n] / 1e3:.1f} |")
    return "\n".join(out)


def _pmc_values(path: str, counter: str) -> dict[str, list[float]]:
    vals: dict[str, list[float]] = {}
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name") != counter:
            continue
        vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return vals


def _pick(vals: dict[str, list[float]], pattern: str) -> list[float]:
    hits = [v for k, v in vals.items() if re.search(pattern, k)]
    if not hits:
        raise SystemExit(f"no dispatch matches {pattern!r}")
    return [x for h in hits for x in h]


def pmc(fetch: str, write: str, kernel: str, calib_kernel: str, calib_bytes: float) -> str:
    f = _pmc_values(fetch, "FETCH_SIZE")
    w = _pmc_values(write, "WRITE_SIZE")
    kf, kw = _pick(f, kernel), _pick(w, kernel)
    cf, cw = _pick(f, calib_kernel), _pick(w, calib_kernel)
    # the calibration kernel reads and writes calib_bytes; the factor maps
    # the counter's KiB to bytes of HBM traffic
    scale_f = calib_bytes / statistics.median(cf)
    scale_w = calib_bytes / statistics.median(cw)
    fb = statistics.mean(kf) * scale_f
    wb = statistics.mean(kw) * scale_w
    lines = [
        f"kernel pattern: {kernel}  ({len(kf)} dispatches with FETCH_SIZE, {len(kw)} with WRITE_SIZE)",
        f"calibration: {calib_kernel}, {calib_bytes:.0f} bytes read + written per dispatch; "
        f"FETCH_SIZE median {statistics.median(cf):.1f} KiB -> {scale_f:.2f} B/KiB-unit, "
        f"WRITE_SIZE median {statistics.median(cw):.1f} KiB -> {scale_w:.2f} B/KiB-unit",
        f"FETCH_SIZE per dispatch: mean {statistics.mean(kf):.1f} KiB "
        f"(min {min(kf):.1f}, max {max(kf):.1f})",
        f"WRITE_SIZE per dispatch: mean {statistics.mean(kw):.1f} KiB "
        f"(min {min(kw):.1f}, max {max(kw):.1f})",
        f"fetch_bytes_per_dispatch: {fb:.1f}",
        f"write_bytes_per_dispatch: {wb:.1f}",
        f"traffic_bytes_per_dispatch: {fb + wb:.1f}",
    ]
    return "\n".join(lines)


def _dispatch_means(path: str, counter: str) -> dict[str, tuple[float, int]]:
    vals: dict[str, list[float]] = {}
    for r in csv.DictReader(open(path)):
        if r.get("Counter_Name") != counter:
            continue
        vals.setdefault(_short(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return {k: (statistics.mean(v), len(v)) for k, v in vals.items()}


def table(stats_csv: str, fetch: str, write: str, sq: str | None, calib_kernel: str,
          calib_bytes: float, min_pct: float) -> str:
    rows = list(csv.DictReader(open(stats_csv)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    f = _dispatch_means(fetch, "FETCH_SIZE")
    w = _dispatch_means(write, "WRITE_SIZE")
    cf = [v for k, v in f.items() if re.search(calib_kernel, k)]
    cw = [v for k, v in w.items() if re.search(calib_kernel, k)]
    if not cf or not cw:
        raise SystemExit(f"calibration kernel {calib_kernel!r} not in the counter runs")
    scale_f = calib_bytes / cf[0][0]
    scale_w = calib_bytes / cw[0][0]
    sqv = {}
    if sq:
        for name in ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_INST_ANY",
                     "SQ_INSTS_VMEM_RD", "SQ_INSTS_LDS"):
            sqv[name] = _dispatch_means(sq, name)
    out = [f"calibration: {calib_kernel} reads and writes {calib_bytes:.0f} B per dispatch; "
           f"FETCH_SIZE x {scale_f:.1f}, WRITE_SIZE x {scale_w:.1f} give bytes", "",
           f"kernels above {min_pct:g}% of the kernel time ({tot / 1e6:.2f} ms in all):", "",
           "| kernel | % time | calls | avg us | HBM MB/dispatch (read + write) | achieved GB/s |"
           + (" VALU insts/wave | wait % of wave cycles |" if sq else ""),
           "|---|---:|---:|---:|---:|---:|" + ("---:|---:|" if sq else "")]
    for r in rows:
        pct = 100.0 * float(r["TotalDurationNs"]) / tot
        if pct < min_pct:
            continue
        k = _short(r["Name"])
        avg_ns = float(r["AverageNs"])
        fb = f[k][0] * scale_f if k in f else float("nan")
        wb = w[k][0] * scale_w if k in w else float("nan")
        line = (f"| `{k}` | {pct:.1f} | {int(r['Calls'])} | {avg_ns / 1e3:.1f} | "
                f"{fb / 1e6:.2f} + {wb / 1e6:.2f} | {(fb + wb) / avg_ns:.0f} |")
        if sq:
            waves = sqv["SQ_WAVES"].get(k, (float("nan"), 0))[0]
            valu = sqv["SQ_INSTS_VALU"].get(k, (float("nan"), 0))[0]
            cyc = sqv["SQ_WAVE_CYCLES"].get(k, (float("nan"), 0))[0]
            wait = sqv["SQ_WAIT_INST_ANY"].get(k, (float("nan"), 0))[0]
            line += f" {valu / waves:.0f} | {100.0 * wait / cyc:.0f} |" if waves and cyc else " | |"
        out.append(line)
    return "\n".join(out)


def step(frames: int, batch: int, steps: int, shake_loops: int, calib_mib: int,
         learned: bool = False) -> int:
    import concurrent.futures as cf
    import multiprocessing as mp
    import time

    import numpy as np
    import torch

    from . import synth as S
    from .context import Context, to_device_image

    ex = cf.ProcessPoolExecutor(max_workers=min(16, frames),
                                mp_context=mp.get_context("spawn"))
    try:
        ws = list(ex.map(S.workload_c3_frame, range(frames)))
    finally:
        ex.shutdown(wait=True)
    w0 = ws[0]
    H, W = w0.images.shape[1:]
    d, code = to_device_image(np.stack([w.images for w in ws]))
    otf, _keep = S.make_otf(len(w0.cams), w0.otf_abca)
    ctx = Context(w0.cams, H, W, precision="f32", max_2d=1 << 16, max_cand=1 << 20,
                  max_obj=1 << 17)
    ctx.set_otf(otf)
    ctx.set_tuning(n_streams=1, frames_per_batch=batch)
    params = S.OlptIprParams.from_buffer_copy(w0.params)
    params.n_shake_loop = shake_loops
    if learned:
        # the learned leg's schedule: a briefly trained proposal network, then
        # one plain iteration from 0.03 voxel
        from .learned import ShakeNet, make_training_set, train
        tctx = Context(w0.cams, H, W, precision="f32")
        tctx.set_otf(otf)
        X, T, _ = make_training_set(tctx, w0.params, w0.voxel_mm, n_particles=500_000, seed=9001)
        net = ShakeNet(max_batch=8192, seed=0)
        train(net, X, T, steps=500, batch=8192, lr=3e-3, ghost_weight=0.5)
        tctx.close()
        del X, T
        ctx.set_shake_model(net, w0.voxel_mm, 0.6)
        params.n_shake_loop = 1
        params.shake_width = 0.03 * w0.voxel_mm
    outs = ctx.alloc_outputs(frames, 1 << 17, d.device)
    for s in range(steps):
        t0 = time.perf_counter()
        ctx.ipr_run(d, code, params, n_frames=frames, cap_obj=1 << 17, outputs=outs)
        torch.cuda.synchronize()
        print(f"step {s}: {1000 * (time.perf_counter() - t0):.1f} ms, "
              f"{int(outs[2].sum())} objects over {frames} frames", flush=True)
    if calib_mib > 0:
        src = torch.ones(calib_mib * (1 << 20) // 4, dtype=torch.float32, device=d.device)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(f"calibration: 3 x neg over {calib_mib} MiB", flush=True)
    ctx.close()
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--top", type=int, default=25)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--top", type=int, default=25)
    t.add_argument("--after", default=None,
                   help="only dispatches after the last one matching this pattern")
    b = sub.add_parser("pmc")
    b.add_argument("--fetch", required=True)
    b.add_argument("--write", required=True)
    b.add_argument("--kernel", required=True)
    b.add_argument("--calib-kernel", required=True)
    b.add_argument("--calib-bytes", type=float, required=True)
    c = sub.add_parser("table")
    c.add_argument("--stats", required=True)
    c.add_argument("--fetch", required=True)
    c.add_argument("--write", required=True)
    c.add_argument("--sq")
    c.add_argument("--calib-kernel", required=True)
    c.add_argument("--calib-bytes", type=float, required=True)
    c.add_argument("--min-pct", type=float, default=2.0)
    d = sub.add_parser("step")
    d.add_argument("--frames", type=int, default=16)
    d.add_argument("--batch", type=int, default=16)
    d.add_argument("--steps", type=int, default=1)
    d.add_argument("--shake-loops", type=int, default=6)
    d.add_argument("--calib-mib", type=int, default=0)
    d.add_argument("--learned", action="store_true",
                   help="the learned leg's schedule (a briefly trained proposal, 1 iteration)")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats(args.csv, args.top))
    elif args.cmd == "trace":
        print(trace(args.csv, args.top, args.after))
    elif args.cmd == "pmc":
        print(pmc(args.fetch, args.write, args.kernel, args.calib_kernel, args.calib_bytes))
    elif args.cmd == "table":
        print(table(args.stats, args.fetch, args.write, args.sq, args.calib_kernel,
                    args.calib_bytes, args.min_pct))
    else:
        return step(args.frames, args.batch, args.steps, args.shake_loops, args.calib_mib,
                    args.learned)
    return 0


if __name__ == "__main__":
    sys.exit(main())
