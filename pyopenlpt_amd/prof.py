"""Summaries of rocprofv3 CSV output for the profiles/ folder.

  python -m pyopenlpt_amd.prof stats <run_kernel_stats.csv> [--top 25]
      per-kernel calls / total / average duration (from --kernel-trace --stats)

  python -m pyopenlpt_amd.prof pmc --fetch <fetch_counter_collection.csv>
      --write <write_counter_collection.csv> --kernel k_shake
      --calib-kernel neg --calib-bytes N
      HBM traffic per dispatch of one kernel from FETCH_SIZE / WRITE_SIZE
      (KiB per dispatch), each scaled by what the same counter reads for a
      plain elementwise kernel that reads N bytes and writes N bytes.

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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--top", type=int, default=25)
    b = sub.add_parser("pmc")
    b.add_argument("--fetch", required=True)
    b.add_argument("--write", required=True)
    b.add_argument("--kernel", required=True)
    b.add_argument("--calib-kernel", required=True)
    b.add_argument("--calib-bytes", type=float, required=True)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats(args.csv, args.top))
    else:
        print(pmc(args.fetch, args.write, args.kernel, args.calib_kernel, args.calib_bytes))
    return 0


if __name__ == "__main__":
    sys.exit(main())
