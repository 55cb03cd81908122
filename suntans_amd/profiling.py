"""Summaries of rocprofv3 output (CSV format) as short text files.

* ``stats``: the ``*_kernel_stats.csv`` of ``rocprofv3 --kernel-trace --stats
  --output-format csv`` -> one line per kernel (calls, average and total
  duration).
* ``pmc``: two ``*_counter_collection.csv`` files of ``rocprofv3 --pmc
  FETCH_SIZE`` and ``--pmc WRITE_SIZE`` runs -> HBM bytes per dispatch per
  kernel.  The counters are in KiB and are scaled by what they read for the
  calibration copy kernel ``k_copy_probe`` (1 GiB read + 1 GiB written per
  launch), so the bytes are the copy-calibrated HBM traffic.

    python -m suntans_amd.profiling stats <dir> [--cmd "..."] > profiles/x.txt
    python -m suntans_amd.profiling pmc <fetch dir> <write dir> [--cmd "..."]

The pmc summary ends with machine-readable ``traffic <kernel> <fetch bytes>
<write bytes>`` lines that bench.py reads for the roofline ``traffic`` field.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import sys
from collections import defaultdict

COPY_KERNEL = "k_copy_probe"
COPY_BYTES = 1 << 30


def short_name(name: str) -> str:
    """'void k_qapply<2, 3, 4, double, double>(sunhip::Geo, ...)' -> 'k_qapply<2, 3, 4, double, double>'."""
    n = re.sub(r"^void ", "", name.strip())
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def _find(d: str, suffix: str) -> list[str]:
    return sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))


def read_stats(d: str) -> list[dict]:
    rows = []
    for f in _find(d, "kernel_stats.csv"):
        for r in csv.DictReader(open(f)):
            rows.append({"name": short_name(r["Name"]), "calls": int(r["Calls"]),
                         "total_ns": float(r["TotalDurationNs"]), "avg_ns": float(r["AverageNs"]),
                         "pct": float(r["Percentage"])})
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def read_counter(d: str, counter: str) -> dict[str, tuple[float, int]]:
    """kernel -> (mean counter value per dispatch, dispatches)."""
    acc = defaultdict(list)
    for f in _find(d, "counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] != counter:
                continue
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def pmc_traffic(fetch_dir: str, write_dir: str) -> tuple[dict, float, float]:
    f = read_counter(fetch_dir, "FETCH_SIZE")
    w = read_counter(write_dir, "WRITE_SIZE")
    if COPY_KERNEL not in f or COPY_KERNEL not in w:
        raise SystemExit(f"calibration kernel {COPY_KERNEL} missing from the counter runs")
    kf = COPY_BYTES / f[COPY_KERNEL][0]
    kw = COPY_BYTES / w[COPY_KERNEL][0]
    out = {}
    for k in sorted(set(f) | set(w)):
        fb = f.get(k, (0.0, 0))[0] * kf
        wb = w.get(k, (0.0, 0))[0] * kw
        out[k] = {"fetch": fb, "write": wb, "n": max(f.get(k, (0, 0))[1], w.get(k, (0, 0))[1])}
    return out, kf, kw


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--cmd", default="")
    a.add_argument("--top", type=int, default=25)
    b = sub.add_parser("pmc")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--cmd", default="")
    args = ap.parse_args(argv)
    w = sys.stdout.write
    if args.what == "stats":
        rows = read_stats(args.dir)
        if args.cmd:
            w(f"command: {args.cmd}\n\n")
        w(f"{'kernel':<60} {'calls':>7} {'avg_us':>10} {'total_ms':>10} {'pct':>6}\n")
        for r in rows[: args.top]:
            w(f"{r['name'][:60]:<60} {r['calls']:>7} {r['avg_ns'] / 1e3:>10.2f} {r['total_ns'] / 1e6:>10.2f} "
              f"{r['pct']:>6.2f}\n")
    else:
        t, kf, kw = pmc_traffic(args.fetch_dir, args.write_dir)
        if args.cmd:
            w(f"command: {args.cmd}\n\n")
        w(f"calibration: {COPY_KERNEL} moves {COPY_BYTES} B each way per launch; bytes per KiB of counter: "
          f"FETCH_SIZE x{kf:.1f} (={kf / 1024:.3f} KiB), WRITE_SIZE x{kw:.1f} (={kw / 1024:.3f} KiB)\n\n")
        w(f"{'kernel':<60} {'dispatches':>10} {'fetch_GB':>9} {'write_GB':>9}\n")
        for k, v in sorted(t.items(), key=lambda kv: -(kv[1]["fetch"] + kv[1]["write"])):
            w(f"{k[:60]:<60} {v['n']:>10} {v['fetch'] / 1e9:>9.4f} {v['write'] / 1e9:>9.4f}\n")
        w("\n")
        for k, v in t.items():
            w(f"traffic {k.replace(' ', '')} {v['fetch']:.0f} {v['write']:.0f}\n")


def load_traffic(path: str) -> dict[str, tuple[float, float]]:
    """Parse the ``traffic`` lines of a pmc summary: kernel -> (fetch, write) bytes."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        p = line.split()
        if len(p) == 4 and p[0] == "traffic":
            out[p[1]] = (float(p[2]), float(p[3]))
    return out


if __name__ == "__main__":
    main()
