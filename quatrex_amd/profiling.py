"""Profiling helpers: HBM-counter calibration and rocprofv3 summaries.

``python -m quatrex_amd.profiling copy --gib G`` runs a plain device copy of
``G`` GiB (torch ``copy_``) so that rocprofv3 ``--pmc FETCH_SIZE`` /
``WRITE_SIZE`` readings of a kernel can be scaled by what the same counters
report for a copy of known size.

``python -m quatrex_amd.profiling summarize <csv...>`` condenses rocprofv3
CSV outputs (kernel_stats / kernel_trace / counter_collection) into the short
text tables kept under ``profiles/``.
"""

from __future__ import annotations

import argparse
import csv
import sys
from collections import defaultdict
from pathlib import Path


def copy_calibration(gib: float, reps: int) -> None:
    import torch

    n = int(gib * (1 << 30)) // 16
    a = torch.empty(n, dtype=torch.complex128, device="cuda")
    b = torch.empty_like(a)
    a.real.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {reps} x {n * 16} bytes")


def _rows(path: Path):
    with open(path, newline="") as f:
        yield from csv.DictReader(f)


def summarize(paths: list[str]) -> None:
    for p in map(Path, paths):
        rows = list(_rows(p))
        if not rows:
            continue
        cols = rows[0].keys()
        print(f"# {p.name}")
        if "TotalDurationNs" in cols:  # kernel_stats.csv
            tot = sum(float(r["TotalDurationNs"]) for r in rows)
            print(f"{'kernel':70s} {'calls':>7s} {'total ms':>10s} {'avg us':>10s} {'%':>6s}")
            for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:20]:
                t = float(r["TotalDurationNs"])
                print(f"{r['Name'][:70]:70s} {int(r['Calls']):7d} {t / 1e6:10.2f} "
                      f"{float(r['AverageNs']) / 1e3:10.2f} {100 * t / tot:6.2f}")
        elif "Counter_Name" in cols:  # counter_collection.csv
            acc = defaultdict(lambda: [0.0, 0])
            for r in rows:
                key = (r["Kernel_Name"], r["Counter_Name"])
                acc[key][0] += float(r["Counter_Value"])
                acc[key][1] += 1
            disp = defaultdict(set)
            for r in rows:
                disp[(r["Kernel_Name"], r["Counter_Name"])].add(r.get("Dispatch_Id", ""))
            print(f"{'kernel':70s} {'counter':>12s} {'dispatches':>10s} {'avg/dispatch':>14s}")
            for (k, c), (v, _) in sorted(acc.items(), key=lambda kv: -kv[1][0])[:20]:
                nd = max(len(disp[(k, c)]), 1)
                print(f"{k[:70]:70s} {c:>12s} {nd:10d} {v / nd:14.1f}")
        elif "Start_Timestamp" in cols:  # kernel_trace.csv
            acc = defaultdict(lambda: [0.0, 0])
            for r in rows:
                d = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
                acc[r["Kernel_Name"]][0] += d
                acc[r["Kernel_Name"]][1] += 1
            for k, (t, n) in sorted(acc.items(), key=lambda kv: -kv[1][0])[:20]:
                print(f"{k[:70]:70s} {n:7d} {t / 1e6:10.2f} ms {t / n / 1e3:10.2f} us")
        print()


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--gib", type=float, default=1.0)
    c.add_argument("--reps", type=int, default=4)
    s = sub.add_parser("summarize")
    s.add_argument("paths", nargs="+")
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy_calibration(args.gib, args.reps)
    else:
        summarize(args.paths)


if __name__ == "__main__":
    main(sys.argv[1:])
