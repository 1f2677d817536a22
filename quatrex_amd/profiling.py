"""Profiling helpers: HBM-counter calibration and rocprofv3 summaries.

``python -m quatrex_amd.profiling copy --gib G`` runs a plain device copy of
``G`` GiB (torch ``copy_``) so that rocprofv3 ``--pmc FETCH_SIZE`` /
``WRITE_SIZE`` readings of a kernel can be scaled by what the same counters
report for a copy of known size.

``python -m quatrex_amd.profiling summarize <csv...>`` condenses rocprofv3
CSV outputs (kernel_stats / kernel_trace / counter_collection) into the short
text tables kept under ``profiles/``.

``python -m quatrex_amd.profiling traffic ...`` turns the FETCH_SIZE and
WRITE_SIZE counter files of a kernel and of the calibration copy into HBM
bytes per launch (the JSON that bench.py reports as ``roofline.traffic``).

``python -m quatrex_amd.profiling sq <counter_collection.csv>`` prints the
matrix-core busy fraction and wait fractions per kernel from SQ counters.

``python -m quatrex_amd.profiling dbstats <results.db>`` prints per-kernel
call counts and durations from a rocprofv3 ``--kernel-trace`` database.

``python -m quatrex_amd.profiling inv --n N --energies E`` times the batched
complex inverse (``rgf_inv_z``) of E diagonally dominant N x N matrices and
checks the first result against numpy.

``python -m quatrex_amd.profiling sweep --n-blocks N --block-size B --energies E``
runs the FP64 selected solve of the synthetic device (default tiles or
``--tile``, no autotuner launches; ``--persistent 0/1`` picks the path for
blocks up to 32, ``--mixed`` the complex64-product sweep without its check)
and prints energies/s and the sweep products' average launch time, flops and
algorithmic bytes per launch from the library's HIP-event profile, so that a
kernel trace and HBM counters of the same command can be set beside them.
"""

from __future__ import annotations

import argparse
import csv
import sys
import time
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


def time_inverse(n: int, energies: int, reps: int) -> None:
    import numpy as np
    import torch

    from quatrex_amd import solve as S

    dev = torch.device("cuda", 0)
    a = torch.complex(torch.rand(energies, n, n, dtype=torch.float64, device=dev) - 0.5,
                      torch.rand(energies, n, n, dtype=torch.float64, device=dev) - 0.5)
    a += 4.0 * torch.eye(n, dtype=torch.complex128, device=dev)
    out = torch.empty_like(a)
    S.inv(a, out=out)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        t0.record()
        S.inv(a, out=out)
        t1.record()
        t1.synchronize()
        best = min(best, t0.elapsed_time(t1))
    x = out[0].cpu().numpy()
    err = np.abs(a[0].cpu().numpy() @ x - np.eye(n)).max()
    flops = 8.0 * n ** 3 * energies
    print(f"inverse n={n} energies={energies}: {best:.3f} ms, {flops / best / 1e9:.2f} TFLOP/s "
          f"(8 n^3 per matrix), max |A X - I| = {err:.2e}")


def _rows(path: Path):
    with open(path, newline="") as f:
        yield from csv.DictReader(f)


def _short(name: str) -> str:
    """Kernel name without the return type, namespaces and argument list."""
    for pre in ("void ", "qx::(anonymous namespace)::", "qx::", "(anonymous namespace)::"):
        if name.startswith(pre):
            name = name[len(pre):]
    name = name.replace("qx::(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def summarize(paths: list[str]) -> None:
    for p in map(Path, paths):
        rows = list(_rows(p))
        if not rows:
            continue
        cols = rows[0].keys()
        print(f"# {p.name}")
        if "TotalDurationNs" in cols:  # kernel_stats.csv
            tot = sum(float(r["TotalDurationNs"]) for r in rows)
            print(f"{'kernel':64s} {'calls':>7s} {'total ms':>10s} {'avg us':>10s} {'%':>6s}")
            for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:20]:
                t = float(r["TotalDurationNs"])
                print(f"{_short(r['Name'])[:64]:64s} {int(r['Calls']):7d} {t / 1e6:10.2f} "
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


def sq_summary(path: str, top: int = 6) -> None:
    """Derived SQ metrics per kernel from a ``--pmc`` counter file holding
    SQ_WAVES, SQ_WAVE_CYCLES, SQ_BUSY_CYCLES, SQ_WAIT_INST_ANY,
    SQ_VALU_MFMA_BUSY_CYCLES, SQ_WAIT_INST_LDS, SQ_INSTS_LDS and one
    SQ_INSTS_VALU_MFMA_* counter.  SQ_BUSY_CYCLES counts per shader engine
    (32 on MI355X, 1024 SIMDs), so the SIMD cycles of a kernel are
    SQ_BUSY_CYCLES x 32 and the matrix-core busy fraction is
    SQ_VALU_MFMA_BUSY_CYCLES over that; wave-cycle counters are in units of
    four cycles on both sides of the wait fractions."""
    acc = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    for r in _rows(Path(path)):
        acc[r["Kernel_Name"]][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[r["Kernel_Name"]].add(r.get("Dispatch_Id", ""))
    order = sorted(acc, key=lambda k: -acc[k].get("SQ_BUSY_CYCLES", 0.0))[:top]
    print(f"{'kernel':52s} {'disp':>5s} {'mfma busy':>9s} {'wait any':>8s} {'wait lds':>8s} "
          f"{'mfma/disp':>10s} {'lds/disp':>10s}")
    for k in order:
        c = acc[k]
        simd = c.get("SQ_BUSY_CYCLES", 0.0) * 32
        wc = c.get("SQ_WAVE_CYCLES", 0.0) or 1.0
        nd = max(len(disp[k]), 1)
        mf = sum(v for n, v in c.items() if n.startswith("SQ_INSTS_VALU_MFMA_"))
        print(f"{_short(k)[:52]:52s} {nd:5d} {c.get('SQ_VALU_MFMA_BUSY_CYCLES', 0.0) / max(simd, 1.0):9.3f} "
              f"{c.get('SQ_WAIT_INST_ANY', 0.0) / wc:8.3f} {c.get('SQ_WAIT_INST_LDS', 0.0) / wc:8.3f} "
              f"{mf / nd:10.0f} {c.get('SQ_INSTS_LDS', 0.0) / nd:10.0f}")


def _per_dispatch(path: Path, kernel: str, counter: str) -> tuple[float, int]:
    """Average counter value (KiB) per dispatch of kernels whose name contains ``kernel``."""
    tot = defaultdict(float)
    for r in _rows(path):
        if r["Counter_Name"] == counter and kernel in r["Kernel_Name"]:
            tot[r.get("Dispatch_Id", "")] += float(r["Counter_Value"])
    if not tot:
        raise SystemExit(f"no {counter} rows for {kernel!r} in {path}")
    return sum(tot.values()) / len(tot), len(tot)


def traffic(args) -> dict:
    """HBM bytes per launch of ``args.kernel``, scaled by the calibration copy."""
    copy_bytes = args.copy_gib * (1 << 30)
    fk, nd = _per_dispatch(Path(args.fetch), args.kernel, "FETCH_SIZE")
    wk, _ = _per_dispatch(Path(args.write), args.kernel, "WRITE_SIZE")
    cf, _ = _per_dispatch(Path(args.copy_fetch), "copyBuffer", "FETCH_SIZE")
    cw, _ = _per_dispatch(Path(args.copy_write), "copyBuffer", "WRITE_SIZE")
    fs, ws = copy_bytes / (cf * 1024), copy_bytes / (cw * 1024)
    out = {
        "kernel": args.kernel,
        "workload": args.workload,
        "fetch_kib_per_dispatch": round(fk, 1),
        "write_kib_per_dispatch": round(wk, 1),
        "dispatches": nd,
        "copy_calibration": {"bytes_per_copy": copy_bytes, "fetch_kib": round(cf, 1),
                             "write_kib": round(cw, 1), "fetch_scale": round(fs, 4),
                             "write_scale": round(ws, 4)},
        "hbm_bytes_per_launch": int(fk * 1024 * fs + wk * 1024 * ws),
    }
    if args.algorithmic_bytes:
        out["algorithmic_operand_bytes_per_launch"] = int(args.algorithmic_bytes)
    return out


def dbstats(path: str) -> None:
    import sqlite3

    con = sqlite3.connect(path)
    tot = con.execute("select sum(end - start) from kernels").fetchone()[0] or 1
    print(f"{'kernel':78s} {'calls':>6s} {'avg us':>10s} {'total ms':>10s} {'%':>6s}")
    for name, n, avg, t in con.execute(
            "select name, count(*), avg(end - start), sum(end - start) from kernels "
            "group by name order by 4 desc limit 12"):
        print(f"{name[:78]:78s} {n:6d} {avg / 1e3:10.1f} {t / 1e6:10.2f} {100 * t / tot:6.1f}")


def time_sweep(n_blocks: int, b: int, energies: int, reps: int, tile: str | None = None,
               persistent: int | None = None, mixed: bool = False) -> None:
    import torch

    from quatrex_amd import _lib
    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    dev = torch.device("cuda", 0)
    g = Geometry.uniform(n_blocks, b)
    model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
    A = W.system_matrix_torch(model, W.draw_energies(energies), 1e-3, dev)
    sld = torch.as_tensor(model.sl_diag, device=dev).expand(energies, -1).contiguous()
    sgd = torch.as_tensor(model.sg_diag, device=dev).expand(energies, -1).contiguous()
    slu = torch.zeros((energies, g.offdiag_size), dtype=torch.complex128, device=dev)
    sgu = torch.zeros_like(slu)
    GR = S.packed_empty(g, energies, dev)
    GL = S.packed_empty(g, energies, dev, parts=("d", "u"))
    GG = S.packed_empty(g, energies, dev, parts=("d", "u"))
    ctx = S.default_context(dev)
    if tile:
        ctx.set_option(_lib.RGF_OPT_TILE_CFG, _lib.TILE_CFG_NAMES.index(tile))
    if persistent is not None:
        ctx.set_option(_lib.RGF_OPT_PERSISTENT, int(persistent))
    # mixed: the complex64-product sweep without the FP64 check (its products
    # are the kernels profiled)
    kw = {"precision": "mixed", "mixed": S.MixedOptions(refine=0, fallback_tol=0.0)} if mixed else {}
    ctx.profile(True)
    t = 0.0
    for r in range(reps + 1):
        if r == 1:
            ctx.profile_read(reset=True)  # the first solve is a warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                         ctx=ctx, **kw)
    torch.cuda.synchronize()
    if reps > 0:
        t = time.perf_counter() - t0
    p = ctx.profile_read(reset=True)
    ctx.status()
    line = f"sweep {n_blocks}x{b} energies={energies}"
    if t > 0:
        line += f": {reps * energies / t:.1f} energies/s"
    n = p["gemm_launches"]
    if n and p["gemm_ms"] > 0:
        ms = p["gemm_ms"] / n
        line += (f"; {n} product launches, {ms * 1e3:.1f} us, {p['gemm_flops'] / n / 1e9:.4f} GFLOP "
                 f"and {p['gemm_bytes'] / n / 1e6:.2f} MB (algorithmic) per launch: "
                 f"{p['gemm_flops'] / p['gemm_ms'] / 1e9:.2f} TFLOP/s, "
                 f"{p['gemm_bytes'] / p['gemm_ms'] / 1e6:.1f} GB/s algorithmic")
    else:
        line += "; no product launches (on-chip sweep)"
    print(line, flush=True)


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--gib", type=float, default=1.0)
    c.add_argument("--reps", type=int, default=4)
    s = sub.add_parser("summarize")
    s.add_argument("paths", nargs="+")
    t = sub.add_parser("traffic")
    t.add_argument("--kernel", required=True, help="substring of the kernel name")
    t.add_argument("--workload", default="")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-gib", type=float, default=1.0)
    t.add_argument("--algorithmic-bytes", type=float, default=0.0)
    q = sub.add_parser("sq")
    q.add_argument("path")
    q.add_argument("--top", type=int, default=6)
    d = sub.add_parser("dbstats")
    d.add_argument("path")
    v = sub.add_parser("inv")
    v.add_argument("--n", type=int, default=256)
    v.add_argument("--energies", type=int, default=512)
    v.add_argument("--reps", type=int, default=4)
    w = sub.add_parser("sweep")
    w.add_argument("--n-blocks", type=int, default=32)
    w.add_argument("--block-size", type=int, default=32)
    w.add_argument("--energies", type=int, default=4096)
    w.add_argument("--reps", type=int, default=2)
    w.add_argument("--tile", default=None, help="GEMM tile configuration name (default: 64x64)")
    w.add_argument("--persistent", type=int, default=None,
                   help="1/0: on-chip sweep for blocks <= 32 on/off (default: library default)")
    w.add_argument("--mixed", action="store_true", help="complex64 products, no FP64 check")
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy_calibration(args.gib, args.reps)
    elif args.cmd == "traffic":
        import json

        print(json.dumps(traffic(args), indent=1))
    elif args.cmd == "sq":
        sq_summary(args.path, args.top)
    elif args.cmd == "dbstats":
        dbstats(args.path)
    elif args.cmd == "inv":
        time_inverse(args.n, args.energies, args.reps)
    elif args.cmd == "sweep":
        time_sweep(args.n_blocks, args.block_size, args.energies, args.reps, args.tile,
                   args.persistent, args.mixed)
    else:
        summarize(args.paths)


if __name__ == "__main__":
    main(sys.argv[1:])
