#!/usr/bin/env python
"""RGF energy points/s on one MI355X: selected solve at 32 x 256 blocks, complex128.

A step is one selected solve (G^R with off-diagonal blocks, G^<, G^>) over all
energies of the workload, whose inputs (A, Sigma^<, Sigma^> for every energy)
are already resident in HBM.  Prints ONE JSON line (rank 0).

The workload: 32 diagonal blocks of 256, 512 energies E ~ U(-1,1)
(default_rng(777)), eta = 1e-3, tight-binding H and Sigma drawn with
default_rng(20262); A = (E + i eta) I - H - Sigma_B.

``roofline`` is measured live: every sweep GEMM launch is bracketed by HIP
events on the library's stream; achieved = algorithmic flops per launch
(8 m n k per complex product term) / average launch duration, against the
FP64 matrix peak.  ``cpu_baseline`` times the numpy restatement of the
reference algorithm (oracle/) on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def flops_per_energy(n_blocks: int, b: int) -> float:
    """Algorithmic work of one energy (reference op sequence, return_retarded, no current)."""
    return 8.0 * b**3 * (34 * n_blocks - 29)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-blocks", type=int, default=32)
    ap.add_argument("--block-size", type=int, default=256)
    ap.add_argument("--energies", type=int, default=512)
    ap.add_argument("--eta", type=float, default=1e-3)
    ap.add_argument("--chunk", type=int, default=0, help="max energies per internal sweep (0=auto)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision legs")
    ap.add_argument("--plain-sample", type=int, default=8,
                    help="energies for the one-energy-at-a-time unfused FP64 reference point")
    ap.add_argument("--traffic-file", default=str(ROOT / "profiles" / "zgemm_traffic_32x256.json"))
    return ap.parse_args()


def cpu_baseline(model, energies, eta, budget_s):
    """Time the oracle's numpy restatement on the host (bounded sample)."""
    from oracle.rgf_ref import BlockTri, selected_solve
    from quatrex_amd import workloads as W

    try:
        from threadpoolctl import threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_limits = None
    cores = os.cpu_count() or 1
    g = model.geometry

    def run(n):
        ad, au, al = W.system_matrix(model, energies[:n], eta)
        sld, slu, sgd, sgu = W.sigmas(model, n)
        bs = list(g.block_sizes)

        def bt(d, u, l):
            m = BlockTri(bs, (n,))
            for i in range(g.n_blocks):
                m._data[(i, i)][...] = g.diag(d, i)
                if i + 1 < g.n_blocks:
                    m._data[(i, i + 1)][...] = g.upper(u, i)
                    m._data[(i + 1, i)][...] = g.lower(l, i) if l is not None else -np.conj(
                        np.swapaxes(g.upper(u, i), -1, -2))
            return m

        A, L, G = bt(ad, au, al), bt(sld, slu, None), bt(sgd, sgu, None)
        t0 = time.perf_counter()
        selected_solve(A, L, G, return_retarded=True, max_batch_size=n)
        return time.perf_counter() - t0

    import contextlib

    limiter = threadpool_limits(limits=cores) if threadpool_limits else contextlib.nullcontext()
    with limiter:
        t1 = run(1)
        n = int(max(1, min(16, budget_s // max(t1, 1e-3))))
        tn = run(n) if n > 1 else t1
    return {
        "value": n / tn,
        "unit": "energy_points/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} energies of the {g.n_blocks}x{g.block_sizes[0]} workload (numpy complex128 "
                  f"restatement of rgf.py, {cores} BLAS threads, {tn:.1f} s)",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    g = Geometry.uniform(args.n_blocks, args.block_size)
    seed = W.DEVICE_SEED_BASE + 2
    model = W.make_device(g, seed)
    energies = W.draw_energies(args.energies)
    E = args.energies

    # inputs resident in HBM
    A = W.system_matrix_torch(model, energies, args.eta, dev)
    sld = torch.as_tensor(model.sl_diag, device=dev).expand(E, -1).contiguous()
    sgd = torch.as_tensor(model.sg_diag, device=dev).expand(E, -1).contiguous()
    slu = torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=dev)
    sgu = torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=dev)
    # outputs: G^R (diag, upper, lower), G^< and G^> (diag, upper)
    GR = S.packed_empty(g, E, dev)
    GL = S.packed_empty(g, E, dev, parts=("d", "u"))
    GG = S.packed_empty(g, E, dev, parts=("d", "u"))
    ctx = S.default_context(dev)
    if args.chunk:
        ctx.set_option(1, args.chunk)

    def step():
        S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                         ctx=ctx)

    for _ in range(args.warmup):
        step()
    ctx.status()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read(reset=True)
    ctx.profile(False)
    ctx.status()

    elapsed = t1 - t0
    eps = E * args.steps / elapsed
    f_e = flops_per_energy(args.n_blocks, args.block_size)
    gemm_avg_ms = prof["gemm_ms"] / max(prof["gemm_launches"], 1)
    gemm_flops_launch = prof["gemm_flops"] / max(prof["gemm_launches"], 1)
    achieved = gemm_flops_launch / (gemm_avg_ms * 1e-3) / 1e12
    traffic = None
    tf = Path(args.traffic_file)
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    extra = {}
    if not args.no_mixed:
        # FP64 results of a sample of energies are the reference for the mixed errors
        ns = min(16, E)
        ref = [X[0][:ns].clone() for X in (GR, GL, GG)]

        def rel_err(X, R):
            d = (X[0][:ns] - R).reshape(ns, g.n_blocks, -1)
            r = R.reshape(ns, g.n_blocks, -1)
            return float((torch.linalg.norm(d, dim=-1) / torch.linalg.norm(r, dim=-1)).max())

        mixed = S.MixedOptions(refine=0, fallback_tol=0.0)

        def mstep():
            S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                             ctx=ctx, precision="mixed", mixed=mixed)

        mstep()
        ctx.status()
        torch.cuda.synchronize()
        tm0 = time.perf_counter()
        for _ in range(args.steps):
            mstep()
        torch.cuda.synchronize()
        tm = time.perf_counter() - tm0
        ctx.status()
        eps_mixed = E * args.steps / tm
        errs = {k: rel_err(X, R) for k, X, R in zip(("G^R", "G^<", "G^>"), (GR, GL, GG), ref)}
        # mixed + FP64 a-posteriori check, FP64 fallback above 1e-8
        flags = torch.zeros(E, dtype=torch.int32, device=dev)
        resid = torch.zeros(E, dtype=torch.float64, device=dev)
        checked = S.MixedOptions(refine=1, fallback_tol=1e-8, fallback_flags=flags, residuals=resid)
        torch.cuda.synchronize()
        tc0 = time.perf_counter()
        S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                         ctx=ctx, precision="mixed", mixed=checked)
        torch.cuda.synchronize()
        tc = time.perf_counter() - tc0
        ctx.status()
        errs_fb = {k: rel_err(X, R) for k, X, R in zip(("G^R", "G^<", "G^>"), (GR, GL, GG), ref)}
        # time-weighted roofline: F_gemm at the FP32 matrix peak, F_inv at the FP64 peak
        b3 = 8.0 * args.block_size ** 3
        t_roof = b3 * ((33 * args.n_blocks - 29) / (FP32_PEAK_TFLOPS * 1e12)
                       + args.n_blocks / (FP64_PEAK_TFLOPS * 1e12))
        extra["mixed"] = {
            "value": round(eps_mixed, 3),
            "unit": "energy_points/s",
            "speedup_vs_fp64": round(eps_mixed / eps, 3),
            "max_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
            "error_sample_energies": ns,
            "roofline_frac": round(t_roof * eps_mixed, 4),
            "with_check_and_fp64_fallback": {
                "value": round(E / tc, 3),
                "fallback_fraction": round(float(flags.float().mean()), 4),
                "median_residual": float(f"{float(resid.median()):.3e}"),
                "max_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs_fb.items()},
            },
        }
    if args.plain_sample > 0:
        # plain FP64 GPU RGF: one energy per sweep, one product group per launch
        ns = min(args.plain_sample, E)
        sl_ = slice(0, ns)
        view = lambda T: tuple(t[sl_] if t is not None else None for t in T)  # noqa: E731
        pctx = S.Context(dev)
        pctx.set_option(1, 1)
        pctx.set_option(5, 1)

        def pstep():
            S.selected_solve(g, view(A), (sld[sl_], slu[sl_]), (sgd[sl_], sgu[sl_]), g_lesser=view(GL),
                             g_greater=view(GG), g_retarded=view(GR), ctx=pctx)

        pstep()
        pctx.status()
        tp0 = time.perf_counter()
        pstep()
        torch.cuda.synchronize()
        tp = time.perf_counter() - tp0
        pctx.status()
        eps_plain = ns / tp
        extra["plain_fp64"] = {
            "value": round(eps_plain, 3),
            "unit": "energy_points/s",
            "definition": "one energy per sweep, one product group per launch (FP64)",
            "sample_energies": ns,
            "batched_fp64_speedup": round(eps / eps_plain, 2),
        }
        if "mixed" in extra:
            extra["mixed"]["speedup_vs_plain_fp64"] = round(extra["mixed"]["value"] / eps_plain, 2)
        pctx.close()

    out = {
        "metric": "RGF energy-points/s at 32x256 c128 (FP64 vs mixed); p50/p99 solve latency",
        "value": round(eps, 3),
        "unit": "energy_points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded tight-binding H, Sigma^<, Sigma^>; energies U(-1,1))",
        "config": {
            "workload": f"RGF selected solve, {args.n_blocks} blocks x {args.block_size}, complex128, "
                        f"{E} energies, eta={args.eta}",
            "n_blocks": args.n_blocks,
            "block_size": args.block_size,
            "energies": E,
            "eta": args.eta,
            "precision": "fp64",
            "outputs": "G^R diag+upper+lower, G^< and G^> diag+upper",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "zgemm_grouped_kernel<64,64,0> (sweep products)",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "avg_launch_ms": round(gemm_avg_ms, 4),
            "flops_per_launch": gemm_flops_launch,
            "launches": prof["gemm_launches"],
        },
        "sweep_roofline": {
            "flops_per_energy": f_e,
            "achieved_tflops": round(f_e * eps / 1e12, 3),
            "frac_fp64_peak": round(f_e * eps / 1e12 / FP64_PEAK_TFLOPS, 4),
            "inverse_ms_share": round(prof["inv_ms"] / max(prof["inv_ms"] + prof["gemm_ms"], 1e-9), 4),
            "inverse_tflops": round(prof["inv_flops"] / max(prof["inv_ms"], 1e-9) / 1e9, 3),
        },
        **extra,
    }

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(model, energies, args.eta, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
