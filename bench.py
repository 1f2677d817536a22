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
    ap.add_argument("--no-autotune", action="store_true",
                    help="untuned GEMM tiles (FP64: 3M 64x64, 3M 32x32 below 64; FP32: 64x64)")
    ap.add_argument("--tile", default=None,
                    help="FP64 GEMM tile configuration of the main leg (skips its autotuning)")
    ap.add_argument("--torch-sample", type=int, default=100,
                    help="energies of the torch (rocBLAS/rocSOLVER) reference-algorithm leg (0 = skip)")
    ap.add_argument("--plain-sample", type=int, default=8,
                    help="energies for the one-energy-at-a-time unfused FP64 reference point")
    ap.add_argument("--serving-requests", type=int, default=1000,
                    help="requests of the serving leg (a prefix of the Poisson stream; 0 = skip)")
    ap.add_argument("--serving-rate", type=float, default=2000.0, help="offered load, requests/s")
    ap.add_argument("--serving-budget-gib", type=float, default=96.0)
    ap.add_argument("--eta-energies", "--eta-sample", dest="eta_energies", type=int, default=1024,
                    help="energies of the 48-block mixed-precision leg, eta cycled (0 = skip)")
    ap.add_argument("--eta-chunk", type=int, default=128, help="energies per solve in that leg")
    ap.add_argument("--dsdb-energies", type=int, default=64,
                    help="energies of the block-sparse densify/sparsify leg (0 = skip)")
    ap.add_argument("--uneven-energies", type=int, default=2048,
                    help="energies of the uneven-batching leg over 4 geometries (0 = skip)")
    ap.add_argument("--uneven-capacity", type=int, default=64)
    ap.add_argument("--uneven-budget-gib", type=float, default=200.0)
    ap.add_argument("--obc-energies", type=int, default=512,
                    help="energies of the open-boundary (Sancho-Rubio) leg (0 = skip)")
    ap.add_argument("--host-energies", type=int, default=64,
                    help="energies of the host-resident (PCIe-inclusive) drop-in leg (0 = skip)")
    ap.add_argument("--sections", type=int, default=4,
                    help="block sections of the per-section (partitioned solve) leg (0 = skip)")
    ap.add_argument("--small-energies", type=int, default=4096,
                    help="energies of the small-block (b = 32) shapes of the HBM-bound leg; b = 64 "
                         "runs a quarter of them (0 = skip)")
    ap.add_argument("--only-leg", choices=("torch", "eta", "uneven", "serving", "dsdb", "obc", "section",
                                           "host", "small"),
                    default=None,
                    help=argparse.SUPPRESS)
    ap.add_argument("--main-eps", type=float, default=0.0, help=argparse.SUPPRESS)
    ap.add_argument("--leg-timeout", type=float, default=300.0)
    ap.add_argument("--traffic-file", default=str(ROOT / "profiles" / "zgemm_traffic_32x256.json"))
    return ap.parse_args()


def cpu_baseline(model, energies, eta, budget_s):
    """Time the oracle's numpy restatement on the host (bounded sample)."""
    import contextlib

    from oracle.rgf_ref import BlockTri, selected_solve
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

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

    limiter = threadpool_limits(limits=cores) if threadpool_limits else contextlib.nullcontext()
    with limiter:
        t1 = run(1)
        n = int(max(1, min(64, budget_s // max(t1, 1e-3))))
        tn = run(n) if n > 1 else t1
    # the same restatement on one core: one energy
    single = None
    if threadpool_limits:
        with threadpool_limits(limits=1):
            ts = run(1)
        single = {"value": 1.0 / ts, "unit": "energy_points/s", "cores": 1,
                  "sample": f"1 energy, 1 BLAS thread, {ts:.1f} s"}
    # the CPU-runnable configuration: 8 blocks of 32, one energy, repeated
    # (median over the repeats; 1 BLAS thread, the op sequence is too small
    # to gain from more)
    small = None
    try:
        g1 = Geometry.uniform(8, 32)
        m1 = W.make_device(g1, W.DEVICE_SEED_BASE + 1)
        e1 = W.draw_energies(1)
        a1 = W.system_matrix(m1, e1, eta)
        s1 = W.sigmas(m1, 1)
        bs1 = list(g1.block_sizes)

        def bt1(d, u, l):
            m = BlockTri(bs1, (1,))
            for i in range(g1.n_blocks):
                m._data[(i, i)][...] = g1.diag(d, i)
                if i + 1 < g1.n_blocks:
                    m._data[(i, i + 1)][...] = g1.upper(u, i)
                    m._data[(i + 1, i)][...] = g1.lower(l, i) if l is not None else -np.conj(
                        np.swapaxes(g1.upper(u, i), -1, -2))
            return m

        A1, L1, G1 = bt1(*a1), bt1(s1[0], s1[1], None), bt1(s1[2], s1[3], None)
        reps = []
        lim = threadpool_limits(limits=1) if threadpool_limits else contextlib.nullcontext()
        with lim:
            tstart = time.perf_counter()
            while len(reps) < 1000 and time.perf_counter() - tstart < min(budget_s, 20.0):
                t0 = time.perf_counter()
                selected_solve(A1, L1, G1, return_retarded=True, max_batch_size=1)
                reps.append(time.perf_counter() - t0)
        med = float(np.median(reps))
        small = {"value": round(1.0 / med, 3), "unit": "energy_points/s", "cores": 1,
                 "median_ms": round(med * 1e3, 4), "repeats": len(reps),
                 "sample": "8 blocks x 32, one energy, eta=1e-3, median over the repeats "
                           "(numpy restatement of rgf.py, 1 BLAS thread)"}
    except Exception as ex:  # pragma: no cover - reported, not fatal
        small = {"error": repr(ex)}
    cpu_model = None
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    blas = None
    if threadpool_limits:
        from threadpoolctl import threadpool_info

        blas = sorted({f"{i.get('internal_api')} {i.get('version')}" for i in threadpool_info()
                       if i.get("user_api") == "blas"})
    return {
        "value": n / tn,
        "unit": "energy_points/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} energies of the {g.n_blocks}x{g.block_sizes[0]} workload (numpy complex128 "
                  f"restatement of rgf.py, {cores} BLAS threads, {tn:.1f} s)",
        "single_core": single,
        "config_8x32_one_energy": small,
        "cpu_model": cpu_model,
        "blas": blas,
    }


def torch_rgf_fp64(Ad, Au, Al, SLd, SLu, SGd, SGu, GRd, GLd, GLu, GGd, GGu):
    """The reference selected solve (rgf.py:200-398, no OBC, no current) on torch tensors.

    All arguments are (E, n, b, b) complex128 device views (uniform block size);
    products go to rocBLAS batched ZGEMM and inverses to torch.linalg.inv, with
    the energies of the call as the batch.  Outputs: G^R diag, G^< and G^> diag
    (projected) and upper blocks.
    """
    import torch

    H = lambda x: x.conj().transpose(-2, -1)  # noqa: E731
    inv = torch.linalg.inv
    N = Ad.shape[1]
    xr, xl, xg = [None] * N, [None] * N, [None] * N
    xr[0] = inv(Ad[:, 0])
    xl[0] = xr[0] @ SLd[:, 0] @ H(xr[0])
    xg[0] = xr[0] @ SGd[:, 0] @ H(xr[0])
    for i in range(N - 1):
        j = i + 1
        a_ji = Al[:, i]
        a_ji_xr = a_ji @ xr[i]
        xr[j] = inv(Ad[:, j] - a_ji_xr @ Au[:, i])
        for x, s_d, s_u in ((xl, SLd, SLu), (xg, SGd, SGu)):
            t = a_ji_xr @ s_u[:, i]
            x[j] = xr[j] @ (s_d[:, j] + a_ji @ x[i] @ H(a_ji) + H(t) - t) @ H(xr[j])
    GLd[:, N - 1] = 0.5 * (xl[-1] - H(xl[-1]))
    GGd[:, N - 1] = 0.5 * (xg[-1] - H(xg[-1]))
    GRd[:, N - 1] = xr[-1]
    for i in range(N - 2, -1, -1):
        j = i + 1
        a_ij, a_ji = Au[:, i], Al[:, i]
        xr_ii_a_ij = xr[i] @ a_ij
        xr_jj_a_ji = xr[j] @ a_ji
        xr_ii_a_ij_xr_jj = xr_ii_a_ij @ xr[j]
        m4 = xr_ii_a_ij @ xr_jj_a_ji
        for x, s_u, gd, gu in ((xl, SLu, GLd, GLu), (xg, SGu, GGd, GGu)):
            t1 = m4 @ x[i] - xr[i] @ s_u[:, i] @ H(xr_ii_a_ij_xr_jj)
            t1 = t1 - H(t1)
            t2 = xr_ii_a_ij @ x[j]
            gu[:, i] = -t2 - x[i] @ H(xr_jj_a_ji) + xr[i] @ s_u[:, i] @ H(xr[j])
            x[i] = x[i] + t2 @ H(xr_ii_a_ij) + t1
            gd[:, i] = 0.5 * (x[i] - H(x[i]))
        xr[i] = xr[i] + m4 @ xr[i]
        GRd[:, i] = xr[i]


def torch_leg(args, model, energies, eps, dev) -> dict:
    """The reference algorithm on torch (rocBLAS batched ZGEMM + torch.linalg.inv).

    Runs one energy batch of the reference's default max_batch_size (100) on
    its own inputs (the first energies of the workload) and compares the
    result with this library's FP64 solve of the same inputs.
    """
    import torch

    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    ns = min(args.torch_sample, len(energies))
    nb, bs = args.n_blocks, args.block_size
    g = Geometry.uniform(nb, bs)
    A = W.system_matrix_torch(model, energies[:ns], args.eta, dev)
    sld = torch.as_tensor(model.sl_diag, device=dev).expand(ns, -1).contiguous()
    sgd = torch.as_tensor(model.sg_diag, device=dev).expand(ns, -1).contiguous()
    slu = torch.zeros((ns, g.offdiag_size), dtype=torch.complex128, device=dev)
    sgu = torch.zeros((ns, g.offdiag_size), dtype=torch.complex128, device=dev)
    GR = S.packed_empty(g, ns, dev)
    GL = S.packed_empty(g, ns, dev, parts=("d", "u"))
    GG = S.packed_empty(g, ns, dev, parts=("d", "u"))
    S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR)
    S.default_context(dev).status()
    ours = [X.clone() for X in (GR[0], GL[0], GG[0], GL[1], GG[1])]
    v4 = lambda t, n: t.view(ns, n, bs, bs)  # noqa: E731
    tviews = (v4(A[0], nb), v4(A[1], nb - 1), v4(A[2], nb - 1), v4(sld, nb), v4(slu, nb - 1),
              v4(sgd, nb), v4(sgu, nb - 1), v4(GR[0], nb), v4(GL[0], nb), v4(GL[1], nb - 1),
              v4(GG[0], nb), v4(GG[1], nb - 1))
    torch_rgf_fp64(*tviews)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    torch_rgf_fp64(*tviews)
    torch.cuda.synchronize()
    tt = time.perf_counter() - t0
    theirs = (GR[0], GL[0], GG[0], GL[1], GG[1])
    agree = max(float((torch.linalg.norm((X - Y).reshape(ns, -1), dim=-1)
                       / torch.linalg.norm(Y.reshape(ns, -1), dim=-1)).max())
                for X, Y in zip(theirs, ours))
    return {
        "value": round(ns / tt, 3),
        "unit": "energy_points/s",
        "definition": "reference algorithm (rgf.py selected_solve) on torch: rocBLAS batched "
                      "ZGEMM + torch.linalg.inv, one batch of 100 energies",
        "sample_energies": ns,
        "batched_fp64_speedup": round(eps / (ns / tt), 2),
        "max_rel_diff_vs_this_library": float(f"{agree:.3e}"),
    }


def eta_leg(args, model_seed, dev) -> dict:
    """Mixed precision at 48 x 256 over 1024 energies, eta cycled over
    {1e-2, 1e-3, 1e-4} (energy k takes the (k mod 3)-th value).

    In chunks that fit HBM: the FP64 solve (timed), the complex64-product
    solve with one FP64 refinement step and no check (timed), its error
    against the FP64 result per output (max over energies and blocks of the
    relative Frobenius error per block, per eta), the same with FP64 Schur
    complements instead of the refinement step (refine=-1, timed, errors per
    eta), and the mixed solve with the FP64 check, whose fallback fraction at
    1e-8 is reported per eta.
    """
    import torch

    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    g = Geometry.uniform(48, args.block_size)
    model = W.make_device(g, model_seed)
    ne = args.eta_energies
    etas = (1e-2, 1e-3, 1e-4)
    energies = W.draw_energies(ne)
    eta_of = np.array([etas[k % 3] for k in range(ne)])
    ctx = S.default_context(dev)
    errs = {eta: {"G^R": 0.0, "G^<": 0.0, "G^>": 0.0} for eta in etas}
    errs_s = {eta: {"G^R": 0.0, "G^<": 0.0, "G^>": 0.0} for eta in etas}
    fb = {eta: [] for eta in etas}
    res = {eta: [] for eta in etas}
    t64 = t32 = ts = 0.0
    for c0 in range(0, ne, args.eta_chunk):
        sl = slice(c0, min(ne, c0 + args.eta_chunk))
        n = sl.stop - sl.start
        A = W.system_matrix_torch(model, energies[sl], eta_of[sl], dev)
        sld = torch.as_tensor(model.sl_diag, device=dev).expand(n, -1).contiguous()
        sgd = torch.as_tensor(model.sg_diag, device=dev).expand(n, -1).contiguous()
        slu = torch.zeros((n, g.offdiag_size), dtype=torch.complex128, device=dev)
        sgu = torch.zeros_like(slu)
        outs = (S.packed_empty(g, n, dev), S.packed_empty(g, n, dev, parts=("d", "u")),
                S.packed_empty(g, n, dev, parts=("d", "u")))

        def solve(**kw):
            S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_retarded=outs[0], g_lesser=outs[1],
                             g_greater=outs[2], ctx=ctx, **kw)

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solve()
        torch.cuda.synchronize()
        t64 += time.perf_counter() - t0
        ctx.status()
        ref = [tuple(p.clone() for p in X if p is not None) for X in outs]

        def accumulate(into):
            for name, X, R in zip(("G^R", "G^<", "G^>"), outs, ref):
                worst = torch.zeros(n, dtype=torch.float64, device=dev)
                for part, rp in zip([p for p in X if p is not None], R):
                    d = (part - rp).reshape(n, -1, args.block_size ** 2)
                    r = rp.reshape(n, -1, args.block_size ** 2)
                    worst = torch.maximum(worst, (torch.linalg.norm(d, dim=-1)
                                                  / torch.linalg.norm(r, dim=-1)).amax(-1))
                w = worst.cpu().numpy()
                for eta in etas:
                    m = eta_of[sl] == eta
                    if m.any():
                        into[eta][name] = max(into[eta][name], float(w[m].max()))

        t0 = time.perf_counter()
        solve(precision="mixed", mixed=S.MixedOptions(refine=1, fallback_tol=0.0))
        torch.cuda.synchronize()
        t32 += time.perf_counter() - t0
        ctx.status()
        accumulate(errs)
        t0 = time.perf_counter()
        solve(precision="mixed", mixed=S.MixedOptions(refine=-1, fallback_tol=0.0))
        torch.cuda.synchronize()
        ts += time.perf_counter() - t0
        ctx.status()
        accumulate(errs_s)
        flags = torch.zeros(n, dtype=torch.int32, device=dev)
        resid = torch.zeros(n, dtype=torch.float64, device=dev)
        solve(precision="mixed", mixed=S.MixedOptions(refine=1, fallback_tol=1e-8, fallback_flags=flags,
                                                      residuals=resid))
        ctx.status()
        f, rr = flags.cpu().numpy(), resid.cpu().numpy()
        for eta in etas:
            m = eta_of[sl] == eta
            fb[eta].extend(f[m].tolist())
            res[eta].extend(rr[m].tolist())
        del A, sld, sgd, slu, sgu, outs, ref
        torch.cuda.empty_cache()
    f_e = flops_per_energy(48, args.block_size)
    out = {"workload": f"48 blocks x {args.block_size}, complex128, {ne} energies, eta cycled over "
                       f"{list(etas)}, chunks of {args.eta_chunk}",
           "fp64": {"energies_per_s": round(ne / t64, 2),
                    "frac_fp64_peak": round(f_e * ne / t64 / 1e12 / FP64_PEAK_TFLOPS, 4)},
           "mixed_unchecked": {"energies_per_s": round(ne / t32, 2), "speedup_vs_fp64": round(t64 / t32, 3)},
           "mixed_fp64_schur": {"energies_per_s": round(ne / ts, 2), "speedup_vs_fp64": round(t64 / ts, 3),
                                "max_rel_err_per_block": {f"eta={eta:g}": {k: float(f"{v:.3e}")
                                                                           for k, v in errs_s[eta].items()}
                                                          for eta in etas}}}
    for eta in etas:
        out[f"eta={eta:g}"] = {"energies": len(fb[eta]),
                               "max_rel_err_per_block": {k: float(f"{v:.3e}") for k, v in errs[eta].items()},
                               "fallback_fraction_at_1e-8": round(float(np.mean(fb[eta])), 4),
                               "median_residual": float(f"{float(np.median(res[eta])):.3e}")}
    return out


def uneven_leg(args, dev) -> dict:
    """Uneven batching: 2048 energies over 4 geometries, interleaved, FP64.

    Each geometry is resident (H, boundary and scattering self-energies); a
    slot per geometry assembles A on the device and replays a captured sweep
    graph per batch, with the GEMM tile per block size picked by the
    autotuner.  Batches of different geometries run concurrently on their own
    streams; outputs are overwritten batch by batch.
    """
    from quatrex_amd import serving as SV
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    shapes = ((16, 128), (32, 256), (24, 384), (12, 512))
    geos = [Geometry.uniform(nb, b) for nb, b in shapes]
    seeds = [W.DEVICE_SEED_BASE + 10 + k for k in range(len(geos))]
    per = args.uneven_energies // len(geos)
    srv = SV.Server(geos, seeds, budget_bytes=int(args.uneven_budget_gib * 2**30),
                    capacity=args.uneven_capacity, slots_per_geometry=1, precision="fp64",
                    use_graphs=True, autotune=True, device=dev)
    try:
        cap = min(s.capacity for s in srv.slots)
        # warm-up: one full batch per geometry captures the sweep graphs
        warm = [SV.Request(0.0, -1 - k, k % len(geos), -0.5 + k / (4.0 * cap), 1e-3)
                for k in range(len(geos) * cap)]
        srv.run(warm, realtime=False)
        rng = np.random.default_rng(W.ENERGY_SEED)
        en = rng.uniform(-1.0, 1.0, size=per * len(geos))
        reqs = [SV.Request(0.0, k, k % len(geos), float(en[k]), 1e-3) for k in range(len(en))]
        res = srv.run(reqs, realtime=False)
        st = SV.latency_stats(res)
    finally:
        srv.close()
    elapsed = max(r.done for r in res) - min(r.arrival for r in res)
    flops = sum(flops_per_energy(nb, b) for nb, b in shapes) * per
    return {
        "value": round(len(res) / elapsed, 3),
        "unit": "energy_points/s",
        "workload": f"{len(res)} energies, {per} per geometry (n_blocks, block) in {list(shapes)}, "
                    "interleaved, eta=1e-3, FP64, G^R/G^</G^> selected blocks",
        "achieved_tflops": round(flops / elapsed / 1e12, 3),
        "frac_fp64_peak": round(flops / elapsed / 1e12 / FP64_PEAK_TFLOPS, 4),
        "batch_capacity": sorted({s.capacity for s in srv.slots}),
        "batches": st["batches"],
        "tiles": srv.tiles,
        "elapsed_s": round(elapsed, 3),
    }


def _tridiagonal_coo(g, energies_vals):
    """Block-sorted (rows, cols) of the full tridiagonal pattern and the matching
    data of packed (diag, upper, lower) per energy."""
    bs = list(g.block_sizes)
    off = np.concatenate(([0], np.cumsum(bs)))
    rows, cols, pieces = [], [], []
    d, u, l = energies_vals
    for i in range(g.n_blocks):
        for j, part in ((i - 1, "l"), (i, "d"), (i + 1, "u")):
            if j < 0 or j >= g.n_blocks:
                continue
            r, c = np.meshgrid(np.arange(bs[i]), np.arange(bs[j]), indexing="ij")
            rows.append((off[i] + r).ravel())
            cols.append((off[j] + c).ravel())
            if part == "d":
                pieces.append(g.diag(d, i).reshape(d.shape[0], -1))
            elif part == "u":
                pieces.append(g.upper(u, i).reshape(u.shape[0], -1))
            else:
                pieces.append(g.lower(l, j).reshape(l.shape[0], -1))
    return np.concatenate(rows), np.concatenate(cols), pieces


def dsdb_leg(args, dev) -> dict:
    """Block-sparse stack <-> packed tridiagonal blocks (rgf_densify_z /
    rgf_sparsify_z) at the bench geometry with the full tridiagonal pattern."""
    import torch

    from quatrex_amd import dsdb as D
    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    g = Geometry.uniform(args.n_blocks, args.block_size)
    E = args.dsdb_energies
    model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
    A = W.system_matrix_torch(model, W.draw_energies(E), args.eta, dev)
    rows, cols, pieces = _tridiagonal_coo(g, A)
    data = torch.cat(pieces, dim=1).contiguous()
    del pieces
    a = D.DSDBCOO(data, rows, cols, g.block_sizes, (E,))
    ctx = S.default_context(dev)
    out = {"workload": f"{g.n_blocks} blocks x {args.block_size}, full tridiagonal pattern "
                       f"({a.nnz} stored entries per energy), {E} energies, complex128"}
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0.record()
            fn()
            t1.record()
            t1.synchronize()
            best = min(best, t0.elapsed_time(t1))
        return best * 1e-3

    parts = {}

    def dens():
        parts["p"], _ = D.densify_batch(a, slice(None), g, "dul", dev, ctx)

    t = timed(dens)
    moved = 2 * 16 * a.nnz * E  # read every stored entry, write every dense element (same count)
    out["densify"] = {"ms": round(t * 1e3, 3), "achieved_gbs": round(moved / t / 1e9, 1),
                      "frac_hbm_peak": round(moved / t / 1e9 / HBM_PEAK_GBS, 4)}
    x = D.DSDBCOO.zeros_like(a)
    bufs = parts["p"]

    def spar():
        D.sparsify_batch(x, slice(None), g, bufs, "dul", dev, ctx)

    t = timed(spar)
    out["sparsify"] = {"ms": round(t * 1e3, 3), "achieved_gbs": round(moved / t / 1e9, 1),
                       "frac_hbm_peak": round(moved / t / 1e9 / HBM_PEAK_GBS, 4)}
    out["round_trip_exact"] = bool(torch.equal(x.data, a.data))
    return out


def obc_leg(args, dev) -> dict:
    """Open boundary conditions of both contacts of the bench device for one
    energy batch: boundary system, Sancho-Rubio surface Green's function,
    refinement step with residuals, retarded/lesser/greater self-energies."""
    import torch

    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.boundary_conditions import OBCSystem, SanchoRubio, compute_contact_obc
    from quatrex_amd.layout import Geometry

    g = Geometry.uniform(args.n_blocks, args.block_size)
    E, b, n = args.obc_energies, args.block_size, args.n_blocks - 1
    model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
    z = torch.as_tensor(W.draw_energies(E) + 1j * args.eta, device=dev)
    eye = torch.eye(b, dtype=torch.complex128, device=dev)

    def blk(packed, view, i):
        return torch.as_tensor(view(packed, i).copy(), device=dev)

    # boundary blocks of A = (E + i eta) I - H (no boundary term)
    def diag(i):
        return (z[:, None, None] * eye - blk(model.h_diag, g.diag, i)).contiguous()

    def coupling(view, i):
        return (-blk(view[0], view[1], i)).expand(E, b, b).contiguous()

    up, lo = (model.h_upper, g.upper), (model.h_lower, g.lower)
    left = (coupling(lo, 0), diag(0), coupling(up, 0))
    right = (coupling(up, n - 1), diag(n), coupling(lo, n - 1))
    fl = torch.full((E,), 0.9, dtype=torch.float64, device=dev)
    fr = torch.full((E,), 0.1, dtype=torch.float64, device=dev)
    ctx = S.default_context(dev)
    out = {"workload": f"both contacts of {args.n_blocks} blocks x {b}, {E} energies, eta={args.eta}, "
                       "eta_obc=0, Sancho-Rubio(max_iterations=100, convergence_tol=1e-6)"}

    def run(osys):
        res = []
        for name, (m10, m00, m01), f in (("left", left, fl), ("right", right, fr)):
            res.append(compute_contact_obc(m10, m00, m01, f, 0.0, osys, contact=name))
        return res

    fresh = SanchoRubio(100, 1e-6)
    run(OBCSystem(fresh, memoization_mode="off"))  # warm-up
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    t0 = time.perf_counter()
    run(OBCSystem(fresh, memoization_mode="off"))
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    prof = ctx.profile_read(reset=True)
    ctx.profile(False)
    it = fresh.iterations
    # algorithmic work: per decimation step 6 products and 1 inverse (8 b^3 each),
    # plus the final inverse, the refinement step (2 products + 1 inverse) and
    # the self-energy (2 products), per contact and energy
    flops = 2 * E * 8.0 * b**3 * (7 * it + 1 + 3 + 2)
    out["sancho_rubio"] = {
        "ms": round(t * 1e3, 3), "energies_per_s": round(E / t, 1), "iterations_last_contact": it,
        "delta_last_contact": float(f"{fresh.delta:.3e}"),
        "achieved_tflops": round(flops / t / 1e12, 3),
        "frac_fp64_peak": round(flops / t / 1e12 / FP64_PEAK_TFLOPS, 4),
        "inverse_ms_share": round(prof["inv_ms"] / max(prof["inv_ms"] + prof["gemm_ms"], 1e-9), 4),
        "gemm_tflops": round(prof["gemm_flops"] / max(prof["gemm_ms"], 1e-9) / 1e9, 3),
    }
    # memoised: the second SCBA-like call reuses the cached surface function
    memo = OBCSystem(SanchoRubio(100, 1e-6), memoization_mode="force-after-first")
    run(memo)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = run(memo)
    torch.cuda.synchronize()
    tm = time.perf_counter() - t0
    out["memoized"] = {"ms": round(tm * 1e3, 3), "speedup_vs_fresh": round(t / tm, 2)}
    out["spectral"] = spectral_run(args, dev)
    out["assembly"] = assembly_run(args, dev)
    # size-independent self-check: the boundary self-energy of a converged
    # surface function has Gamma = i(S - S^H) Hermitian, and lesser + greater
    # = i (2f - 1) Gamma is anti-Hermitian
    sig = res[0][0]
    herm = float((torch.linalg.norm(res[0][1] + res[0][1].conj().transpose(-1, -2)) /
                  torch.linalg.norm(res[0][1])).item())
    out["lesser_antihermitian_rel"] = float(f"{herm:.3e}")
    out["sigma_norm"] = float(f"{float(torch.linalg.norm(sig[0])):.6e}")
    return out


def spectral_run(args, dev) -> dict:
    """The reference's default OBC (Spectral with Beyn's contour integral) on
    a lead at the bench's block size: a 1-D chain (on-site 0, hopping 1) cut
    into cells of b sites, energies inside its band, so each energy has two
    finite non-zero modes, both on the unit circle (up to eta), inside the
    annulus 0.5 < |w| < 2; the surface Green's function is checked against
    the Sancho-Rubio decimation of the same lead."""
    import torch

    from quatrex_amd import workloads as W
    from quatrex_amd.boundary_conditions import SanchoRubio, Spectral
    from quatrex_amd.nevp import Beyn

    E, b = args.obc_energies, args.block_size
    z = torch.as_tensor(1.5 * W.draw_energies(E) + 1j * args.eta, device=dev)
    eye = torch.eye(b, dtype=torch.complex128, device=dev)
    h00 = torch.diag(torch.ones(b - 1, dtype=torch.complex128, device=dev), 1)
    h00 = h00 + h00.T
    h01 = torch.zeros((b, b), dtype=torch.complex128, device=dev)
    h01[b - 1, 0] = 1.0
    a_ii = (z[:, None, None] * eye - h00).contiguous()
    a_ij = (-h01).expand(E, b, b).contiguous()
    a_ji = (-h01.T).expand(E, b, b).contiguous()
    r_o, r_i = 2.0, 0.5
    sp = Spectral(Beyn(r_o=r_o, r_i=r_i, m_0=4, num_quad_points=32), max_decay=1.5 * np.log(r_o))
    sp(a_ii, a_ij, a_ji, contact="left")  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = sp(a_ii, a_ij, a_ji, contact="left")
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ref = SanchoRubio(200, 1e-12)(a_ii, a_ij, a_ji, contact="left")
    rel = float((torch.linalg.norm((x - ref).reshape(E, -1), dim=1) /
                 torch.linalg.norm(ref.reshape(E, -1), dim=1)).max())
    return {"workload": f"left contact, 1-D chain in cells of {b} sites, {E} energies in (-1.5, 1.5), "
                        f"eta={args.eta}; Spectral(Beyn(r_o={r_o}, r_i={r_i}, m_0=4, "
                        "num_quad_points=32), num_ref_iterations=2)",
            "ms": round(t * 1e3, 3), "energies_per_s": round(E / t, 1),
            "reflected_modes_min_max": [int(sp.n_reflected.min()), int(sp.n_reflected.max())],
            "max_rel_diff_vs_sancho_rubio": float(f"{rel:.3e}")}


def assembly_run(args, dev, energies: int = 64) -> dict:
    """The electron solver's system matrix on the tridiagonal blocks
    (`solve.assemble_system`): A = z S - P - (Sigma^R_herm + (Sigma^> -
    Sigma^<)/2 + H) for every energy, with a shared H, overlap S and
    potential and per-energy self-energies; HBM-bound (three self-energy
    stacks read, A written: 4 x 16 B per stored element)."""
    import torch

    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    g = Geometry.uniform(args.n_blocks, args.block_size)
    E = energies
    model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=dev)  # noqa: E731
    h = (t(model.h_diag), t(model.h_upper), t(model.h_lower))
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)

    def stack():
        return tuple(torch.randn((E, n), dtype=torch.complex128, device=dev, generator=gen)
                     for n in (g.diag_size, g.offdiag_size, g.offdiag_size))

    sr, sl, sg = stack(), stack(), stack()
    ov = tuple(torch.zeros(n, dtype=torch.complex128, device=dev)
               for n in (g.diag_size, g.offdiag_size, g.offdiag_size))
    for i in range(g.n_blocks):
        g.diag(ov[0][None], i)[0].copy_(torch.eye(args.block_size, dtype=torch.complex128, device=dev))
    pot = torch.linspace(-0.1, 0.1, g.n, dtype=torch.float64, device=dev)
    z = torch.as_tensor(W.draw_energies(E) + 1j * args.eta, device=dev)
    A = tuple(torch.empty_like(x) for x in sr)
    ctx = S.default_context(dev)

    def run():
        S.assemble_system(g, z, h, a=A, overlap=ov, potential=pot, sigma_retarded_herm=sr,
                          sigma_lesser=sl, sigma_greater=sg, ctx=ctx)

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    moved = 4 * 16 * E * (g.diag_size + 2 * g.offdiag_size)
    # size-independent check of one block of one energy against the formula
    e, i = E // 3, args.n_blocks // 2
    ref = (z[e] * g.diag(ov[0][None], i)[0] - torch.diag(pot[g.block_offsets[i]:g.block_offsets[i + 1]]
                                                          .to(torch.complex128))
           - (g.diag(sr[0], i)[e] + 0.5 * (g.diag(sg[0], i)[e] - g.diag(sl[0], i)[e]) + g.diag(h[0][None], i)[0]))
    err = float((torch.linalg.norm(g.diag(A[0], i)[e] - ref) / torch.linalg.norm(ref)).item())
    return {"workload": f"{args.n_blocks} blocks x {args.block_size}, {E} energies, shared H, overlap and "
                        "potential, per-energy Sigma^R_herm, Sigma^<, Sigma^> (all tridiagonal blocks)",
            "ms": round(el * 1e3, 3), "achieved_gbs": round(moved / el / 1e9, 1),
            "frac_hbm_peak": round(moved / el / 1e9 / HBM_PEAK_GBS, 4),
            "rel_err_one_block": float(f"{err:.3e}")}


def section_leg(args, dev) -> dict:
    """The per-section work of a block-partitioned (multi-GPU) solve at its
    real shape: the FP64 sweep over n_blocks / sections blocks of the same
    size and the same energies, as one GPU of the partition would run it."""
    import torch

    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    nb = args.n_blocks // args.sections
    g = Geometry.uniform(nb, args.block_size)
    E = args.energies
    model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
    A = W.system_matrix_torch(model, W.draw_energies(E), args.eta, dev)
    sld = torch.as_tensor(model.sl_diag, device=dev).expand(E, -1).contiguous()
    sgd = torch.as_tensor(model.sg_diag, device=dev).expand(E, -1).contiguous()
    slu = torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=dev)
    sgu = torch.zeros_like(slu)
    GR = S.packed_empty(g, E, dev)
    GL = S.packed_empty(g, E, dev, parts=("d", "u"))
    GG = S.packed_empty(g, E, dev, parts=("d", "u"))
    ctx = S.default_context(dev)

    def step():
        S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                         ctx=ctx)

    step()
    ctx.status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    ctx.status()
    f = flops_per_energy(nb, args.block_size)
    return {"workload": f"one of {args.sections} sections: {nb} blocks x {args.block_size}, {E} energies, "
                        "FP64 selected solve",
            "ms": round(t * 1e3, 3), "energies_per_s": round(E / t, 2),
            "achieved_tflops": round(f * E / t / 1e12, 3),
            "frac_fp64_peak": round(f * E / t / 1e12 / FP64_PEAK_TFLOPS, 4),
            "ms_per_block_step": round(t * 1e3 / nb, 3)}


def small_block_leg(args, dev) -> dict:
    """Small blocks, where a complex product's intensity (about b/6 flop per
    byte with its operands in HBM) is below the FP64 ridge (78.6 / 8 = 9.8):
    the FP64 selected solve over 8 x 32 (the one-energy CPU configuration's
    shape, here with many energies), 32 x 32 and 32 x 64.  Energies/s from
    plain steps; GEMM time, flops and algorithmic bytes (each operand of each
    problem read once, each output written once) from one step with the
    library's HIP-event profiling on."""
    import torch

    from quatrex_amd import autotune
    from quatrex_amd import solve as S
    from quatrex_amd import workloads as W
    from quatrex_amd.layout import Geometry

    out = {"definition": "GEMM achieved_gbs = algorithmic operand and output bytes of the sweep "
                         "products / their HIP-event time; intensity = their flops / those bytes"}
    ctx = S.default_context(dev)
    for nb, b in ((8, 32), (32, 32), (32, 64)):
        E = args.small_energies if b <= 32 else max(1, args.small_energies // 4)
        g = Geometry.uniform(nb, b)
        model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
        A = W.system_matrix_torch(model, W.draw_energies(E), args.eta, dev)
        sld = torch.as_tensor(model.sl_diag, device=dev).expand(E, -1).contiguous()
        sgd = torch.as_tensor(model.sg_diag, device=dev).expand(E, -1).contiguous()
        slu = torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=dev)
        sgu = torch.zeros_like(slu)
        GR = S.packed_empty(g, E, dev)
        GL = S.packed_empty(g, E, dev, parts=("d", "u"))
        GG = S.packed_empty(g, E, dev, parts=("d", "u"))
        tile = autotune.tune(ctx, b, n_energies=min(E, 1024))["best_name"]

        def step():
            S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG,
                             g_retarded=GR, ctx=ctx)

        step()
        ctx.status()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        ctx.status()
        ctx.profile(True)
        ctx.profile_read(reset=True)
        step()
        prof = ctx.profile_read(reset=True)
        ctx.profile(False)
        ctx.status()
        f = flops_per_energy(nb, b)
        gms = max(prof["gemm_ms"], 1e-9)
        gbs = prof["gemm_bytes"] / gms / 1e6
        out[f"{nb}x{b}"] = {
            "energies": E, "tile": tile,
            "energies_per_s": round(E / t, 1),
            "sweep_tflops": round(f * E / t / 1e12, 3),
            "sweep_frac_fp64_peak": round(f * E / t / 1e12 / FP64_PEAK_TFLOPS, 4),
            "gemm": {"launches": prof["gemm_launches"],
                     "achieved_tflops": round(prof["gemm_flops"] / gms / 1e9, 3),
                     "achieved_gbs": round(gbs, 1),
                     "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                     "intensity_flop_per_byte": round(prof["gemm_flops"] / max(prof["gemm_bytes"], 1), 2)},
            "inverse_ms_share": round(prof["inv_ms"] / max(prof["inv_ms"] + prof["gemm_ms"], 1e-9), 4),
        }
        if b <= 32:
            # the same solve on the launch-per-product-group sweep, for comparison
            out[f"{nb}x{b}"]["path"] = "persistent on-chip sweep"
            pctx = S.Context(dev)
            pctx.set_option(_lib_opt("RGF_OPT_PERSISTENT"), 0)
            autotune.tune(pctx, b, n_energies=min(E, 1024))

            def step2():
                S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG,
                                 g_retarded=GR, ctx=pctx)

            step2()
            pctx.status()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step2()
            torch.cuda.synchronize()
            t2 = (time.perf_counter() - t0) / args.steps
            pctx.status()
            out[f"{nb}x{b}"]["launch_per_product_energies_per_s"] = round(E / t2, 1)
            pctx.close()
        del A, sld, sgd, slu, sgu, GR, GL, GG
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return out


def _lib_opt(name: str) -> int:
    from quatrex_amd import _lib

    return getattr(_lib, name)


def _host_tridiagonal_dsdb(g, packed, rows_cols=None):
    """A CPU block-sparse stack (DSDBCOO, torch data) of the full tridiagonal
    pattern holding packed (diag, upper, lower) numpy stacks."""
    import torch

    from quatrex_amd import dsdb as D

    d, u, l = packed
    E = d.shape[0]
    bs = list(g.block_sizes)
    if rows_cols is None:
        rows, cols, _ = _tridiagonal_coo(g, (d[:1], u[:1], l[:1]))
    else:
        rows, cols = rows_cols
    data = torch.empty((E, rows.size), dtype=torch.complex128)
    k = 0
    for i in range(g.n_blocks):
        for j, part in ((i - 1, "l"), (i, "d"), (i + 1, "u")):
            if j < 0 or j >= g.n_blocks:
                continue
            blk = g.diag(d, i) if part == "d" else (g.upper(u, i) if part == "u" else g.lower(l, j))
            n = bs[i] * bs[j]
            data[:, k:k + n] = torch.from_numpy(np.ascontiguousarray(blk).reshape(E, n))
            k += n
    return D.DSDBCOO(data, rows, cols, g.block_sizes, (E,)), (rows, cols)


def host_io_leg(args, dev) -> dict:
    """The drop-in RGF on host-resident (CPU torch) stacks, block-tridiagonal
    and block-sparse (full tridiagonal pattern): energy chunks staged through
    pinned memory, moved over PCIe and solved with the copies of both
    directions and the solves overlapping - the PCIe-inclusive rate of a
    caller that hands over host buffers (the device-resident rate is `value`)."""
    import torch

    from quatrex_amd import workloads as W
    from quatrex_amd.datastructures import BlockTridiagonal
    from quatrex_amd.greens_function_solver import RGF
    from quatrex_amd.layout import Geometry

    g = Geometry.uniform(args.n_blocks, args.block_size)
    E = args.host_energies
    model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
    ad, au, al = W.system_matrix(model, W.draw_energies(E), args.eta)
    sld, slu, sgd, sgu = W.sigmas(model, E)
    t = torch.from_numpy
    bs = list(g.block_sizes)
    A = BlockTridiagonal.from_packed(bs, t(ad), t(au), t(al))
    # the workload's Sigma^<> have no off-diagonal blocks
    SL = BlockTridiagonal.from_packed(bs, t(sld), t(slu), torch.zeros_like(t(slu)))
    SG = BlockTridiagonal.from_packed(bs, t(sgd), t(sgu), torch.zeros_like(t(sgu)))
    outs = tuple(BlockTridiagonal(bs, (E,)) for _ in range(3))
    solver = RGF(max_batch_size=E)
    solver.selected_solve(A, SL, SG, out=outs, return_retarded=True)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.selected_solve(A, SL, SG, out=outs, return_retarded=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    full = g.diag_size + 2 * g.offdiag_size  # all three parts
    # in: A (3 parts), Sigma^<> (diag and upper: the lower parts are not
    # read); out: G^R, G^<, G^> (3 parts each)
    moved = 16 * E * (full + 2 * (g.diag_size + g.offdiag_size) + 3 * full)
    res = {"workload": f"{g.n_blocks} blocks x {args.block_size}, {E} energies, complex128, CPU-resident "
                       "A, Sigma^<, Sigma^> and outputs (G^R, G^<, G^> all tridiagonal blocks)",
           "energies_per_s": round(E / el, 2), "ms": round(el * 1e3, 1),
           "host_bytes_in_and_out": moved,
           "effective_gbs": round(moved / el / 1e9, 2)}
    nl = g.block_sizes[-1] ** 2
    ref = outs[1].packed()[0][-1, -nl:].clone()  # G^< of the last energy's last diagonal block
    del A, SL, SG, outs, solver
    # the same solve on block-sparse stacks (DSDBCOO, CPU torch data): the
    # stored entries move, (de)densified on the device per chunk
    a, rc = _host_tridiagonal_dsdb(g, (ad, au, al))
    del ad, au, al
    z = np.zeros_like(slu)
    sl, _ = _host_tridiagonal_dsdb(g, (sld, slu, z), rc)
    sg, _ = _host_tridiagonal_dsdb(g, (sgd, sgu, z), rc)
    del sld, slu, sgd, sgu, z
    xs = [type(a).zeros_like(a) for _ in range(3)]
    solver = RGF(max_batch_size=E)
    solver.selected_solve(a, sl, sg, out=xs, return_retarded=True)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.selected_solve(a, sl, sg, out=xs, return_retarded=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the last stored entries are the last diagonal block, row by row
    moved_sp = 16 * E * 6 * int(a.nnz)  # every stored entry of the three inputs and three outputs
    res["block_sparse"] = {"energies_per_s": round(E / el, 2), "ms": round(el * 1e3, 1),
                           "stored_entries_per_energy": int(a.nnz),
                           "host_bytes_in_and_out": moved_sp,
                           "effective_gbs": round(moved_sp / el / 1e9, 2),
                           "equal_to_block_tridiagonal_run": bool(torch.equal(xs[1].data[-1, -nl:], ref))}
    return res


def serving_leg(args, dev) -> dict:
    """Continuous batching over 6 resident geometries, Poisson arrivals, FP64 (the mixed path
    cannot meet 1e-8 on these workloads, so FP64 is the serving precision)."""
    from quatrex_amd import serving as SV
    from quatrex_amd import workloads as W

    geos = SV.draw_geometries(6, W.GEOMETRY_SEED)
    seeds = [W.GEOMETRY_SEED + 1 + k for k in range(len(geos))]
    srv = SV.Server(geos, seeds, budget_bytes=int(args.serving_budget_gib * 2**30), capacity=8,
                    slots_per_geometry=2, precision="fp64", fallback_tol=1e-8, refine=1,
                    autotune=True, device=dev)
    try:
        # warm every slot (graph capture, first launches) outside the measured stream
        warm = [SV.Request(0.0, -1 - k, k % len(geos), 0.1 * k, 1e-3) for k in range(4 * len(geos))]
        srv.run(warm, realtime=False)
        reqs = SV.poisson_requests(args.serving_requests, args.serving_rate, len(geos), eta=1e-3)
        res = srv.run(reqs, realtime=True)
        st = SV.latency_stats(res)
        # the same server below capacity: 60% of the throughput just measured
        rate2 = 0.6 * st["throughput_rps"]
        n2 = int(min(600, max(100, rate2 * 8)))
        res2 = srv.run(SV.poisson_requests(n2, rate2, len(geos), seed=W.ARRIVAL_SEED + 1, eta=1e-3),
                       realtime=True)
        st2 = SV.latency_stats(res2)
    finally:
        srv.close()
    cost = {f"{g.n_blocks}x{g.max_block}": round(flops_per_energy(g.n_blocks, g.max_block) / 1e12, 4)
            for g in geos}
    return {
        "workload": f"{len(geos)} resident geometries (n_blocks x block: TFLOP per energy) {cost}, "
                    f"Poisson arrivals at {args.serving_rate:g}/s, first {args.serving_requests} "
                    f"requests of the stream, one energy point each, eta=1e-3",
        "offered_rps": args.serving_rate,
        "precision": "fp64",
        "requests_served": args.serving_requests,
        "requests_in_config": 20000,
        "hbm_budget_gib": args.serving_budget_gib,
        "slots": len(srv.slots),
        "slot_capacity": sorted({s.capacity for s in srv.slots}),
        "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "4 (HIP default)"),
        "tiles": srv.tiles,
        **{k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()},
        "below_capacity": {"offered_rps": round(rate2, 2),
                           **{k: (round(v, 4) if isinstance(v, float) else v) for k, v in st2.items()}},
    }


LEG_ARGS = ("n_blocks", "block_size", "energies", "eta", "torch_sample", "eta_energies", "eta_chunk",
            "dsdb_energies",
            "obc_energies", "sections", "steps", "host_energies", "small_energies",
            "uneven_energies", "uneven_capacity", "uneven_budget_gib", "serving_requests",
            "serving_rate", "serving_budget_gib")


def _progress(msg: str) -> None:
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_leg(leg: str, args, eps: float) -> dict:
    """Run one secondary leg in a child process; its JSON line, or the error."""
    import subprocess

    cmd = [sys.executable, "-u", str(Path(__file__).resolve()), "--only-leg", leg,
           "--main-eps", repr(eps)]
    for name in LEG_ARGS:
        cmd += ["--" + name.replace("_", "-"), str(getattr(args, name))]
    env = dict(os.environ)
    if leg in ("serving", "uneven"):
        # a hardware queue per slot stream (12 and 4 slots); HIP reads this at start-up
        env["GPU_MAX_HW_QUEUES"] = "16"
    try:
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=args.leg_timeout, env=env)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {args.leg_timeout} s"}
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("{")]
    if not lines:
        tail = (res.stderr or "").strip().splitlines()[-12:]
        return {"error": f"exit code {res.returncode}", "stderr_tail": tail}
    out = json.loads(lines[-1])
    if res.returncode != 0:  # the leg finished, the process did not end cleanly
        out["process_exit_code"] = res.returncode
    return out


def leg_main(args) -> None:
    """Child-process entry of one secondary leg: prints its JSON line."""
    import faulthandler

    import torch

    faulthandler.enable(all_threads=True)  # a native crash prints every thread's stack

    from quatrex_amd import workloads as W

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    seed = W.DEVICE_SEED_BASE + 2
    if args.only_leg == "torch":
        # torch's default batched inverse on ROCm (hipBLAS getrfBatched) fails with
        # an allocation error inside its fixed rocBLAS workspace; MAGMA is used
        torch.backends.cuda.preferred_linalg_library("magma")
        from quatrex_amd.layout import Geometry

        model = W.make_device(Geometry.uniform(args.n_blocks, args.block_size), seed)
        res = torch_leg(args, model, W.draw_energies(args.energies), args.main_eps, dev)
        res["linalg_backend"] = "magma"
    elif args.only_leg == "eta":
        res = eta_leg(args, seed, dev)
    elif args.only_leg == "uneven":
        res = uneven_leg(args, dev)
    elif args.only_leg == "dsdb":
        res = dsdb_leg(args, dev)
    elif args.only_leg == "obc":
        res = obc_leg(args, dev)
    elif args.only_leg == "section":
        res = section_leg(args, dev)
    elif args.only_leg == "host":
        res = host_io_leg(args, dev)
    elif args.only_leg == "small":
        res = small_block_leg(args, dev)
    else:
        res = serving_leg(args, dev)
    print(json.dumps(res), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.only_leg:
        leg_main(args)
        return

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
    # GEMM tile per block size from the autotuner (timed on this device, cached)
    from quatrex_amd import _lib
    from quatrex_amd import autotune as AT

    tiles = {"f64": "64x64-3m" if args.block_size >= 64 else "32x32-3m",
             "f32": "64x64" if args.block_size >= 64 else "32x32"}
    if args.tile:
        tiles["f64"] = args.tile
        ctx.set_option(_lib.RGF_OPT_TILE_FOR_BLOCK + args.block_size, _lib.TILE_CFG_NAMES.index(args.tile))
    elif not args.no_autotune:
        tiles["f64"] = AT.tune(ctx, args.block_size, n_energies=args.energies)["best_name"]
    if not args.no_mixed and not args.no_autotune:
        tiles["f32"] = AT.tune(ctx, args.block_size, n_energies=args.energies, f32=True)["best_name"]
    tile_m, tile_n = tiles["f64"].removesuffix("-3m").split("x")
    tile_3m = tiles["f64"].endswith("-3m")

    def step():
        S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                         ctx=ctx)

    for _ in range(args.warmup):
        step()
    ctx.status()
    torch.cuda.synchronize()
    _progress(f"main leg: {args.warmup} warm-up steps done, timing {args.steps}")
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
    achieved = gemm_flops_launch / (gemm_avg_ms * 1e-3) / 1e12 if gemm_avg_ms > 0 else 0.0
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
        # the two ways of keeping the inversion input accurate, unchecked: one
        # FP64 Newton step per inverted block (refine=1), or the Schur
        # complements themselves in FP64 (refine=-1)
        refined = {}
        for name, r in (("newton_step", 1), ("fp64_schur", -1)):
            mo = S.MixedOptions(refine=r, fallback_tol=0.0)
            S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                             ctx=ctx, precision="mixed", mixed=mo)
            ctx.status()
            torch.cuda.synchronize()
            tr0 = time.perf_counter()
            S.selected_solve(g, A, (sld, slu), (sgd, sgu), g_lesser=GL, g_greater=GG, g_retarded=GR,
                             ctx=ctx, precision="mixed", mixed=mo)
            torch.cuda.synchronize()
            tr = time.perf_counter() - tr0
            ctx.status()
            refined[name] = {"refine": r, "value": round(E / tr, 3),
                             "max_rel_err_vs_fp64": {k: float(f"{rel_err(X, R):.3e}") for k, X, R in
                                                     zip(("G^R", "G^<", "G^>"), (GR, GL, GG), ref)}}
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
            "refined": refined,
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
        pctx.set_option(6, 0)

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
            "gemm_tiles": tiles,
            "serving_requests": f"{args.serving_requests} of the 20000-request Poisson stream "
                                f"(serving leg; 0 = leg off)",
        },
        "roofline": {
            "bound": "mfma",
            "kernel": f"zgemm_grouped_kernel<{tile_m}, {tile_n}, 0, false, {str(tile_3m).lower()}> "
                      f"(sweep products)",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "avg_launch_ms": round(gemm_avg_ms, 4),
            "flops_per_launch": gemm_flops_launch,
            "launches": prof["gemm_launches"],
            # 3M tiles execute 6 of the 8 m n k real flops of a complex product
            # on the matrix cores: their busy fraction is 3/4 of frac
            "mfma_executed_frac": round(achieved / FP64_PEAK_TFLOPS * (0.75 if tile_3m else 1.0), 4),
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

    del A, sld, sgd, slu, sgu, GR, GL, GG
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    # The secondary legs run one child process each: a failure there (a third
    # party library included) is reported in the JSON instead of ending the run.
    legs = []
    if args.torch_sample > 0:
        legs.append(("torch_fp64", "torch"))
    if args.eta_energies > 0 and not args.no_mixed:
        legs.append(("mixed_error_vs_eta", "eta"))
    if args.dsdb_energies > 0:
        legs.append(("block_sparse_io", "dsdb"))
    if args.obc_energies > 0:
        legs.append(("open_boundary", "obc"))
    if args.sections > 1 and args.n_blocks // args.sections >= 2:
        legs.append(("partition_section", "section"))
    if args.host_energies > 0:
        legs.append(("host_resident_dropin", "host"))
    if args.small_energies > 0:
        legs.append(("small_blocks", "small"))
    if args.uneven_energies > 0:
        legs.append(("uneven_batching", "uneven"))
    if args.serving_requests > 0:
        legs.append(("serving", "serving"))
    for key, leg in legs:
        _progress(f"leg {key} ...")
        t_leg = time.perf_counter()
        out[key] = run_leg(leg, args, eps)
        _progress(f"leg {key} done in {time.perf_counter() - t_leg:.1f} s")

    # speed-ups against the reference algorithm on vendor libraries (rocBLAS
    # ZGEMM + torch.linalg.inv at the reference's default batch of 100)
    tv = out.get("torch_fp64", {}).get("value") if isinstance(out.get("torch_fp64"), dict) else None
    if tv:
        out["speedup_vs_torch_fp64"] = round(eps / tv, 3)
        if "mixed" in out:
            out["mixed"]["speedup_vs_torch_fp64"] = round(out["mixed"]["value"] / tv, 3)
            chk = out["mixed"]["with_check_and_fp64_fallback"]
            chk["speedup_vs_torch_fp64"] = round(chk["value"] / tv, 3)
    if not args.no_cpu_baseline:
        _progress("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(model, energies, args.eta, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
