"""Mixed-precision selected solve (complex64 products on FP32 matrix cores,
FP64 inversions, FP64 a-posteriori check with FP64 fallback)."""

import numpy as np
import pytest

from oracle.dense import dense_solve
from quatrex_amd import _lib
from quatrex_amd import solve as S
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry
from tests.helpers import block_rel_errors, edge_case, tridiagonal_mask

pytestmark = pytest.mark.gpu


def _case(device, nb, b, ne, eta, seed):
    import torch

    g = Geometry.uniform(nb, b)
    model = W.make_device(g, seed)
    en = W.draw_energies(ne)
    A = W.system_matrix(model, en, eta)
    sld, slu, sgd, sgu = W.sigmas(model, ne)
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=device)  # noqa: E731
    return g, A, (sld, slu), (sgd, sgu), t


def _run(device, g, A, SL, SG, t, precision, mixed=None):
    E = A[0].shape[0]
    GR, GL, GG = (S.packed_empty(g, E, device) for _ in range(3))
    S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)), g_lesser=GL,
                     g_greater=GG, g_retarded=GR, precision=precision, mixed=mixed)
    S.default_context(device).status()
    return [tuple(x.cpu().numpy() for x in X) for X in (GR, GL, GG)]


def _errors(g, outs, A, SL, SG):
    E = A[0].shape[0]
    z = np.zeros_like(SL[1])
    ref = dense_solve(g.to_dense(*A), g.to_dense(SL[0], SL[1], z), g.to_dense(SG[0], SG[1], z))
    bs = list(g.block_sizes)
    m = tridiagonal_mask(bs)
    return [block_rel_errors(g.to_dense(*X), r * m, bs) for X, r in zip(outs, ref)]


def _cpu_residual(g, A, SL, SG, outs):
    """Same identities as the library's check, evaluated in numpy (block 0, N/2, N-1)."""
    GR, GL, GG = outs
    N = g.n_blocks
    E = A[0].shape[0]
    worst = np.zeros(E)
    for i in sorted({0, N // 2, N - 1}):
        b = g.block_sizes[i]
        D = lambda x, k: g.diag(x, k)  # noqa: E731
        U = lambda x, k: g.upper(x, k)  # noqa: E731
        L = lambda x, k: g.lower(x, k)  # noqa: E731
        H = lambda x: np.conj(np.swapaxes(x, -1, -2))  # noqa: E731
        rr = D(A[0], i) @ D(GR[0], i) - np.eye(b)
        if i > 0:
            rr = rr + L(A[2], i - 1) @ U(GR[1], i - 1)
        if i < N - 1:
            rr = rr + U(A[1], i) @ L(GR[2], i)
        res = [np.linalg.norm(rr, axis=(-2, -1)) / np.sqrt(b)]
        for X, Sg in ((GL, SL), (GG, SG)):
            r = D(A[0], i) @ D(X[0], i) - D(Sg[0], i) @ H(D(GR[0], i))
            if i > 0:
                r = r + L(A[2], i - 1) @ U(X[1], i - 1) + H(U(Sg[1], i - 1)) @ H(L(GR[2], i - 1))
            if i < N - 1:
                r = r - U(A[1], i) @ H(U(X[1], i)) - U(Sg[1], i) @ H(U(GR[1], i))
            scale = (np.linalg.norm(D(A[0], i), axis=(-2, -1)) * np.linalg.norm(D(X[0], i), axis=(-2, -1))
                     + np.linalg.norm(D(Sg[0], i), axis=(-2, -1)) * np.linalg.norm(D(GR[0], i), axis=(-2, -1)))
            res.append(np.linalg.norm(r, axis=(-2, -1)) / scale)
        worst = np.maximum(worst, np.max(res, axis=0))
    return worst


@pytest.mark.parametrize("eta", [1e-2, 1e-3])
def test_mixed_error_is_fp32_level(device, eta):
    g, A, SL, SG, t = _case(device, 6, 64, 3, eta, 3001)
    outs = _run(device, g, A, SL, SG, t, "mixed", S.MixedOptions(refine=1, fallback_tol=0.0))
    errs = _errors(g, outs, A, SL, SG)
    assert max(errs) < 1e-2, errs
    assert max(errs) > 1e-12, errs  # it really ran in reduced precision


@pytest.mark.parametrize("cfg", [0, 1, 3])
def test_mixed_every_f32_tile_config(device, cfg):
    """Each tile configuration built for the FP32-compute products (64x64,
    32x32, 128x64) gives FP32-level results at b = 200 (partial tiles)."""
    g, A, SL, SG, t = _case(device, 4, 200, 2, 1e-2, 3004)
    ctx = S.default_context(device)
    ctx.set_option(_lib.RGF_OPT_TILE_CFG, cfg)
    try:
        outs = _run(device, g, A, SL, SG, t, "mixed", S.MixedOptions(refine=1, fallback_tol=0.0))
    finally:
        ctx.set_option(_lib.RGF_OPT_TILE_CFG, -1)
    errs = _errors(g, outs, A, SL, SG)
    assert max(errs) < 1e-3, errs
    assert max(errs) > 1e-12, errs


@pytest.mark.parametrize("bs", [[40], [1] * 8])
def test_mixed_edge_geometries(device, bs):
    """The mixed-precision sweep on a single block and on 1 x 1 blocks, with
    both contacts: FP32-level agreement with the dense FP64 oracle; with the
    FP64 check at 1e-8, every energy either fell back (then <= 1e-10 per
    block) or reports a residual <= 1e-8."""
    import torch

    from oracle.dense import dense_solve

    g, A, SL, SG, obc = edge_case(bs, 600 + len(bs))
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=device)  # noqa: E731
    obc_t = [[t(x) if x is not None else None for x in lst] for lst in obc]
    E = A[0].shape[0]

    def lower(u):
        out = np.empty_like(u)
        for i in range(g.n_blocks - 1):
            g.lower(out, i)[...] = -np.conj(np.swapaxes(g.upper(u, i), -1, -2))
        return out

    ref = dense_solve(g.to_dense(*A), g.to_dense(SL[0], SL[1], lower(SL[1])),
                      g.to_dense(SG[0], SG[1], lower(SG[1])), g.block_offsets, obc)
    m = tridiagonal_mask(list(g.block_sizes))
    bs_ = list(g.block_sizes)
    for tol in (0.0, 1e-8):
        GR, GL, GG = (S.packed_empty(g, E, device) for _ in range(3))
        res = torch.zeros(E, dtype=torch.float64, device=device)
        flags = torch.zeros(E, dtype=torch.int32, device=device)
        S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)), g_lesser=GL,
                         g_greater=GG, g_retarded=GR, obc=obc_t, precision="mixed",
                         mixed=S.MixedOptions(refine=1, fallback_tol=tol, residuals=res,
                                              fallback_flags=flags))
        S.default_context(device).status()
        outs = [g.to_dense(*(x.cpu().numpy() for x in X)) for X in (GR, GL, GG)]
        if tol == 0.0:
            errs = [block_rel_errors(o, r * m, bs_) for o, r in zip(outs, ref)]
            assert max(errs) < 1e-3, errs
            continue
        fl, rs = flags.cpu().numpy(), res.cpu().numpy()
        for e in range(E):
            errs = [block_rel_errors(o[e:e + 1], (r * m)[e:e + 1], bs_) for o, r in zip(outs, ref)]
            if fl[e]:
                assert max(errs) <= 1e-10, (e, errs)
            else:
                assert rs[e] <= 1e-8 and max(errs) < 1e-3, (e, rs[e], errs)


@pytest.mark.parametrize("eta", [1e-2, 1e-3])
def test_mixed_fp64_schur_mode(device, eta):
    """refine = -1 computes each inversion's input (the Schur complement and
    its factor A_ji g_i) in FP64 instead of taking Newton steps: G^R at least
    as close to the dense FP64 oracle as with one Newton step (the steps'
    fixed point), G^< and G^> at FP32 level, and still reduced precision."""
    g, A, SL, SG, t = _case(device, 6, 64, 3, eta, 3005)
    e1 = _errors(g, _run(device, g, A, SL, SG, t, "mixed", S.MixedOptions(refine=1, fallback_tol=0.0)),
                 A, SL, SG)
    es = _errors(g, _run(device, g, A, SL, SG, t, "mixed", S.MixedOptions(refine=-1, fallback_tol=0.0)),
                 A, SL, SG)
    assert max(es) < 1e-2 and max(es) > 1e-12, es
    assert es[0] <= 2.0 * e1[0] + 1e-13, (es, e1)
    with pytest.raises(ValueError):
        _run(device, g, A, SL, SG, t, "mixed", S.MixedOptions(refine=-2, fallback_tol=0.0))


def test_mixed_residual_matches_cpu_and_flags_fallback(device):
    import torch

    g, A, SL, SG, t = _case(device, 8, 48, 4, 1e-3, 3002)
    E = 4
    res = torch.zeros(E, dtype=torch.float64, device=device)
    flags = torch.zeros(E, dtype=torch.int32, device=device)
    # no fallback: residuals of the mixed result, checked against numpy
    outs = _run(device, g, A, SL, SG, t, "mixed",
                S.MixedOptions(refine=1, fallback_tol=0.0, residuals=res, fallback_flags=flags))
    cpu = _cpu_residual(g, A, SL, SG, outs)
    gpu = res.cpu().numpy()
    assert np.allclose(gpu, cpu, rtol=1e-6, atol=1e-15), (gpu, cpu)
    assert np.all(gpu > 1e-9)
    assert flags.sum().item() == 0
    # exact (FP64) results satisfy the identities to rounding
    exact = _run(device, g, A, SL, SG, t, "fp64")
    assert np.all(_cpu_residual(g, A, SL, SG, exact) < 1e-12)
    # with the 1e-8 tolerance every energy falls back to FP64
    outs = _run(device, g, A, SL, SG, t, "mixed",
                S.MixedOptions(refine=1, fallback_tol=1e-8, residuals=res, fallback_flags=flags))
    assert flags.cpu().numpy().tolist() == [1] * E
    assert max(_errors(g, outs, A, SL, SG)) <= 1e-10


def test_mixed_single_energy_fallback_path(device):
    """Few failing energies take the per-energy FP64 path; the rest stay mixed."""
    import torch

    g, A, SL, SG, t = _case(device, 5, 32, 8, 1e-2, 3003)
    E = 8
    res = torch.zeros(E, dtype=torch.float64, device=device)
    flags = torch.zeros(E, dtype=torch.int32, device=device)
    outs = _run(device, g, A, SL, SG, t, "mixed",
                S.MixedOptions(refine=0, fallback_tol=0.0, residuals=res, fallback_flags=flags))
    r = res.cpu().numpy()
    tol = 0.5 * float(np.sort(r)[-2] + np.sort(r)[-1])  # only the worst energy exceeds it
    outs = _run(device, g, A, SL, SG, t, "mixed",
                S.MixedOptions(refine=0, fallback_tol=tol, residuals=res, fallback_flags=flags))
    f = flags.cpu().numpy()
    assert f.sum() == 1 and f[np.argmax(r)] == 1
    worst = int(np.argmax(r))
    one = [tuple(p[worst:worst + 1] for p in X) for X in outs]
    errs = _errors(g, one, tuple(x[worst:worst + 1] for x in A), tuple(x[worst:worst + 1] for x in SL),
                   tuple(x[worst:worst + 1] for x in SG))
    assert max(errs) <= 1e-10
