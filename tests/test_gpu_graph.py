"""hipGraph-captured sweeps (replayed with new data, checked against eager
solves after every replay) and the block-size GEMM autotuner."""

import json

import numpy as np
import pytest

from quatrex_amd import _lib, autotune
from quatrex_amd import solve as S
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry

pytestmark = pytest.mark.gpu


def _resident(g, seed, device):
    import torch

    m = W.make_device(g, seed)
    t = lambda x: torch.as_tensor(x, device=device)  # noqa: E731
    sb = [None] * g.n_blocks
    for i in W.boundary_blocks(g):
        sb[i] = t(W.BOUNDARY_SIGMA * np.eye(g.block_sizes[i], dtype=np.complex128))
    return m, (t(m.h_diag), t(m.h_upper), t(m.h_lower)), sb


@pytest.mark.parametrize("blocks", [(48, 40, 48, 32, 48), (40,)])
@pytest.mark.parametrize("precision,refine", [("fp64", 0), ("mixed", 1), ("mixed", -1)])
def test_graph_replays_match_eager(device, precision, refine, blocks):
    import torch

    g = Geometry(blocks)
    E = 3
    m, h, sb = _resident(g, 77, device)
    z = torch.zeros(E, dtype=torch.complex128, device=device)
    A = S.packed_empty(g, E, device)
    SL = (torch.as_tensor(m.sl_diag, device=device).expand(E, -1).contiguous(),
          torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=device))
    SG = (torch.as_tensor(m.sg_diag, device=device).expand(E, -1).contiguous(),
          torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=device))
    outs = [S.packed_empty(g, E, device), S.packed_empty(g, E, device, ("d", "u")),
            S.packed_empty(g, E, device, ("d", "u"))]
    ref = [S.packed_empty(g, E, device), S.packed_empty(g, E, device, ("d", "u")),
           S.packed_empty(g, E, device, ("d", "u"))]
    ctx = S.Context(device)
    graph = S.SolveGraph(g, A, SL, SG, g_retarded=outs[0], g_lesser=outs[1], g_greater=outs[2],
                         ctx=ctx, precision=precision, refine=refine)
    rng = np.random.default_rng(3)
    for rep in range(3):
        en = rng.uniform(-1, 1, E)
        z.copy_(torch.as_tensor(en + 1e-3j, dtype=torch.complex128))
        S.assemble(g, z, h, a=A, sigma_b=sb, ctx=ctx)
        graph.launch()
        ctx.status()
        # eager reference on the same inputs (other buffers), with a sync in between
        kw = {} if precision == "fp64" else dict(precision="mixed",
                                                 mixed=S.MixedOptions(refine=refine, fallback_tol=0.0))
        S.selected_solve(g, A, SL, SG, g_retarded=ref[0], g_lesser=ref[1], g_greater=ref[2], ctx=ctx, **kw)
        ctx.status()
        for X, R in zip(outs, ref):
            for x, r in zip(X, R):
                if x is not None:
                    assert torch.equal(x, r), rep
        # eager work between replays
        (A[0] * 2).sum().item()
    graph.close()
    ctx.close()


def test_autotune_picks_valid_config_and_caches(device, tmp_path):
    ctx = S.Context(device)
    path = tmp_path / "tiles.json"
    e1 = autotune.tune(ctx, 64, n_energies=32, path=path)
    assert e1["best"] in range(_lib.RGF_NUM_TILE_CFGS) and e1["best_name"] in e1["ms"]
    assert len(e1["ms"]) == _lib.RGF_NUM_TILE_CFGS
    assert all(v > 0 for v in e1["ms"].values())
    e2 = autotune.tune(ctx, 64, n_energies=32, f32=True, path=path)
    assert e2["best"] in (0, 1)
    db = json.loads(path.read_text())
    assert len(next(iter(db.values()))) == 2
    # cached: no re-timing at the same energy count
    assert autotune.tune(ctx, 64, n_energies=32, path=path) == e1
    # a tuned context still solves correctly
    g = Geometry.uniform(4, 64)
    m, h, sb = _resident(g, 5, device)
    import torch

    z = torch.tensor([0.2 + 1e-3j], dtype=torch.complex128, device=device)
    A = S.packed_empty(g, 1, device)
    S.assemble(g, z, h, a=A, sigma_b=sb, ctx=ctx)
    SL = (torch.as_tensor(m.sl_diag, device=device)[None].contiguous(),
          torch.zeros((1, g.offdiag_size), dtype=torch.complex128, device=device))
    GR, GL, GG = S.packed_empty(g, 1, device), S.packed_empty(g, 1, device, ("d", "u")), S.packed_empty(g, 1, device, ("d", "u"))
    S.selected_solve(g, A, SL, SL, g_retarded=GR, g_lesser=GL, g_greater=GG, ctx=ctx)
    ctx.status()
    Ah = tuple(x.cpu().numpy() for x in A)
    dense = np.linalg.inv(g.to_dense(*Ah))[0]
    blk = g.diag(GR[0].cpu().numpy(), 2)[0]
    bo = g.block_offsets
    ref = dense[bo[2]:bo[3], bo[2]:bo[3]]
    assert np.linalg.norm(blk - ref) / np.linalg.norm(ref) < 1e-10
    ctx.close()


@pytest.mark.parametrize("precision,refine", [("fp64", 0), ("mixed", 1)])
def test_graph_with_current_matches_eager_and_dense(device, precision, refine):
    """A captured sweep that also returns the Meir-Wingreen layer currents
    (the trace-job table is written by a kernel node of the graph): every
    replay equals the eager solve with the current, with eager work and
    synchronisations in between, and the FP64 currents match the brute-force
    dense current (left-connected inverses, no RGF recurrence)."""
    import torch

    from oracle.dense import dense_current

    rng = np.random.default_rng(17)
    g = Geometry((24, 40, 16, 33, 24))
    N, E = g.n_blocks, 3
    m = W.make_device(g, 78)
    h = tuple(torch.as_tensor(x, device=device) for x in (m.h_diag, m.h_upper, m.h_lower))

    def contact(b, f):
        br = -0.5j * np.eye(b) + 0.05 * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b)))
        gam = 1j * (br - br.conj().T)
        rep = lambda x: np.broadcast_to(x, (E, b, b)).copy()  # noqa: E731
        return rep(br), rep(1j * gam * f), rep(1j * gam * (f - 1))

    L, R = contact(g.block_sizes[0], 0.8), contact(g.block_sizes[-1], 0.3)
    obc_np = [[L[k]] + [None] * (N - 2) + [R[k]] for k in range(3)]
    obc = [[torch.as_tensor(x, device=device) if x is not None else None for x in lst] for lst in obc_np]
    z = torch.zeros(E, dtype=torch.complex128, device=device)
    A = S.packed_empty(g, E, device)
    SL = (torch.as_tensor(m.sl_diag, device=device).expand(E, -1).contiguous(),
          torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=device))
    SG = (torch.as_tensor(m.sl_diag, device=device).expand(E, -1).contiguous(),
          torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=device))
    mk = lambda: [S.packed_empty(g, E, device), S.packed_empty(g, E, device, ("d", "u")),  # noqa: E731
                  S.packed_empty(g, E, device, ("d", "u"))]
    outs, ref = mk(), mk()
    cur = torch.empty((E, N + 1), dtype=torch.complex128, device=device)
    cur_ref = torch.empty_like(cur)
    ctx = S.Context(device)
    graph = S.SolveGraph(g, A, SL, SG, g_retarded=outs[0], g_lesser=outs[1], g_greater=outs[2],
                         obc=obc, current=cur, ctx=ctx, precision=precision, refine=refine)
    kw = {} if precision == "fp64" else dict(precision="mixed",
                                             mixed=S.MixedOptions(refine=refine, fallback_tol=0.0))
    for rep in range(3):
        en = rng.uniform(-1, 1, E)
        z.copy_(torch.as_tensor(en + 1e-2j, dtype=torch.complex128))
        S.assemble(g, z, h, a=A, ctx=ctx)  # the boundary enters as OBC blocks
        graph.launch()
        ctx.status()
        S.selected_solve(g, A, SL, SG, g_retarded=ref[0], g_lesser=ref[1], g_greater=ref[2], obc=obc,
                         current=cur_ref, ctx=ctx, **kw)
        ctx.status()
        assert torch.equal(cur, cur_ref), rep
        for X, Rf in zip(outs, ref):
            for x, r in zip(X, Rf):
                if x is not None:
                    assert torch.equal(x, r), rep
        if precision == "fp64":
            a = g.to_dense(*(x.cpu().numpy() for x in A))
            sl = g.to_dense(SL[0].cpu().numpy(), None, None)
            sg = g.to_dense(SG[0].cpu().numpy(), None, None)
            cd = dense_current(a, sl, sg, g.block_offsets, obc_np)
            err = np.abs(cur.cpu().numpy() - cd).max() / np.abs(cd).max()
            assert err <= 1e-10, (rep, err)
        (A[0] * 2).sum().item()  # eager work between replays
    graph.close()
    ctx.close()
