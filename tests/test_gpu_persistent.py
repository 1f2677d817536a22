"""The persistent on-chip sweep (geometries whose blocks are all <= 32: one
workgroup per energy point, every intermediate block in LDS, energy points of
several geometries in one device work queue) against the dense oracle and
against the launch-per-product-group sweep."""

import numpy as np
import pytest

from oracle.dense import dense_solve
from quatrex_amd import _lib
from quatrex_amd import solve as S
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry
from tests.helpers import block_rel_errors, skew_tridiagonal, tridiagonal_mask
from tests.test_gpu_rgf import _check_solve, _generator_case, _t

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _ctx(device, persistent):
    ctx = S.Context(device)
    ctx.set_option(_lib.RGF_OPT_PERSISTENT, 1 if persistent else 0)
    return ctx


@pytest.mark.parametrize("nb,b,ne", [(8, 32, 3), (32, 32, 2), (5, 16, 4), (3, 7, 2), (1, 32, 2),
                                     (6, 1, 2)])
def test_generator_workloads_vs_dense(device, nb, b, ne):
    g = Geometry.uniform(nb, b)
    _, A, SL, SG = _generator_case(g, ne, W.DEVICE_SEED_BASE + 70 + nb)
    errs = _check_solve(device, g, A, SL, SG)
    assert max(errs) <= TOL, errs


def _mixed_case(bs, E, seed, rng):
    g = Geometry(tuple(bs))
    model = W.make_device(g, seed)
    en = np.linspace(-0.7, 0.6, E)
    A = W.system_matrix(model, en, 1e-2, fold_boundary=False)
    sl = np.broadcast_to(skew_tridiagonal(bs, rng), (E, g.n, g.n))
    sg = np.broadcast_to(skew_tridiagonal(bs, rng), (E, g.n, g.n))
    SL = g.from_dense(sl)[:2]
    SG = g.from_dense(sg)[:2]

    def one(b, f):
        br = -0.5j * np.eye(b) + 0.05 * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b)))
        gam = 1j * (br - br.conj().T)
        rep = lambda x: np.broadcast_to(x, (E, b, b)).copy()  # noqa: E731
        return rep(br), rep(1j * gam * f), rep(1j * gam * (f - 1))

    L, R = one(bs[0], 0.8), one(bs[-1], 0.3)
    if len(bs) == 1:
        obc = [[L[k]] for k in range(3)]
    else:
        obc = [[L[k]] + [None] * (len(bs) - 2) + [R[k]] for k in range(3)]
    return g, A, SL, SG, obc


@pytest.mark.parametrize("bs", [[24, 32, 16, 31, 24], [2, 2, 2, 4, 4, 2, 2, 2], [32], [5, 32, 1, 17]])
def test_mixed_sizes_offdiag_sigma_obc_current_vs_dense(device, bs):
    """Mixed block sizes (padded inside the kernel), off-diagonal Sigma^<>,
    OBC on both ends and the Meir-Wingreen current against the brute-force
    dense current."""
    rng = np.random.default_rng(31 + len(bs))
    g, A, SL, SG, obc = _mixed_case(bs, 3, 91, rng)
    errs = _check_solve(device, g, A, SL, SG, obc_np=obc, with_current=True)
    assert max(errs) <= TOL, errs


def _solve_with(device, ctx, g, A, SL, SG, obc=None, current=False, retarded=True):
    import torch

    E = A[0].shape[0]
    t = lambda x: _t(x, device)  # noqa: E731
    GR = S.packed_empty(g, E, device) if retarded else None
    GL = S.packed_empty(g, E, device)
    GG = S.packed_empty(g, E, device, ("d", "u"))
    cur = torch.empty((E, g.n_blocks + 1), dtype=torch.complex128, device=device) if current else None
    ob = [[t(x) if x is not None else None for x in lst] for lst in obc] if obc is not None else None
    S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)), g_lesser=GL,
                     g_greater=GG, g_retarded=GR, obc=ob, current=cur, ctx=ctx)
    ctx.status()
    return GR, GL, GG, cur


def _rel(x, r):
    import torch

    return float(torch.linalg.norm(x - r) / torch.linalg.norm(r))


@pytest.mark.parametrize("retarded", [True, False])
def test_persistent_equals_launch_per_product_sweep(device, retarded):
    """Many more energy points than workgroups (every workgroup takes several
    from the queue), OBC and the current, with and without the retarded
    output (without it g_i lives in the per-workgroup scratch): the persistent
    sweep against the launch-per-product-group sweep of the same inputs."""
    rng = np.random.default_rng(5)
    bs = [16, 32, 24, 32, 8, 32]
    g = Geometry(tuple(bs))
    E = 700
    model = W.make_device(g, 93)
    en = rng.uniform(-1, 1, E)
    A = W.system_matrix(model, en, 1e-2, fold_boundary=False)
    sld, slu, sgd, sgu = W.sigmas(model, E)
    obc = _mixed_case(bs, E, 93, rng)[4]
    outs = []
    for persistent in (True, False):
        ctx = _ctx(device, persistent)
        outs.append(_solve_with(device, ctx, g, A, (sld, slu), (sgd, sgu), obc=obc, current=True,
                                retarded=retarded))
        ctx.close()
    (GRp, GLp, GGp, cp), (GRs, GLs, GGs, cs) = outs
    if retarded:
        for x, r in zip(GRp, GRs):
            assert _rel(x, r) <= 1e-12
    for X, R in ((GLp, GLs), (GGp, GGs)):
        for x, r in zip(X, R):
            if x is not None:
                assert _rel(x, r) <= 1e-12
    assert _rel(cp, cs) <= 1e-11


def test_uneven_batch_one_persistent_launch(device):
    """Energy points of three small geometries and one with 48-row blocks in
    one batch call: the small ones share one persistent launch (one work
    queue), the large one takes the batched sweep; every result against the
    dense oracle, and the small ones bit for bit equal to solving each
    geometry alone."""
    import torch

    cases = [((8, 32, 3), 3), ((4, 16, 2, 16), 5), ((1, 1, 1), 2), ((48, 40, 48), 2)]
    probs, refs, alone = [], [], []
    for k, (bs, E) in enumerate(cases):
        g = Geometry(bs)
        _, A, SL, SG = _generator_case(g, E, W.DEVICE_SEED_BASE + 80 + k)
        t = lambda x: _t(x, device)  # noqa: E731
        GR, GL, GG = S.packed_empty(g, E, device), S.packed_empty(g, E, device), S.packed_empty(g, E, device)
        probs.append(S.SolveProblem(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)),
                                    g_lesser=GL, g_greater=GG, g_retarded=GR))
        z = np.zeros_like(SL[1])
        refs.append((g, dense_solve(g.to_dense(*A), g.to_dense(SL[0], SL[1], z),
                                    g.to_dense(SG[0], SG[1], z))))
        alone.append(_solve_with(device, S.default_context(device), g, A, SL, SG))
    S.selected_solve_batch(probs)
    S.default_context(device).status()
    for p, (g, ref), one in zip(probs, refs, alone):
        bs = list(g.block_sizes)
        m = tridiagonal_mask(bs)
        for X, r in zip((p.g_retarded, p.g_lesser, p.g_greater), ref):
            xd = g.to_dense(*(x.cpu().numpy() for x in X))
            assert block_rel_errors(xd, r * m, bs) <= TOL
        if max(bs) <= 32:
            for X, R in zip((p.g_retarded, p.g_lesser, p.g_greater), one[:3]):
                for x, r in zip(X, R):
                    if x is not None and r is not None:
                        assert torch.equal(x, r)


def test_graph_capture_of_persistent_sweep(device):
    """A captured graph of a small geometry records the persistent launch
    (table uploads and the queue reset included): replays with new energies
    and eager work in between equal the eager persistent solve."""
    import torch

    g = Geometry((32, 24, 32, 16))
    E = 5
    m = W.make_device(g, 97)
    h = tuple(torch.as_tensor(x, device=device) for x in (m.h_diag, m.h_upper, m.h_lower))
    sb = [None] * g.n_blocks
    for i in W.boundary_blocks(g):
        sb[i] = torch.as_tensor(W.BOUNDARY_SIGMA * np.eye(g.block_sizes[i], dtype=np.complex128),
                                device=device)
    z = torch.zeros(E, dtype=torch.complex128, device=device)
    A = S.packed_empty(g, E, device)
    SL = (torch.as_tensor(m.sl_diag, device=device).expand(E, -1).contiguous(),
          torch.zeros((E, g.offdiag_size), dtype=torch.complex128, device=device))
    SG = (torch.as_tensor(m.sg_diag, device=device).expand(E, -1).contiguous(), SL[1])
    mk = lambda: [S.packed_empty(g, E, device), S.packed_empty(g, E, device, ("d", "u")),  # noqa: E731
                  S.packed_empty(g, E, device, ("d", "u"))]
    outs, ref = mk(), mk()
    ctx = S.Context(device)
    graph = S.SolveGraph(g, A, SL, SG, g_retarded=outs[0], g_lesser=outs[1], g_greater=outs[2], ctx=ctx)
    rng = np.random.default_rng(8)
    for rep in range(3):
        z.copy_(torch.as_tensor(rng.uniform(-1, 1, E) + 1e-3j, dtype=torch.complex128))
        S.assemble(g, z, h, a=A, sigma_b=sb, ctx=ctx)
        graph.launch()
        ctx.status()
        S.selected_solve(g, A, SL, SG, g_retarded=ref[0], g_lesser=ref[1], g_greater=ref[2], ctx=ctx)
        ctx.status()
        for X, R in zip(outs, ref):
            for x, r in zip(X, R):
                if x is not None:
                    assert torch.equal(x, r), rep
        (A[0] * 2).sum().item()
    graph.close()
    ctx.close()


def test_singular_block_is_reported(device):
    """An exactly singular block to invert (the first diagonal block of one
    energy set to zero) in the persistent path sets the status word like the
    batched inverse does."""
    import torch

    g = Geometry((8, 8, 8))
    E = 2
    _, A, SL, SG = _generator_case(g, E, W.DEVICE_SEED_BASE + 99)
    A = [x.copy() for x in A]
    blk = g.diag(A[0], 0)
    blk[1, :, :] = 0.0  # energy 1, block 0: a zero matrix
    ctx = S.Context(device)
    t = lambda x: _t(x, device)  # noqa: E731
    GR, GL, GG = (S.packed_empty(g, E, device) for _ in range(3))
    S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)), g_lesser=GL,
                     g_greater=GG, g_retarded=GR, ctx=ctx)
    with pytest.raises(np.linalg.LinAlgError, match="energy 1, block 0"):
        ctx.status()
    ctx.close()
    torch.cuda.synchronize()
