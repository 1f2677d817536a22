"""GPU parity of the HIP RGF path against the CPU oracle.

Bars: relative Frobenius error per selected block <= 1e-10 against the dense
oracle (numpy.linalg.inv of the full matrix) for the FP64 path; the drop-in
class is exercised on the reference test axes; full-size (32 x 256) runs are
checked through size-independent identities.
"""

import numpy as np
import pytest

from oracle.dense import dense_current, dense_solve
from oracle.rgf_ref import BlockTri
from oracle.rgf_ref import selected_solve as oracle_solve
from quatrex_amd import _lib
from quatrex_amd import solve as S
from quatrex_amd import workloads as W
from quatrex_amd.datastructures import BlockTridiagonal
from quatrex_amd.greens_function_solver import RGF, OBCBlocks
from quatrex_amd.layout import Geometry
from tests.helpers import (BATCHING, BLOCK_SIZES, STACK_SHAPES, block_rel_errors, bt_dense,
                           edge_case, skew_tridiagonal, tridiagonal_mask)

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _t(x, device):
    import torch

    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.complex128, device=device)


# --------------------------------------------------------------------------- inverse


@pytest.mark.parametrize("n", [1, 2, 3, 17, 32, 33, 64, 65, 100, 128, 129, 200, 255, 256, 300, 512,
                               640])
def test_batched_inverse_vs_numpy(device, n):
    rng = np.random.default_rng(n)
    nm = 3
    a = rng.standard_normal((nm, n, n)) + 1j * rng.standard_normal((nm, n, n))
    a[1] += 4 * np.eye(n)
    out = S.inv(_t(a, device)).cpu().numpy()
    S.default_context(device).status()
    ref = np.linalg.inv(a)
    for k in range(nm):
        err = np.linalg.norm(out[k] - ref[k]) / np.linalg.norm(ref[k])
        cond = np.linalg.cond(a[k])
        assert err <= 1e-14 * cond * n, (k, err, cond)


def test_inverse_needs_pivoting(device):
    """Zero leading diagonal: only row pivoting makes the elimination possible."""
    rng = np.random.default_rng(5)
    n = 96
    perm = rng.permutation(n)
    a = np.zeros((2, n, n), np.complex128)
    a[:, np.arange(n), perm] = 1.0 + 0.5j
    a += 1e-3 * (rng.standard_normal((2, n, n)))
    a[:, 0, 0] = 0.0
    out = S.inv(_t(a, device)).cpu().numpy()
    S.default_context(device).status()
    assert np.abs(out @ a - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("n", [4, 48, 96, 200, 300])
def test_singular_inverse_raises(device, n):
    """A zero row: the last elimination step meets an exactly zero pivot, in
    every panel layout (two threads per row, 64-column panels, paired
    32-column panels)."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    a[1, n // 3, :] = 0.0
    ctx = S.default_context(device)
    S.inv(_t(a, device))
    with pytest.raises(np.linalg.LinAlgError):
        ctx.status()
    ctx.status()  # status resets after reporting


def test_singular_block_raises(device):
    g = Geometry((4, 4))
    a = bt_dense([4, 4], seed=3)
    a[0:4, :] = 0.0  # zero first block row -> exactly singular A_00
    d, u, l = g.from_dense(a[None])
    x = S.packed_empty(g, 1, device)
    ctx = S.default_context(device)
    S.selected_inv(g, (_t(d, device), _t(u, device), _t(l, device)), x=x, ctx=ctx)
    with pytest.raises(np.linalg.LinAlgError):
        ctx.status()
    ctx.status()  # status resets after reporting


# --------------------------------------------------------------------------- drop-in


@pytest.mark.parametrize("bs_name", list(BLOCK_SIZES))
@pytest.mark.parametrize("max_batch_size", BATCHING)
@pytest.mark.parametrize("stack", STACK_SHAPES)
@pytest.mark.parametrize("container", ["device", "numpy"])
def test_dropin_selected_inv(device, bs_name, max_batch_size, stack, container):
    bs = BLOCK_SIZES[bs_name]
    a = bt_dense(bs, seed=11)
    ref = np.linalg.inv(a) * tridiagonal_mask(bs)
    if container == "device":
        A = BlockTridiagonal.from_sparray(a, bs, stack, device=device)
    else:
        A = BlockTri.from_dense(np.broadcast_to(a, (*stack, *a.shape)), bs, stack)
    x = RGF(max_batch_size=max_batch_size).selected_inv(A)
    xd = x.to_dense()
    xd = xd.cpu().numpy() if hasattr(xd, "cpu") else xd
    assert block_rel_errors(xd, np.broadcast_to(ref, xd.shape), bs) <= TOL


@pytest.mark.parametrize("bs_name", list(BLOCK_SIZES))
@pytest.mark.parametrize("max_batch_size", BATCHING)
@pytest.mark.parametrize("stack", STACK_SHAPES)
@pytest.mark.parametrize("out", [True, False])
@pytest.mark.parametrize("return_retarded", [True, False])
def test_dropin_selected_solve(device, bs_name, max_batch_size, stack, out, return_retarded):
    """The reference's test_selected_solve, at the stricter per-block bar."""
    bs = BLOCK_SIZES[bs_name]
    a = bt_dense(bs, seed=12)
    bl = a - a.conj().T
    gr, gl, gg = dense_solve(a, bl, bl)
    A = BlockTridiagonal.from_sparray(a, bs, stack, device=device)
    Bl = BlockTridiagonal.from_sparray(bl, bs, stack, device=device)
    Bg = BlockTridiagonal.from_sparray(bl, bs, stack, device=device)
    solver = RGF(max_batch_size=max_batch_size)
    if out:
        Xr, Xl, Xg = (BlockTridiagonal.zeros_like(A) for _ in range(3))
        assert solver.selected_solve(A, Bl, Bg, out=[Xl, Xg, Xr], return_retarded=return_retarded) is None
    else:
        res = solver.selected_solve(A, Bl, Bg, return_retarded=return_retarded)
        Xl, Xg = res[:2]
        if return_retarded:
            Xr = res[2]
    mask = tridiagonal_mask(bs)
    for X, ref in ((Xl, gl), (Xg, gg)) + (((Xr, gr),) if return_retarded else ()):
        xd = X.to_dense().cpu().numpy()
        assert block_rel_errors(xd, np.broadcast_to(ref * mask, xd.shape), bs) <= TOL


def test_dropin_numpy_container_and_symmetric_output(device):
    """A numpy-backed container with the DSDBSparse interface; symmetric G^< storage."""
    bs = [3, 5, 2, 4]
    a = bt_dense(bs, seed=13)
    rng = np.random.default_rng(1)
    sl = skew_tridiagonal(bs, rng)
    sg = skew_tridiagonal(bs, rng)
    st = (4,)
    mk = lambda m: BlockTri.from_dense(np.broadcast_to(m, (*st, *m.shape)), bs, st)  # noqa
    xl = BlockTri(bs, st, symmetry=True, symmetry_op=lambda x: -x.conj())
    xg = BlockTri(bs, st, symmetry=True, symmetry_op=lambda x: -x.conj())
    xr = BlockTri(bs, st)
    RGF(max_batch_size=2).selected_solve(mk(a), mk(sl), mk(sg), out=[xl, xg, xr], return_retarded=True)
    gr, gl, gg = dense_solve(a, sl, sg)
    m = tridiagonal_mask(bs)
    assert block_rel_errors(xl.to_dense(), np.broadcast_to(gl * m, xl.shape), bs) <= TOL
    assert block_rel_errors(xg.to_dense(), np.broadcast_to(gg * m, xg.shape), bs) <= TOL
    assert block_rel_errors(xr.to_dense(), np.broadcast_to(gr * m, xr.shape), bs) <= TOL


def test_dropin_wrong_out_length_raises(device):
    bs = [2, 2]
    A = BlockTridiagonal.from_sparray(bt_dense(bs), bs, (2,), device=device)
    X = BlockTridiagonal.zeros_like(A)
    with pytest.raises(ValueError):
        RGF().selected_solve(A, A, A, out=[X, X], return_retarded=True)
    with pytest.raises(TypeError):
        RGF().selected_solve(A, A, A, return_current=True)


# --------------------------------------------------------------------------- workloads


def _generator_case(g, n_e, seed, eta=1e-3, fold=True):
    model = W.make_device(g, seed)
    en = W.draw_energies(n_e)
    ad, au, al = W.system_matrix(model, en, eta, fold_boundary=fold)
    sld, slu, sgd, sgu = W.sigmas(model, n_e)
    return model, (ad, au, al), (sld, slu), (sgd, sgu)


def _check_solve(device, g, A, SL, SG, obc_np=None, with_current=False):
    E = A[0].shape[0]
    t = lambda x: _t(x, device)  # noqa
    GR = S.packed_empty(g, E, device)
    GL = S.packed_empty(g, E, device)
    GG = S.packed_empty(g, E, device)
    obc = None
    if obc_np is not None:
        obc = [[t(x) if x is not None else None for x in lst] for lst in obc_np]
    cur = None
    if with_current:
        import torch

        cur = torch.empty((E, g.n_blocks + 1), dtype=torch.complex128, device=device)
    S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)), g_lesser=GL,
                     g_greater=GG, g_retarded=GR, obc=obc, current=cur)
    S.default_context(device).status()
    a = g.to_dense(*A)
    def lower(u):
        out = np.empty_like(u)
        for i in range(g.n_blocks - 1):
            g.lower(out, i)[...] = -np.conj(np.swapaxes(g.upper(u, i), -1, -2))
        return out

    sl = g.to_dense(SL[0], SL[1], lower(SL[1]))
    sg = g.to_dense(SG[0], SG[1], lower(SG[1]))
    ref = dense_solve(a, sl, sg, g.block_offsets, obc_np)
    bs = list(g.block_sizes)
    m = tridiagonal_mask(bs)
    errs = []
    for X, r in zip((GR, GL, GG), ref):
        xd = g.to_dense(*(x.cpu().numpy() for x in X))
        errs.append(block_rel_errors(xd, r * m, bs))
    if with_current:
        cref = dense_current(a, sl, sg, g.block_offsets, obc_np)
        cerr = np.abs(cur.cpu().numpy() - cref).max() / np.abs(cref).max()
        errs.append(cerr)
    return errs


@pytest.mark.parametrize("nb,b,ne", [(8, 32, 2), (5, 64, 3), (4, 100, 2), (3, 128, 2), (2, 256, 1),
                                     (3, 384, 1), (3, 512, 1), (2, 520, 1)])
def test_generator_workload_vs_dense(device, nb, b, ne):
    g = Geometry.uniform(nb, b)
    _, A, SL, SG = _generator_case(g, ne, W.DEVICE_SEED_BASE + nb)
    errs = _check_solve(device, g, A, SL, SG)
    assert max(errs) <= TOL, errs


@pytest.mark.parametrize("sizes", [(520, 512), (600, 400, 600)])
def test_mixed_panel_widths_vs_dense(device, sizes):
    """Blocks above 512 rows (16-column inversion panels) next to blocks of
    at most 512 (32-column panels, whose per-energy scratch is larger): the
    selected solve and the selected inversion against the dense oracle."""
    g = Geometry(sizes)
    _, A, SL, SG = _generator_case(g, 2, W.DEVICE_SEED_BASE + 31)
    errs = _check_solve(device, g, A, SL, SG)
    assert max(errs) <= TOL, errs
    X = S.packed_empty(g, 2, device)
    S.selected_inv(g, tuple(_t(x, device) for x in A), x=X)
    S.default_context(device).status()
    ref = np.linalg.inv(g.to_dense(*A))
    xd = g.to_dense(*(x.cpu().numpy() for x in X))
    err = block_rel_errors(xd, ref * tridiagonal_mask(list(sizes)), list(sizes))
    assert err <= TOL, err


def test_mixed_blocks_offdiag_sigma_obc_current(device):
    rng = np.random.default_rng(21)
    bs = [24, 40, 16, 33, 24]
    g = Geometry(tuple(bs))
    model = W.make_device(g, 77)
    en = np.array([0.05, -0.6, 0.33])
    A = W.system_matrix(model, en, 1e-2, fold_boundary=False)
    E = len(en)
    sl = np.broadcast_to(skew_tridiagonal(bs, rng), (E, g.n, g.n))
    sg = np.broadcast_to(skew_tridiagonal(bs, rng), (E, g.n, g.n))
    SL = g.from_dense(sl)[:2]
    SG = g.from_dense(sg)[:2]

    def one(b, f):
        br = -0.5j * np.eye(b) + 0.05 * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b)))
        gam = 1j * (br - br.conj().T)
        rep = lambda x: np.broadcast_to(x, (E, b, b)).copy()  # noqa
        return rep(br), rep(1j * gam * f), rep(1j * gam * (f - 1))

    L, R = one(bs[0], 0.8), one(bs[-1], 0.3)
    obc = [[L[k]] + [None] * (len(bs) - 2) + [R[k]] for k in range(3)]
    errs = _check_solve(device, g, A, SL, SG, obc_np=obc, with_current=True)
    assert max(errs) <= TOL, errs


@pytest.mark.parametrize("cfg", range(_lib.RGF_NUM_TILE_CFGS))
def test_every_tile_config_vs_dense(device, cfg):
    """Every GEMM tile configuration the autotuner can install gives the dense
    result: block sizes that leave partial tiles of each shape (24 ... 200),
    off-diagonal Sigma, both contacts and the current."""
    rng = np.random.default_rng(400 + cfg)
    bs = [24, 40, 130, 33, 200]
    g = Geometry(tuple(bs))
    model = W.make_device(g, 78)
    en = np.array([0.2, -0.45])
    A = W.system_matrix(model, en, 1e-2, fold_boundary=False)
    E = len(en)
    sl = np.broadcast_to(skew_tridiagonal(bs, rng), (E, g.n, g.n))
    sg = np.broadcast_to(skew_tridiagonal(bs, rng), (E, g.n, g.n))
    SL = g.from_dense(sl)[:2]
    SG = g.from_dense(sg)[:2]

    def one(b, f):
        br = -0.5j * np.eye(b) + 0.05 * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b)))
        gam = 1j * (br - br.conj().T)
        rep = lambda x: np.broadcast_to(x, (E, b, b)).copy()  # noqa
        return rep(br), rep(1j * gam * f), rep(1j * gam * (f - 1))

    L, R = one(bs[0], 0.7), one(bs[-1], 0.2)
    obc = [[L[k]] + [None] * (len(bs) - 2) + [R[k]] for k in range(3)]
    ctx = S.default_context(device)
    ctx.set_option(_lib.RGF_OPT_TILE_CFG, cfg)
    try:
        errs = _check_solve(device, g, A, SL, SG, obc_np=obc, with_current=True)
    finally:
        ctx.set_option(_lib.RGF_OPT_TILE_CFG, -1)
    assert max(errs) <= TOL, errs


@pytest.mark.parametrize("bs", [[40], [1] * 8, [1, 3, 1, 2], [33]])
def test_edge_geometries_vs_dense(device, bs):
    """A single block (the sweeps are empty; both boundary terms of the
    current read block 0) and 1 x 1 blocks, with off-diagonal Sigma, both
    contacts and the current, against the dense oracle."""
    g, A, SL, SG, obc = edge_case(bs, 500 + len(bs))
    errs = _check_solve(device, g, A, SL, SG, obc_np=obc, with_current=True)
    assert max(errs) <= TOL, errs


def test_zero_energies_are_a_no_op(device):
    """n_energies = 0: both entry points return without launching work and
    leave the (empty) outputs alone."""
    import torch

    g = Geometry((8, 8, 8))
    z = lambda shape: torch.zeros(shape, dtype=torch.complex128, device=device)  # noqa
    A = (z((0, g.diag_size)), z((0, g.offdiag_size)), z((0, g.offdiag_size)))
    SL = (z((0, g.diag_size)), z((0, g.offdiag_size)))
    SG = (z((0, g.diag_size)), z((0, g.offdiag_size)))
    GR, GL, GG = (S.packed_empty(g, 0, device) for _ in range(3))
    S.selected_solve(g, A, SL, SG, g_lesser=GL, g_greater=GG, g_retarded=GR)
    X = S.packed_empty(g, 0, device)
    S.selected_inv(g, A, x=X)
    S.default_context(device).status()
    torch.cuda.synchronize()
    assert GR[0].shape[0] == 0 and X[0].shape[0] == 0


def test_retarded_obc_equals_folded_boundary(device):
    """Sigma_B folded into A, or passed as OBCBlocks.retarded, gives the same result."""
    g = Geometry.uniform(6, 48)
    model, Af, SL, SG = _generator_case(g, 2, 99)
    _, An, _, _ = _generator_case(g, 2, 99, fold=False)
    obc_r = W.boundary_obc(model, 2)
    obc = [[obc_r.get(i) for i in range(6)], [None] * 6, [None] * 6]
    e1 = _check_solve(device, g, Af, SL, SG)
    e2 = _check_solve(device, g, An, SL, SG, obc_np=obc)
    assert max(e1 + e2) <= TOL


def test_energy_alone_equals_in_batch(device):
    """Each energy's result is bitwise the same solved alone or inside a batch."""
    import torch

    g = Geometry.uniform(6, 64)
    _, A, SL, SG = _generator_case(g, 5, 55)
    t = lambda x: _t(x, device)  # noqa
    outs = []
    for sl in (slice(0, 5), slice(2, 3)):
        GL = S.packed_empty(g, sl.stop - sl.start, device)
        GG = S.packed_empty(g, sl.stop - sl.start, device)
        GR = S.packed_empty(g, sl.stop - sl.start, device)
        S.selected_solve(g, tuple(t(x[sl]) for x in A), tuple(t(x[sl]) for x in SL),
                         tuple(t(x[sl]) for x in SG), g_lesser=GL, g_greater=GG, g_retarded=GR)
        outs.append((GL, GG, GR))
    torch.cuda.synchronize()
    for Xb, Xa in zip(outs[0], outs[1]):
        for pb, pa in zip(Xb, Xa):
            assert torch.equal(pb[2:3], pa)


def test_profile_counters(device):
    """The HIP-event profile that bench.py reports: per energy the launched
    product flops lie within 5% of the reference op count 8 b^3 (33 N - 29)
    (this build fuses and adds a few products), the algorithmic bytes give the
    b/6-type intensity of b x b products, inversions count 8 b^3 N, and
    reading with reset clears the counters."""
    import torch

    nb, b, E = 6, 64, 3
    g = Geometry.uniform(nb, b)
    _, A, SL, SG = _generator_case(g, E, 71)
    t = lambda x: _t(x, device)  # noqa
    ctx = S.Context(device)
    try:
        GR, GL, GG = (S.packed_empty(g, E, device) for _ in range(3))
        ctx.profile(True)
        ctx.profile_read(reset=True)
        S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)), g_lesser=GL,
                         g_greater=GG, g_retarded=GR, ctx=ctx)
        p = ctx.profile_read(reset=True)
        ctx.status()
        torch.cuda.synchronize()
        assert p["gemm_launches"] > 0 and p["gemm_ms"] > 0 and p["inv_ms"] > 0
        ref = 8.0 * b**3 * (33 * nb - 29) * E
        assert abs(p["gemm_flops"] - ref) / ref < 0.05, (p["gemm_flops"], ref)
        assert p["inv_calls"] == nb and p["inv_flops"] == 8.0 * b**3 * nb * E
        intensity = p["gemm_flops"] / p["gemm_bytes"]
        assert b / 12 < intensity < b / 4, intensity
        again = ctx.profile_read(reset=True)
        assert again["gemm_launches"] == 0 and again["gemm_bytes"] == 0 and again["inv_calls"] == 0
    finally:
        ctx.close()


def test_caller_workspace_solve_exact_and_chunked(device):
    """A caller-owned workspace of exactly rgf_workspace_size bytes for all
    energies, or for 2 of the 5 (the sweep then runs in chunks), gives a
    context-owned workspace's results bit for bit; one byte less than one
    energy needs raises."""
    import torch

    g = Geometry.uniform(5, 48)
    _, A, SL, SG = _generator_case(g, 5, 61)
    t = lambda x: _t(x, device)  # noqa

    def run(ctx):
        out = [S.packed_empty(g, 5, device) for _ in range(3)]
        S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)),
                         g_lesser=out[1], g_greater=out[2], g_retarded=out[0], ctx=ctx)
        ctx.status()
        torch.cuda.synchronize()
        return out

    ref_ctx = S.Context(device)
    try:
        ref = run(ref_ctx)
    finally:
        ref_ctx.close()
    for ne in (5, 2):
        ctx = S.Context(device)
        try:
            ctx.set_workspace(torch.empty(S.workspace_size(g, ne), dtype=torch.uint8, device=device))
            out = run(ctx)
        finally:
            ctx.close()
        for X, R in zip(out, ref):
            for p, q in zip(X, R):
                assert torch.equal(p, q), ne
    ctx = S.Context(device)
    try:
        ctx.set_workspace(torch.empty(S.workspace_size(g, 1) - 1, dtype=torch.uint8, device=device))
        with pytest.raises(_lib.RgfError):
            run(ctx)
    finally:
        ctx.close()


def test_caller_workspace_inverse_sizes(device):
    """rgf_inv_z on caller-owned workspaces of every size from one matrix to
    about three (256-byte steps): once a size is large enough for one matrix
    and its auxiliaries, every larger size works (exact multiples of the
    per-matrix need included) and the result, in chunks, matches numpy."""
    import torch

    n, nm = 64, 7
    rng = np.random.default_rng(64)
    a = rng.standard_normal((nm, n, n)) + 1j * rng.standard_normal((nm, n, n)) + 8 * np.eye(n)
    ref = np.linalg.inv(a)
    ad = _t(a, device)
    big = torch.empty(3 * (2 * n * n * 16 + 4096), dtype=torch.uint8, device=device)
    ctx = S.Context(device)
    first_ok = None
    try:
        for nbytes in range(n * n * 16, big.numel() + 1, 256):
            ctx.set_workspace(big[:nbytes])
            try:
                out = S.inv(ad, ctx=ctx)
                ctx.status()
            except _lib.RgfError:
                assert first_ok is None, (nbytes, first_ok)
                continue
            if first_ok is None:
                first_ok = nbytes
            if (nbytes - first_ok) % 8192 == 0:
                err = np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max()
                assert err < 1e-13, (nbytes, err)
    finally:
        ctx.close()
    assert first_ok is not None


@pytest.mark.slow
def test_full_size_32x256_identities(device):
    """32 x 256: block-row identity (A G^R)_ii = I, anti-Hermitian G^<, and a
    dense cross-check of the first four energies (G^R all selected blocks,
    G^< diagonal and upper blocks) against torch.linalg.inv of the full
    8192 x 8192 matrix on the device (a checker only)."""
    import torch

    g = Geometry.uniform(32, 256)
    ne = 4
    model = W.make_device(g, W.DEVICE_SEED_BASE + 2)
    en = W.draw_energies(ne)
    A = W.system_matrix_torch(model, en, 1e-3, device)
    sld = torch.as_tensor(model.sl_diag, device=device).expand(ne, -1).contiguous()
    sgd = torch.as_tensor(model.sg_diag, device=device).expand(ne, -1).contiguous()
    zu = torch.zeros((ne, g.offdiag_size), dtype=torch.complex128, device=device)
    GR = S.packed_empty(g, ne, device)
    GL = S.packed_empty(g, ne, device)
    GG = S.packed_empty(g, ne, device)
    S.selected_solve(g, A, (sld, zu), (sgd, zu), g_lesser=GL, g_greater=GG, g_retarded=GR)
    S.default_context(device).status()
    eye = torch.eye(256, dtype=torch.complex128, device=device)
    worst = 0.0
    for i in range(32):
        acc = g.diag(A[0], i) @ g.diag(GR[0], i)
        if i > 0:
            acc = acc + g.lower(A[2], i - 1) @ g.upper(GR[1], i - 1)
        if i < 31:
            acc = acc + g.upper(A[1], i) @ g.lower(GR[2], i)
        worst = max(worst, float((torch.linalg.norm(acc - eye, dim=(-2, -1)) / 16.0).max()))
    assert worst < 1e-10, worst
    for i in range(32):
        d = g.diag(GL[0], i)
        assert torch.equal(d, -d.conj().transpose(-1, -2))
    # dense cross-check of the first four energies
    bo = g.block_offsets

    def rel(x, ref):
        return float(torch.linalg.norm(x - ref) / torch.linalg.norm(ref))

    for e in range(ne):
        a = torch.zeros((g.n, g.n), dtype=torch.complex128, device=device)
        sl = torch.zeros_like(a)
        for i in range(32):
            s_ = slice(bo[i], bo[i + 1])
            a[s_, s_] = g.diag(A[0][e:e + 1], i)[0]
            sl[s_, s_] = g.diag(sld[e:e + 1], i)[0]
            if i < 31:
                t_ = slice(bo[i + 1], bo[i + 2])
                a[s_, t_] = g.upper(A[1][e:e + 1], i)[0]
                a[t_, s_] = g.lower(A[2][e:e + 1], i)[0]
        ginv = torch.linalg.inv(a)
        del a
        gl = ginv @ sl @ ginv.conj().transpose(-1, -2)
        del sl
        for i in range(32):
            s_ = slice(bo[i], bo[i + 1])
            assert rel(g.diag(GR[0][e:e + 1], i)[0], ginv[s_, s_]) < TOL, (e, i, "G^R diag")
            assert rel(g.diag(GL[0][e:e + 1], i)[0], gl[s_, s_]) < TOL, (e, i, "G^< diag")
            if i < 31:
                t_ = slice(bo[i + 1], bo[i + 2])
                assert rel(g.upper(GR[1][e:e + 1], i)[0], ginv[s_, t_]) < TOL, (e, i, "G^R upper")
                assert rel(g.lower(GR[2][e:e + 1], i)[0], ginv[t_, s_]) < TOL, (e, i, "G^R lower")
                assert rel(g.upper(GL[1][e:e + 1], i)[0], gl[s_, t_]) < TOL, (e, i, "G^< upper")
        del ginv, gl
        torch.cuda.empty_cache()
