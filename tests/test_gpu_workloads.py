"""The multi-geometry workloads on the GPU, at their real block sizes:

* uneven batching - four interleaved geometries (16, 128), (32, 256),
  (24, 384), (12, 512) through the server with captured sweeps and
  autotuned tiles;
* conditioning - 48 x 256 with eta cycled over 1e-2, 1e-3, 1e-4, FP64 and
  the complex64-product path (whose accuracy is stated, not assumed);
* serving - the six resident geometries the serving benchmark draws
  (n_blocks 8-64, blocks 64-512), each result equal alone and in the mix.

Full sizes are checked through identities that hold for any N: the
block-row identity (A G^R)_ii = I, its lesser / greater counterpart
(A G^<>)_ii = (Sigma^<> G^A)_ii, and exactly anti-Hermitian stored G^<>_ii.
Reduced-N copies with the same block sizes are checked against the dense
oracle (numpy.linalg.inv of the full matrix) at <= 1e-10 per block.
"""

import numpy as np
import pytest

from oracle.dense import dense_solve
from quatrex_amd import solve as S
from quatrex_amd import serving as SV
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry
from tests.helpers import block_rel_errors, tridiagonal_mask

pytestmark = pytest.mark.gpu

TOL = 1e-10
IDENTITY_TOL = 1e-9
UNEVEN = [(16, 128), (32, 256), (24, 384), (12, 512)]


def _h(x):
    return x.conj().transpose(-1, -2).contiguous()


def identity_residuals(g, A, SL, SG, GR, GL, GG):
    """Max over blocks of the three block-row identity residuals of one energy
    (packed device tensors with a leading axis of 1): retarded
    ||(A G^R)_ii - I|| / sqrt(b_i), lesser / greater
    ||(A G)_ii - (Sigma G^A)_ii|| / (||A_ii|| ||G_ii|| + ||Sigma_ii|| ||G^R_ii||).
    Also checks that the stored G^<_ii, G^>_ii are exactly anti-Hermitian."""
    import torch

    N = g.n_blocks
    D = lambda x, i: g.diag(x, i)[0]  # noqa: E731
    U = lambda x, i: g.upper(x, i)[0]  # noqa: E731
    L = lambda x, i: g.lower(x, i)[0]  # noqa: E731
    nrm = lambda x: float(torch.linalg.norm(x))  # noqa: E731
    worst = [0.0, 0.0, 0.0]
    for i in range(N):
        b = int(g.block_sizes[i])
        acc = D(A[0], i) @ D(GR[0], i) - torch.eye(b, dtype=torch.complex128, device=A[0].device)
        if i > 0:
            acc = acc + L(A[2], i - 1) @ U(GR[1], i - 1)
        if i < N - 1:
            acc = acc + U(A[1], i) @ L(GR[2], i)
        worst[0] = max(worst[0], nrm(acc) / np.sqrt(b))
        for k, (X, Sg) in enumerate(((GL, SL), (GG, SG)), start=1):
            d = D(X[0], i)
            assert torch.equal(d, -_h(d)), ("stored diagonal not anti-Hermitian", k, i)
            lhs = D(A[0], i) @ d
            rhs = D(Sg[0], i) @ _h(D(GR[0], i))
            if i > 0:
                lhs = lhs + L(A[2], i - 1) @ U(X[1], i - 1)
                rhs = rhs - _h(U(Sg[1], i - 1)) @ _h(L(GR[2], i - 1))
            if i < N - 1:
                lhs = lhs - U(A[1], i) @ _h(U(X[1], i))
                rhs = rhs + U(Sg[1], i) @ _h(U(GR[1], i))
            scale = nrm(D(A[0], i)) * nrm(d) + nrm(D(Sg[0], i)) * nrm(D(GR[0], i))
            worst[k] = max(worst[k], nrm(lhs - rhs) / scale)
    return worst


def _dense_reference(g, model, energies, etas):
    A = W.system_matrix(model, energies, etas)
    sld, slu, sgd, sgu = W.sigmas(model, len(energies))
    z = np.zeros_like(slu)
    return A, dense_solve(g.to_dense(*A), g.to_dense(sld, slu, z), g.to_dense(sgd, sgu, z))


def _server_identities(srv, reqs, kept):
    """Identity residuals of every kept request (its energy assembled again)."""
    import torch

    worst = [0.0, 0.0, 0.0]
    for rid in kept:
        req = reqs[rid]
        res = srv.resident[req.geometry]
        g = res.geometry
        dev = srv.device
        z = torch.tensor([req.energy + 1j * req.eta], dtype=torch.complex128, device=dev)
        A = S.packed_empty(g, 1, dev)
        S.assemble(g, z, res.h, a=A, sigma_b=res.sigma_b)
        S.default_context(dev).status()
        SL = (res.sl_diag[None], res.sl_upper[None])
        SG = (res.sg_diag[None], res.sg_upper[None])
        GR, GL, GG = (tuple(t[None] if t is not None else None for t in X) for X in srv.kept[rid])
        w = identity_residuals(g, A, SL, SG, GR, GL, GG)
        worst = [max(a, b) for a, b in zip(worst, w)]
    return worst


def _kept_dense(srv, reqs, kept):
    """Dense-oracle check of every kept request: G^R on all tridiagonal
    blocks, G^<> on the diagonal and upper blocks (the lower ones are
    -upper^H and not stored by the server)."""
    for rid in kept:
        req = reqs[rid]
        res = srv.resident[req.geometry]
        g = res.geometry
        _, ref = _dense_reference(g, res.model, [req.energy], [req.eta])
        bs = list(g.block_sizes)
        m = tridiagonal_mask(bs)
        for k, (X, r) in enumerate(zip(srv.kept[rid], ref)):
            xd = g.to_dense(*(x[None].cpu().numpy() if x is not None else None for x in X))
            mask = m if k == 0 else (m & _block_upper(bs))
            err = block_rel_errors(xd * mask, r * mask, bs)
            assert err <= TOL, (rid, k, err)


def _block_upper(bs):
    off = np.concatenate(([0], np.cumsum(bs)))
    n = int(off[-1])
    m = np.zeros((n, n), bool)
    for i in range(len(bs)):
        for j in range(i, len(bs)):
            m[off[i]:off[i + 1], off[j]:off[j + 1]] = True
    return m


# --------------------------------------------------------------------------- uneven batching


def _interleaved(geoms, per_geometry, eta=1e-3, seed=W.ENERGY_SEED):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(per_geometry):
        for gid in range(len(geoms)):
            reqs.append(SV.Request(0.0, len(reqs), gid, float(rng.uniform(-1, 1)), eta))
    return reqs


def test_uneven_four_geometries_reduced_n_vs_dense(device, tmp_path, monkeypatch):
    """The uneven workload's block sizes (128, 256, 384, 512) at two or three
    blocks each, interleaved through one server with captured sweeps and
    autotuned tiles, against the dense oracle."""
    monkeypatch.setenv("QX_TILE_CACHE", str(tmp_path / "tiles.json"))
    geoms = [Geometry.uniform(3, 128), Geometry.uniform(2, 256), Geometry.uniform(2, 384),
             Geometry.uniform(2, 512)]
    seeds = [W.DEVICE_SEED_BASE + 40 + k for k in range(4)]
    reqs = _interleaved(geoms, 3)
    keep = set(range(len(reqs)))
    srv = SV.Server(geoms, seeds, budget_bytes=16 << 30, capacity=2, slots_per_geometry=1,
                    precision="fp64", keep_outputs=keep, autotune=True, device=device)
    try:
        assert all(len(s.graphs) == s.capacity for s in srv.slots)
        results = srv.run(reqs, realtime=False)
        assert sorted(r.rid for r in results) == list(range(len(reqs)))
        _kept_dense(srv, reqs, keep)
    finally:
        srv.close()


@pytest.mark.slow
def test_uneven_four_geometries_full_size(device, tmp_path, monkeypatch):
    """(16, 128), (32, 256), (24, 384), (12, 512) interleaved round-robin
    through the server (graphs, autotuned tiles): the block-row identities
    of one energy of every geometry."""
    import torch

    monkeypatch.setenv("QX_TILE_CACHE", str(tmp_path / "tiles.json"))
    geoms = [Geometry.uniform(n, b) for n, b in UNEVEN]
    seeds = [W.DEVICE_SEED_BASE + 50 + k for k in range(4)]
    reqs = _interleaved(geoms, 3)
    keep = {4, 5, 6, 7}  # the second energy of every geometry
    srv = SV.Server(geoms, seeds, budget_bytes=48 << 30, capacity=3, slots_per_geometry=1,
                    precision="fp64", keep_outputs=keep, autotune=True, device=device)
    try:
        assert set(srv.tiles) == {f"f64:{b}" for _, b in UNEVEN}
        results = srv.run(reqs, realtime=False)
        assert sorted(r.rid for r in results) == list(range(len(reqs)))
        worst = _server_identities(srv, reqs, keep)
        print("uneven full-size identity residuals (R, <, >):", worst)
        assert max(worst) <= IDENTITY_TOL, worst
    finally:
        srv.close()
        torch.cuda.empty_cache()


# --------------------------------------------------------------------------- conditioning


ETAS = (1e-2, 1e-3, 1e-4)


def _eta_case(device, nb, b, n_e, seed):
    import torch

    g = Geometry.uniform(nb, b)
    model = W.make_device(g, seed)
    en = W.draw_energies(n_e)
    etas = np.array([ETAS[k % 3] for k in range(n_e)])
    A = W.system_matrix_torch(model, en, etas, device)
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=device)  # noqa: E731
    sld, slu, sgd, sgu = W.sigmas(model, n_e)
    return g, model, en, etas, A, (t(sld), t(slu)), (t(sgd), t(sgu))


def _solve(device, g, A, SL, SG, **kw):
    E = int(A[0].shape[0])
    GR = S.packed_empty(g, E, device)
    GL = S.packed_empty(g, E, device, ("d", "u"))
    GG = S.packed_empty(g, E, device, ("d", "u"))
    S.selected_solve(g, A, SL, SG, g_lesser=GL, g_greater=GG, g_retarded=GR, **kw)
    S.default_context(device).status()
    return GR, GL, GG


def test_eta_cycle_reduced_n_vs_dense(device):
    """Block size 256 at 6 blocks, eta 1e-2 / 1e-3 / 1e-4: FP64 against the
    dense oracle at <= 1e-10 per block for every eta."""
    g, model, en, etas, A, SL, SG = _eta_case(device, 6, 256, 6, W.DEVICE_SEED_BASE + 4)
    GR, GL, GG = _solve(device, g, A, SL, SG)
    _, ref = _dense_reference(g, model, en, etas)
    bs = list(g.block_sizes)
    m = tridiagonal_mask(bs)
    mu = m & _block_upper(bs)
    errs = []
    for X, r, mask in zip((GR, GL, GG), ref, (m, mu, mu)):
        parts = tuple(x.cpu().numpy() if x is not None else None for x in X)
        xd = g.to_dense(*parts)
        errs.append(block_rel_errors(xd * mask, r * mask, bs))
    assert max(errs) <= TOL, errs


@pytest.mark.slow
def test_eta_cycle_48x256_fp64_identities(device):
    """The conditioning workload's geometry (48 x 256) with eta cycled over
    1e-2, 1e-3, 1e-4: every energy's block-row identities in FP64."""
    import torch

    g, model, en, etas, A, SL, SG = _eta_case(device, 48, 256, 3, W.DEVICE_SEED_BASE + 3)
    GR, GL, GG = _solve(device, g, A, SL, SG)
    for e in range(3):
        one = lambda X: tuple(x[e:e + 1] if x is not None else None for x in X)  # noqa: E731
        w = identity_residuals(g, one(A), one(SL), one(SG), one(GR), one(GL), one(GG))
        print(f"48x256 eta={etas[e]:g} FP64 identity residuals (R, <, >):", w)
        assert max(w) <= IDENTITY_TOL, (etas[e], w)
    torch.cuda.empty_cache()


@pytest.mark.slow
def test_eta_cycle_48x256_mixed_accuracy_and_fallback(device):
    """The complex64-product path at 48 x 256: its error against FP64 per
    output at each eta (the accuracy it delivers: G^R about 1e-5 - 1e-3,
    G^<> about 1e-5 - 1e-1, growing as eta falls), the FP64 check flagging
    every energy at the 1e-8 tolerance, and the fallback then returning the
    FP64 results bit for bit."""
    import torch

    g, model, en, etas, A, SL, SG = _eta_case(device, 48, 256, 3, W.DEVICE_SEED_BASE + 3)
    ref = _solve(device, g, A, SL, SG)
    unchecked = _solve(device, g, A, SL, SG, precision="mixed",
                       mixed=S.MixedOptions(refine=1, fallback_tol=0.0))
    bound = {1e-2: (1e-4, 2e-4), 1e-3: (1e-3, 5e-3), 1e-4: (1e-2, 3e-1)}
    for e in range(3):
        errs = []
        for X, R in zip(unchecked, ref):
            num = sum(float(torch.linalg.norm(x[e] - r[e]) ** 2) for x, r in zip(X, R) if x is not None)
            den = sum(float(torch.linalg.norm(r[e]) ** 2) for r in R if r is not None)
            errs.append((num / den) ** 0.5)
        print(f"48x256 eta={etas[e]:g} mixed error vs FP64 (R, <, >):", errs)
        br, bl = bound[etas[e]]
        assert errs[0] <= br and max(errs[1:]) <= bl, (etas[e], errs)
        assert max(errs) > 1e-8, "complex64 products cannot be this close to FP64"
    flags = torch.zeros(3, dtype=torch.int32, device=device)
    resid = torch.zeros(3, dtype=torch.float64, device=device)
    checked = _solve(device, g, A, SL, SG, precision="mixed",
                     mixed=S.MixedOptions(refine=1, fallback_tol=1e-8, fallback_flags=flags,
                                          residuals=resid))
    assert flags.cpu().tolist() == [1, 1, 1], resid.cpu().tolist()
    for X, R in zip(checked, ref):
        for x, r in zip(X, R):
            if x is not None:
                assert torch.equal(x, r)
    torch.cuda.empty_cache()


# --------------------------------------------------------------------------- serving


def _alone(srv, req):
    """One request in a fresh one-energy FP64 solve on the default stream."""
    import torch

    res = srv.resident[req.geometry]
    g = res.geometry
    dev = srv.device
    z = torch.tensor([req.energy + 1j * req.eta], dtype=torch.complex128, device=dev)
    A = S.packed_empty(g, 1, dev)
    S.assemble(g, z, res.h, a=A, sigma_b=res.sigma_b)
    SL = (res.sl_diag[None].contiguous(), res.sl_upper[None].contiguous())
    SG = (res.sg_diag[None].contiguous(), res.sg_upper[None].contiguous())
    return _solve(dev, g, A, SL, SG)


@pytest.mark.slow
def test_serving_six_geometries_alone_equals_mix(device):
    """The serving benchmark's six resident geometries (drawn with its seed:
    n_blocks 8-64, blocks 64-512), requests of all six interleaved and
    overlapping on the device (two slots per geometry, captured sweeps): every
    kept result bit for bit equal to the request solved alone, and the
    identities of one energy per geometry."""
    import torch

    geos = SV.draw_geometries(6, W.GEOMETRY_SEED)
    assert {g.max_block for g in geos} <= {64, 128, 256, 384, 512}
    seeds = [W.GEOMETRY_SEED + 1 + k for k in range(len(geos))]
    rng = np.random.default_rng(11)
    reqs = [SV.Request(0.0, k, k % 6, float(rng.uniform(-1, 1)), 1e-3) for k in range(24)]
    keep = set(range(6, 12))  # one energy of every geometry
    srv = SV.Server(geos, seeds, budget_bytes=64 << 30, capacity=2, slots_per_geometry=2,
                    precision="fp64", keep_outputs=keep, device=device)
    try:
        results = srv.run(reqs, realtime=False)
        assert sorted(r.rid for r in results) == list(range(len(reqs)))
        for rid in keep:
            alone = _alone(srv, reqs[rid])
            for X, R in zip(srv.kept[rid], alone):
                for part, rpart in zip(X, R):
                    if part is not None and rpart is not None:
                        assert torch.equal(part, rpart[0]), rid
        worst = _server_identities(srv, reqs, keep)
        print("serving six geometries identity residuals (R, <, >):", worst)
        assert max(worst) <= IDENTITY_TOL, worst
    finally:
        srv.close()
        torch.cuda.empty_cache()
