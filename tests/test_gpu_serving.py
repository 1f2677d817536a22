"""Serving loop on the GPU: concurrent batches of several geometries, and each
request's result equal to the same request solved alone."""

import numpy as np
import pytest

from quatrex_amd import solve as S
from quatrex_amd.layout import Geometry
from quatrex_amd.serving import AdmissionError, Request, Server, latency_stats

pytestmark = pytest.mark.gpu

GEOS = [Geometry.uniform(6, 32), Geometry.uniform(4, 48), Geometry.uniform(5, 64)]
SEEDS = [51, 52, 53]


def _alone(server, gid, req, precision):
    """Solve one request in a fresh single-energy batch on the default stream."""
    import torch

    res = server.resident[gid]
    g = res.geometry
    dev = server.device
    z = torch.tensor([req.energy + 1j * req.eta], dtype=torch.complex128, device=dev)
    A = S.packed_empty(g, 1, dev)
    S.assemble(g, z, res.h, a=A, sigma_b=res.sigma_b)
    GR, GL, GG = S.packed_empty(g, 1, dev), S.packed_empty(g, 1, dev, ("d", "u")), S.packed_empty(g, 1, dev, ("d", "u"))
    SL = (res.sl_diag[None].contiguous(), res.sl_upper[None].contiguous())
    SG = (res.sg_diag[None].contiguous(), res.sg_upper[None].contiguous())
    if precision == "fp64":
        S.selected_solve(g, A, SL, SG, g_lesser=GL, g_greater=GG, g_retarded=GR)
    else:
        S.selected_solve(g, A, SL, SG, g_lesser=GL, g_greater=GG, g_retarded=GR, precision="mixed",
                         mixed=S.MixedOptions(refine=1, fallback_tol=1e-8))
    S.default_context().status()
    return GR, GL, GG


@pytest.mark.parametrize("precision", ["fp64", "adaptive"])
def test_results_same_alone_and_in_the_mix(device, precision):
    import torch

    rng = np.random.default_rng(9)
    n = 36
    reqs = [Request(0.0, k, int(rng.integers(0, 3)), float(rng.uniform(-1, 1))) for k in range(n)]
    keep = {0, 5, 11, 17, 23, 35}
    srv = Server(GEOS, SEEDS, budget_bytes=8 << 30, capacity=4, slots_per_geometry=2,
                 precision=precision, keep_outputs=keep, device=device)
    try:
        results = srv.run(reqs, realtime=False)
        assert sorted(r.rid for r in results) == list(range(n))
        st = latency_stats(results)
        assert st["batches"] >= n // 4
        for rid in keep:
            req = reqs[rid]
            got = srv.kept[rid]
            ref = _alone(srv, req.geometry, req, "fp64" if precision == "fp64" else "mixed")
            for X, R in zip(got, ref):
                for part, rpart in zip(X, R):
                    if part is None or rpart is None:
                        continue
                    r = rpart[0]
                    if precision == "fp64":
                        assert torch.equal(part, r)
                    else:
                        err = float(torch.linalg.norm(part - r) / torch.linalg.norm(r))
                        assert err <= 1e-8, (rid, err)
    finally:
        srv.close()


def test_admission_by_footprint(device):
    with pytest.raises(AdmissionError):
        Server([Geometry.uniform(32, 256)], [1], budget_bytes=64 << 20, capacity=4,
               slots_per_geometry=1, device=device)


def test_slots_per_geometry_list_and_queue_warning(device, monkeypatch):
    """A slot count per geometry; more slot streams than hardware queues warns."""
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "4")
    with pytest.warns(RuntimeWarning, match="hardware queues"):
        srv = Server(GEOS, SEEDS, budget_bytes=8 << 30, capacity=2, slots_per_geometry=[1, 3, 1],
                     precision="fp64", use_graphs=False, device=device)
    try:
        assert [sum(s.gid == k for s in srv.slots) for k in range(3)] == [1, 3, 1]
        res = srv.run([Request(0.0, k, k % 3, 0.05 * k, 1e-3) for k in range(9)], realtime=False)
        assert sorted(r.rid for r in res) == list(range(9))
    finally:
        srv.close()
    with pytest.raises(ValueError):
        Server(GEOS, SEEDS, budget_bytes=8 << 30, capacity=2, slots_per_geometry=[1, 1], device=device)


def test_realtime_poisson_stream(device):
    from quatrex_amd.serving import poisson_requests

    reqs = poisson_requests(60, 400.0, 3, seed=7)
    srv = Server(GEOS, SEEDS, budget_bytes=8 << 30, capacity=4, slots_per_geometry=2,
                 precision="adaptive", device=device)
    try:
        results = srv.run(reqs, realtime=True)
        st = latency_stats(results)
        assert st["requests"] == 60
        assert all(r.done >= r.arrival for r in results)
        assert st["p99_ms"] >= st["p50_ms"] > 0
    finally:
        srv.close()


def test_autotuned_uneven_batching_matches_alone(device, tmp_path, monkeypatch):
    """Geometries of different block sizes in one server with the tile per block
    size picked by the autotuner: every kept result equals the request solved
    alone with the default tiles (to rounding: the tile shape may differ)."""
    import torch

    monkeypatch.setenv("QX_TILE_CACHE", str(tmp_path / "tiles.json"))
    n = 24
    reqs = [Request(0.0, k, k % 3, -0.9 + 0.07 * k) for k in range(n)]
    keep = {0, 1, 2, 13, 22, 23}
    srv = Server(GEOS, SEEDS, budget_bytes=8 << 30, capacity=4, slots_per_geometry=1,
                 precision="fp64", keep_outputs=keep, autotune=True, device=device)
    try:
        assert set(srv.tiles) == {"f64:32", "f64:48", "f64:64"}
        results = srv.run(reqs, realtime=False)
        assert sorted(r.rid for r in results) == list(range(n))
        for rid in keep:
            req = reqs[rid]
            ref = _alone(srv, req.geometry, req, "fp64")
            for X, R in zip(srv.kept[rid], ref):
                for part, rpart in zip(X, R):
                    if part is None or rpart is None:
                        continue
                    r = rpart[0]
                    err = float(torch.linalg.norm(part - r) / torch.linalg.norm(r))
                    assert err <= 1e-12, (rid, err)
    finally:
        srv.close()


def test_serving_returns_currents(device):
    """The server with currents: the boundary enters as OBC blocks, graphs
    record the current, and every kept request's G^R and layer currents match
    the dense oracle (the current by brute force from left-connected dense
    inverses)."""
    from oracle.dense import dense_current, dense_solve
    from quatrex_amd import workloads as W
    from tests.helpers import block_rel_errors, tridiagonal_mask

    geos = [Geometry((24, 40, 16, 33, 24)), Geometry.uniform(4, 48)]
    seeds = [61, 62]
    n = 10
    reqs = [Request(0.0, k, k % 2, -0.8 + 0.17 * k, 1e-2) for k in range(n)]
    keep = set(range(n))
    srv = Server(geos, seeds, budget_bytes=8 << 30, capacity=3, slots_per_geometry=1,
                 precision="fp64", keep_outputs=keep, current=True, device=device)
    try:
        assert all(len(s.graphs) == s.capacity for s in srv.slots)
        results = srv.run(reqs, realtime=False)
        assert sorted(r.rid for r in results) == list(range(n))
        for rid in keep:
            req = reqs[rid]
            res = srv.resident[req.geometry]
            g = res.geometry
            GR, GL, GG, (cur,) = srv.kept[rid]
            A = W.system_matrix(res.model, [req.energy], [req.eta], fold_boundary=False)
            a = g.to_dense(*A)
            sl = g.to_dense(res.model.sl_diag[None], None, None)
            ends = W.boundary_blocks(g)
            obc = [[None] * g.n_blocks for _ in range(3)]
            for i in ends:
                sb = W.BOUNDARY_SIGMA * np.eye(g.block_sizes[i], dtype=np.complex128)[None]
                obc[0][i], obc[1][i], obc[2][i] = sb, 0 * sb, -(sb - np.conj(np.swapaxes(sb, -1, -2)))
            ref = dense_solve(a, sl, sl, g.block_offsets, obc)
            got = g.to_dense(*(x[None].cpu().numpy() for x in GR))
            err = block_rel_errors(got, ref[0] * tridiagonal_mask(list(g.block_sizes)), list(g.block_sizes))
            assert err <= 1e-10, (rid, err)
            cd = dense_current(a, sl, sl, g.block_offsets, obc)[0]
            cerr = np.abs(cur.cpu().numpy() - cd).max() / np.abs(cd).max()
            assert cerr <= 1e-10, (rid, cerr)
    finally:
        srv.close()
