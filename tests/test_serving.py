"""Host-side logic of the serving loop (no GPU)."""

import numpy as np

from quatrex_amd.layout import Geometry
from quatrex_amd.serving import Request, Result, draw_geometries, latency_stats, poisson_requests, slot_bytes


def test_poisson_arrivals_are_seeded_and_have_the_rate():
    a = poisson_requests(20000, 2000.0, 6)
    b = poisson_requests(20000, 2000.0, 6)
    assert [r.arrival for r in a[:50]] == [r.arrival for r in b[:50]]
    t = np.array([r.arrival for r in a])
    assert np.all(np.diff(t) > 0)
    assert abs(len(a) / t[-1] - 2000.0) < 60.0
    geos = np.bincount([r.geometry for r in a], minlength=6)
    assert geos.min() > 3000
    assert all(-1 <= r.energy <= 1 for r in a)


def test_geometry_draws():
    gs = draw_geometries(6, 20265)
    assert len(gs) == 6
    for g in gs:
        assert 8 <= g.n_blocks <= 64 and g.block_sizes[0] in (64, 128, 256, 384, 512)
    assert [g.block_sizes for g in gs] == [g.block_sizes for g in draw_geometries(6, 20265)]


def test_slot_footprint_grows_with_capacity():
    g = Geometry.uniform(16, 128)
    one, four = slot_bytes(g, 1), slot_bytes(g, 4)
    assert four > 3.5 * one
    # one energy's A alone is (3N-2) blocks of 128^2 complex128
    assert one > (3 * 16 - 2) * 128 * 128 * 16


def test_latency_stats():
    res = [Result(k, 0, 0.001 * k, 0.001 * k + 0.01 * (k % 10 + 1), "fp64", False, k // 4)
           for k in range(100)]
    st = latency_stats(res)
    assert st["requests"] == 100
    assert 50 <= st["p50_ms"] <= 60 and 95 <= st["p99_ms"] <= 100
    assert st["fallback_fraction"] == 0.0 and st["batches"] == 25
    assert Request(0.5, 1, 0, 0.1) < Request(0.6, 0, 0, 0.1)


def test_adaptive_precision_backs_off_exponentially():
    """A geometry whose mixed batches keep failing the check is sent to FP64,
    with mixed retried after 8, 16, 32, ... batches; one success resets it."""
    from collections import deque

    from quatrex_amd.serving import Server

    srv = Server.__new__(Server)
    srv.precision = "adaptive"
    srv.retry_every = 8
    srv.recent = [deque(maxlen=4)]
    srv.batches_since_retry = [0]
    srv.retry_gap = [8]
    assert srv._choose_precision(0) == "mixed"  # no history yet
    srv._record_mixed(0, 1.0)
    tries = []
    for k in range(1, 200):
        p = srv._choose_precision(0)
        if p == "mixed":
            tries.append(k)
            srv._record_mixed(0, 1.0)
    assert np.diff([0] + tries).tolist() == [16, 32, 64]
    srv._record_mixed(0, 0.0)
    srv._record_mixed(0, 0.0)
    srv._record_mixed(0, 0.0)
    assert srv.retry_gap[0] == 8
    assert srv._choose_precision(0) == "mixed"  # history now mostly successful
