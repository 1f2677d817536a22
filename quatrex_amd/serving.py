"""Continuous-batching RGF serving loop on one MI355X.

Requests are single energy points for one of several resident device
geometries.  Each geometry keeps its Hamiltonian, boundary self-energy and
scattering self-energies resident in HBM; a geometry owns one or more
*slots*: a HIP stream with its own library context and fixed device buffers
for up to ``capacity`` energies (A, Sigma^<, Sigma^>, outputs, workspace).
Slots are created only if their HBM footprint fits the server budget
(admission by footprint); a request is admitted into a slot's next batch.

A dispatcher thread moves arrived requests into per-geometry queues and
hands a batch (the oldest pending requests of one geometry, up to the slot
capacity) to any idle slot; one worker thread per slot assembles A on the
device (``rgf_assemble_z``) and runs the selected solve on the slot's stream.
ctypes releases the GIL during library calls, so slots of different
geometries - and several slots of one geometry - run concurrently on the
device.  A request's latency is the time from its arrival to the end of the
solve that produced it (results resident in HBM).

Streams and hardware queues: HIP multiplexes a process's streams onto
``GPU_MAX_HW_QUEUES`` hardware queues (4 by default), and kernels in one
hardware queue run in order, so a small geometry's short kernels can wait
behind another slot's long products.  Each slot uses one stream (its
context's inversion overlap, which adds a second stream per slot, is off by
default: the slots already overlap one another); the launcher should set
``GPU_MAX_HW_QUEUES`` to at least the number of slots before the process
starts (the server warns otherwise).  Measured on the six-geometry
workload: 97 -> 130 requests/s at saturation and a 6x lower p50 below
capacity with 16 queues and one stream per slot; 24 or 32 queues lowered the
saturated throughput again (more kernels in flight than the CUs serve well).

Currents: with ``current=True`` every result also carries the N+1
Meir-Wingreen layer currents.  The boundary then enters as OBC blocks
instead of being folded into A: B^R = Sigma_B on the first and last block,
and the boundary part of the generator's Sigma^> (Sigma^> = Sigma^< -
(Sigma_B - Sigma_B^H) there) as B^> with B^< = 0, so the Green's functions are
those of the folded workload and J_0, J_N are the currents into the two
boundaries.

FP64 batches replay a hipGraph of the whole sweep, captured for every batch
size of a slot when the server starts (serially, before any batch runs; the
graphs share the slot's workspace) and replayed on the slot's buffers (A is
re-assembled in place before each replay), which removes the per-launch host
cost of small geometries.

Precision: FP64 is the default.  ``mixed`` runs the complex64-product sweep
with the FP64 check and FP64 fallback above ``fallback_tol``; ``adaptive``
additionally sends a geometry straight to FP64 while its recent fallback rate
is high (the mixed attempt would be wasted), retrying mixed after
``retry_every`` batches, the interval doubling after every failed retry (up
to 1024 batches).  At a 1e-8 tolerance complex64 products fail the check for
every energy of these workloads (their G^<> error is 1e-5 to 1e-1 against
FP64, growing as eta falls), so ``mixed`` then costs the check on top of the
FP64 solve; the modes are kept for looser tolerances.
"""

from __future__ import annotations

import heapq
import os
import queue
import threading
import time
import warnings
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from quatrex_amd import _lib
from quatrex_amd import solve as S
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry


class AdmissionError(RuntimeError):
    """A geometry's slots do not fit the server's HBM budget."""


@dataclass(order=True)
class Request:
    arrival: float
    rid: int = field(compare=False)
    geometry: int = field(compare=False)
    energy: float = field(compare=False)
    eta: float = field(compare=False, default=1e-3)


@dataclass
class Result:
    rid: int
    geometry: int
    arrival: float
    done: float
    precision: str
    fallback: bool
    batch: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def slot_bytes(geometry: Geometry, capacity: int, current: bool = False) -> int:
    """HBM bytes of one slot: I/O buffers for ``capacity`` energies + workspace."""
    g = geometry
    per_e = 16 * (3 * g.diag_size + 3 * g.offdiag_size      # A
                  + 2 * (g.diag_size + g.offdiag_size)       # Sigma^<, Sigma^>
                  + 3 * g.diag_size + 2 * g.offdiag_size     # G^R diag, upper, lower
                  + 2 * (g.diag_size + g.offdiag_size))      # G^<, G^> diag, upper
    per_e += 16 + 8 + 4  # z, residual, flag
    if current:
        ends = [g.block_sizes[i] for i in W.boundary_blocks(g)]
        per_e += 16 * (g.n_blocks + 1 + 3 * sum(b * b for b in ends))  # current, OBC blocks
    ws = S.workspace_size(g, capacity, solve=True, current=current, retarded=True)
    ws += 2 * 16 * g.offdiag_size * capacity + 64 * 8 * capacity + 16 * g.max_block ** 2 + 4096
    return per_e * capacity + ws


class Slot:
    def __init__(self, server: "Server", gid: int, capacity: int, device, overlap: bool = False):
        import torch

        self.server = server
        self.gid = gid
        self.capacity = capacity
        self.device = device
        res = server.resident[gid]
        g = res.geometry
        self.geometry = g
        self.stream = torch.cuda.Stream(device)  # default priority
        self.ctx = S.Context(device, self.stream)
        self.ctx.set_option(_lib.RGF_OPT_OVERLAP, 1 if overlap else 0)
        C = capacity
        with torch.cuda.stream(self.stream):
            self.z = torch.zeros(C, dtype=torch.complex128, device=device)
            self.A = S.packed_empty(g, C, device)
            self.SL = (res.sl_diag.expand(C, -1).contiguous(), res.sl_upper.expand(C, -1).contiguous())
            self.SG = (res.sg_diag.expand(C, -1).contiguous(), res.sg_upper.expand(C, -1).contiguous())
            self.GR = S.packed_empty(g, C, device)
            self.GL = S.packed_empty(g, C, device, parts=("d", "u"))
            self.GG = S.packed_empty(g, C, device, parts=("d", "u"))
            self.resid = torch.zeros(C, dtype=torch.float64, device=device)
            self.flags = torch.zeros(C, dtype=torch.int32, device=device)
            self.cur = None
            self.obc = None
            if server.current:
                # the boundary as OBC blocks (module notes); Sigma^> without its boundary part
                self.cur = torch.zeros((C, g.n_blocks + 1), dtype=torch.complex128, device=device)
                self.obc = ([None] * g.n_blocks, [None] * g.n_blocks, [None] * g.n_blocks)
                for i in W.boundary_blocks(g):
                    sb = res.sigma_b[i]
                    self.obc[0][i] = sb.expand(C, -1, -1).contiguous()
                    self.obc[1][i] = torch.zeros_like(self.obc[0][i])
                    self.obc[2][i] = (-(sb - sb.conj().T)).expand(C, -1, -1).contiguous()
                self.SG = (res.sl_diag.expand(C, -1).contiguous(), res.sg_upper.expand(C, -1).contiguous())
            nbytes = S.workspace_size(g, C, solve=True, current=server.current) \
                + 2 * 16 * g.offdiag_size * C + 64 * 8 * C + 16 * g.max_block ** 2 + 4096
            self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.stream.synchronize()
        self.ctx.set_workspace(self.workspace)
        self.graphs: dict = {}
        self.inbox: queue.Queue = queue.Queue()
        self.busy = False
        self.thread = threading.Thread(target=self._work, daemon=True)
        self.thread.start()

    def _work(self):
        import torch

        while True:
            item = self.inbox.get()
            if item is None:
                return
            batch, precision = item
            try:
                fb = self._solve(batch, precision)
                done = time.perf_counter() - self.server.t0
                self.server._complete(self, batch, done, precision, fb)
            except BaseException as exc:  # pragma: no cover - surfaced by the dispatcher
                self.server._fail(exc)
                return

    def _solve(self, batch: list, precision: str) -> np.ndarray:
        import torch

        # Every torch call of a worker runs on the slot's stream: the thread's
        # default (legacy) stream would synchronise with, and fail against, a
        # sweep another slot is capturing into a graph.
        with torch.cuda.stream(self.stream):
            return self._solve_on_stream(batch, precision)

    def _solve_on_stream(self, batch: list, precision: str) -> np.ndarray:
        import torch

        n = len(batch)
        g = self.geometry
        res = self.server.resident[self.gid]
        zh = torch.tensor([r.energy + 1j * r.eta for r in batch], dtype=torch.complex128)
        self.z[:n].copy_(zh, non_blocking=False)
        view = lambda T: tuple(t[:n] if t is not None else None for t in T)  # noqa: E731
        A = view(self.A)
        S.assemble(g, self.z[:n], res.h, a=A, sigma_b=None if self.obc else res.sigma_b, ctx=self.ctx)
        common = dict(g_lesser=view(self.GL), g_greater=view(self.GG), g_retarded=view(self.GR),
                      ctx=self.ctx)
        if self.obc is not None:
            common.update(obc=self._obc_view(n), current=self.cur[:n])
        SL, SG = view(self.SL), view(self.SG)
        if precision == "fp64":
            graph = self.graphs.get(n)
            if graph is not None:
                graph.launch()
            else:
                S.selected_solve(g, A, SL, SG, **common)
            self.ctx.status()
            return np.zeros(n, dtype=bool)
        opts = S.MixedOptions(refine=self.server.refine, fallback_tol=self.server.fallback_tol,
                              fallback_flags=self.flags[:n], residuals=self.resid[:n])
        S.selected_solve(g, A, SL, SG, precision="mixed", mixed=opts, **common)
        self.ctx.status()
        return self.flags[:n].cpu().numpy().astype(bool)

    def _obc_view(self, n: int):
        return tuple([t[:n] if t is not None else None for t in lst] for lst in self.obc)

    def prepare_graphs(self) -> None:
        """Capture the FP64 sweep for every batch size 1..capacity (server thread,
        before any batch runs): graphs are never captured while other slots
        launch work, and they share the slot's workspace."""
        g = self.geometry
        view = lambda T, n: tuple(t[:n] if t is not None else None for t in T)  # noqa: E731
        for n in range(1, self.capacity + 1):
            extra = {}
            if self.obc is not None:
                extra = dict(obc=self._obc_view(n), current=self.cur[:n])
            self.graphs[n] = S.SolveGraph(g, view(self.A, n), view(self.SL, n), view(self.SG, n),
                                          g_lesser=view(self.GL, n), g_greater=view(self.GG, n),
                                          g_retarded=view(self.GR, n), ctx=self.ctx, **extra)
        self.stream.synchronize()

    def outputs(self, k: int):
        """Device views of energy ``k`` of the last batch: (G^R, G^<, G^>) packed,
        plus the layer currents when the server computes them."""
        outs = tuple(tuple(t[k] if t is not None else None for t in X) for X in (self.GR, self.GL, self.GG))
        if self.cur is not None:
            outs = outs + ((self.cur[k],),)
        return outs

    def close(self):
        self.inbox.put(None)
        self.thread.join(timeout=60)
        for gph in self.graphs.values():
            gph.close()
        self.graphs.clear()
        self.ctx.close()


@dataclass
class ResidentGeometry:
    geometry: Geometry
    model: W.DeviceModel
    h: tuple
    sl_diag: object
    sl_upper: object
    sg_diag: object
    sg_upper: object
    sigma_b: list


class Server:
    """Continuous-batching server over resident geometries.

    Parameters
    ----------
    geometries : list of Geometry
    seeds : list of int, model seeds per geometry
    budget_bytes : HBM bytes the server may use for slots
    capacity : max energies per batch (per slot; reduced to fit the budget)
    slots_per_geometry : concurrent batches per geometry (one count for all, or
        one per geometry)
    precision : "fp64", "mixed" or "adaptive"
    autotune : time the GEMM tile configurations per block size at start-up
        (cached on disk) and use the fastest in every slot
    overlap : run each slot's block inversions on a second stream (off: one
        stream per slot, see the module notes on hardware queues)
    current : also compute the layer currents of every request (the boundary
        then enters as OBC blocks, see the module notes)
    """

    def __init__(self, geometries, seeds, budget_bytes: int, capacity: int = 8,
                 slots_per_geometry=2, precision: str = "fp64",
                 fallback_tol: float = 1e-8, refine: int = 1, retry_every: int = 8,
                 keep_outputs: set | None = None, use_graphs: bool = True, autotune: bool = False,
                 overlap: bool = False, current: bool = False, device=None):
        import torch

        if precision not in ("fp64", "mixed", "adaptive"):
            raise ValueError(precision)
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        self.precision = precision
        self.current = bool(current)
        self.fallback_tol = fallback_tol
        self.refine = refine
        self.retry_every = retry_every
        self.use_graphs = use_graphs
        self.keep_outputs = keep_outputs or set()
        self.kept: dict = {}
        self.resident = []
        self.slots: list[Slot] = []
        self.budget = int(budget_bytes)
        self.used = 0
        for g, seed in zip(geometries, seeds):
            self.resident.append(self._make_resident(g, seed))
            self.used += self._resident_bytes(g)
        if self.used > self.budget:
            raise AdmissionError(f"resident geometries need {self.used} bytes > budget {self.budget}")
        counts = ([int(slots_per_geometry)] * len(geometries) if np.isscalar(slots_per_geometry)
                  else [int(c) for c in slots_per_geometry])
        if len(counts) != len(geometries) or min(counts) < 1:
            raise ValueError("slots_per_geometry: one positive count, or one per geometry")
        for gid, g in enumerate(geometries):
            cap = capacity
            nslots = counts[gid]
            need = slot_bytes(g, cap, self.current) * nslots
            while cap > 1 and self.used + need > self.budget:
                cap //= 2
                need = slot_bytes(g, cap, self.current) * nslots
            if self.used + need > self.budget:
                raise AdmissionError(
                    f"geometry {gid} ({g.n_blocks} blocks, max block {g.max_block}) needs "
                    f"{need} bytes for {nslots} slot(s); {self.budget - self.used} left")
            for _ in range(nslots):
                self.slots.append(Slot(self, gid, cap, self.device, overlap=overlap))
            self.used += need
        streams = len(self.slots) * (2 if overlap else 1)
        queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4)
        if streams > queues:
            warnings.warn(f"{streams} slot streams share {queues} hardware queues: kernels of "
                          f"different slots wait for one another; start the process with "
                          f"GPU_MAX_HW_QUEUES >= {streams}", RuntimeWarning)
        self.tiles: dict = {}
        if autotune:
            self._autotune(geometries, capacity)
        if use_graphs and precision != "mixed":
            # serially, here: the tile choice is final and no batch is running
            for slot in self.slots:
                slot.prepare_graphs()
        self.recent = [deque(maxlen=4) for _ in geometries]
        self.batches_since_retry = [0] * len(geometries)
        self.retry_gap = [retry_every] * len(geometries)
        self._lock = threading.Condition()
        self._error = None

    def _autotune(self, geometries, capacity: int) -> None:
        """Pick the GEMM tile per block size (and precision) once; install it in every slot."""
        from quatrex_amd import _lib
        from quatrex_amd import autotune as AT

        precs = (False,) if self.precision == "fp64" else (False, True)
        sizes = sorted({int(b) for g in geometries for b in g.block_sizes})
        for b in sizes:
            for f32 in precs:
                entry = AT.tune(self.slots[0].ctx, b, n_energies=capacity, f32=f32)
                self.tiles[f"{'f32' if f32 else 'f64'}:{b}"] = entry["best_name"]
                base = _lib.RGF_OPT_TILE_FOR_BLOCK_F32 if f32 else _lib.RGF_OPT_TILE_FOR_BLOCK
                for slot in self.slots:
                    slot.ctx.set_option(base + b, entry["best"])

    def _resident_bytes(self, g: Geometry) -> int:
        return 16 * (3 * g.diag_size + 3 * g.offdiag_size)

    def _make_resident(self, g: Geometry, seed: int) -> ResidentGeometry:
        import torch

        m = W.make_device(g, seed)
        d = self.device
        t = lambda x: torch.as_tensor(x, device=d)  # noqa: E731
        sb = [None] * g.n_blocks
        for i in W.boundary_blocks(g):
            sb[i] = t(W.BOUNDARY_SIGMA * np.eye(g.block_sizes[i], dtype=np.complex128))
        return ResidentGeometry(g, m, (t(m.h_diag), t(m.h_upper), t(m.h_lower)), t(m.sl_diag),
                                t(m.sl_upper), t(m.sg_diag), t(m.sg_upper), sb)

    # ------------------------------------------------------------------
    def _choose_precision(self, gid: int) -> str:
        if self.precision != "adaptive":
            return self.precision
        hist = self.recent[gid]
        if hist and np.mean(hist) > 0.5:
            self.batches_since_retry[gid] += 1
            if self.batches_since_retry[gid] < self.retry_gap[gid]:
                return "fp64"
            self.batches_since_retry[gid] = 0
        return "mixed"

    def _record_mixed(self, gid: int, rate: float) -> None:
        """Fallback rate of a mixed batch; a failed attempt doubles the wait before the next."""
        self.recent[gid].append(rate)
        self.retry_gap[gid] = min(2 * self.retry_gap[gid], 1024) if rate > 0.5 else self.retry_every

    def _complete(self, slot: Slot, batch, done, precision, fallback):
        import torch

        # Copies of kept outputs are made on the slot's stream and finished
        # before the slot is released: its next batch overwrites the buffers.
        kept = {}
        if any(r.rid in self.keep_outputs for r in batch):
            with torch.cuda.stream(slot.stream):
                for k, r in enumerate(batch):
                    if r.rid in self.keep_outputs:
                        kept[r.rid] = tuple(tuple(t.clone() if t is not None else None for t in X)
                                            for X in slot.outputs(k))
            slot.stream.synchronize()
        with self._lock:
            self.kept.update(kept)
            for k, r in enumerate(batch):
                self.results.append(Result(r.rid, r.geometry, r.arrival, done, precision,
                                           bool(fallback[k]), self.nbatches))
            if precision == "mixed":
                self._record_mixed(slot.gid, float(np.mean(fallback)) if len(fallback) else 0.0)
            self.nbatches += 1
            slot.busy = False
            self._lock.notify_all()

    def _fail(self, exc):
        with self._lock:
            self._error = exc
            self._lock.notify_all()

    def run(self, requests: list[Request], realtime: bool = True) -> list[Result]:
        """Serve ``requests`` (sorted by arrival); returns per-request results."""
        reqs = sorted(requests)
        pending = [deque() for _ in self.resident]
        self.results: list[Result] = []
        self.nbatches = 0
        self.t0 = time.perf_counter()
        i = 0
        n = len(reqs)
        with self._lock:
            while len(self.results) < n:
                if self._error is not None:
                    raise self._error
                now = time.perf_counter() - self.t0
                while i < n and (not realtime or reqs[i].arrival <= now):
                    r = reqs[i]
                    if not realtime:
                        r.arrival = now
                    pending[r.geometry].append(r)
                    i += 1
                dispatched = False
                for slot in self.slots:
                    q = pending[slot.gid]
                    if slot.busy or not q:
                        continue
                    batch = [q.popleft() for _ in range(min(slot.capacity, len(q)))]
                    slot.busy = True
                    slot.inbox.put((batch, self._choose_precision(slot.gid)))
                    dispatched = True
                if dispatched:
                    continue
                wait = None
                if i < n:
                    wait = max(0.0, reqs[i].arrival - (time.perf_counter() - self.t0))
                self._lock.wait(timeout=wait if wait is not None else 1.0)
        return self.results

    def close(self):
        for s in self.slots:
            s.close()


def poisson_requests(n: int, rate: float, n_geometries: int, seed: int = W.ARRIVAL_SEED,
                     eta: float = 1e-3) -> list[Request]:
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    geo = rng.integers(0, n_geometries, size=n)
    en = rng.uniform(-1.0, 1.0, size=n)
    return [Request(float(t[k]), k, int(geo[k]), float(en[k]), eta) for k in range(n)]


def draw_geometries(n: int, seed: int, n_range=(8, 64), sizes=(64, 128, 256, 384, 512)):
    """Resident geometries: N ~ U{n_range}, b ~ U(sizes)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nb = int(rng.integers(n_range[0], n_range[1] + 1))
        b = int(rng.choice(sizes))
        out.append(Geometry.uniform(nb, b))
    return out


def latency_stats(results: list[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    done = max(r.done for r in results)
    first = min(r.arrival for r in results)
    return {
        "requests": len(results),
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "throughput_rps": float(len(results) / max(done - first, 1e-9)),
        "fallback_fraction": float(np.mean([r.fallback for r in results])),
        "mixed_fraction": float(np.mean([r.precision == "mixed" for r in results])),
        "batches": len({r.batch for r in results}),
    }
