"""Drop-in ``RGF`` GF solver backed by the gfx950 HIP library.

Mirrors ``RGF`` of ``src/qttools/greens_function_solver/rgf.py`` (constructor
``:21-23``, ``selected_inv`` ``:25-113``, ``selected_solve`` ``:115-438``): same
signatures, the same energy batching (``get_batches``, ``:60,184-186``), the
same output conventions and exceptions:

* ``out`` given -> returns ``None`` (or the current when ``return_current``);
  ``out=None`` -> new ``(xl, xg[, xr][, current])`` (``rgf.py:430-438``);
* a wrong number of outputs raises ``ValueError`` (``rgf.py:189-194``);
* stored diagonal G^<, G^> are anti-Hermitian projected (``rgf.py:305-310``);
* ``G^<_ji = -G^<_ij^H`` is written only for non-symmetric storage (``:357-359``);
* the current without lesser/greater OBC blocks raises ``TypeError`` (``:416``);
* an exactly singular block raises ``numpy.linalg.LinAlgError`` (numpy's
  ``inv`` behaviour, ``kernels/linalg/inv.py:6-11``).

Supersets (documented deviations): off-diagonal G^R blocks are also written
when the retarded output is non-symmetric (the reference writes only the
diagonal, ``rgf.py:311-312,410-412``); the boundary currents are evaluated per
energy batch (the reference evaluates them once after the loop with the last
batch's blocks, ``rgf.py:414-428``, which only matches for a single batch).

Every arithmetic operation runs in the HIP library; this module only moves
blocks between the caller's containers and packed device buffers.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from quatrex_amd import solve as _solve
from quatrex_amd.dsdb import densify_batch as _gather_dsdb
from quatrex_amd.dsdb import sparsify_batch as _scatter_dsdb
from quatrex_amd.dsdb import writes_every_entry as _dsdb_writes_all
from quatrex_amd.dsdb import structure_of as _structure_of
from quatrex_amd.datastructures import BlockTridiagonal
from quatrex_amd.greens_function_solver.solver import GFSolver, OBCBlocks
from quatrex_amd.layout import Geometry
from quatrex_amd.utils.solvers_utils import get_batches


def _torch():
    import torch

    return torch


def _is_device_container(m) -> bool:
    return isinstance(m, BlockTridiagonal) and m.device.type == "cuda"


def _as_tensor(x, device):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.complex128)
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.complex128).to(device, non_blocking=True)


def _geometry(m) -> Geometry:
    return Geometry(tuple(int(b) for b in np.asarray(m.block_sizes)))


def _gather(m, geo: Geometry, parts: str, device):
    """Packed ``(diag, upper, lower)`` device tensors of a stack view ``m``."""
    torch = _torch()
    if isinstance(m, BlockTridiagonal) and not m.symmetry:
        # the packed storage is already the device layout: used in place on
        # the device, one transfer per part from host memory
        d, u, l = m.packed()
        out = []
        for key, t in zip("dul", (d, u, l)):
            out.append(t.to(device=device, dtype=torch.complex128).contiguous() if key in parts else None)
        return tuple(out)
    blk0 = m.blocks[0, 0]
    lead = tuple(blk0.shape[:-2])
    E = int(np.prod(lead)) if lead else 1
    host = {}
    sizes = {"d": geo.diag_size, "u": geo.offdiag_size, "l": geo.offdiag_size}
    for key in parts:
        host[key] = torch.empty((E, sizes[key]), dtype=torch.complex128)
    for i in range(geo.n_blocks):
        if "d" in parts:
            geo.diag(host["d"], i)[...] = _as_tensor(m.blocks[i, i], "cpu").reshape(E, *geo.diag(host["d"], i).shape[1:])
        if i + 1 < geo.n_blocks:
            if "u" in parts:
                tgt = geo.upper(host["u"], i)
                tgt[...] = _as_tensor(m.blocks[i, i + 1], "cpu").reshape(tgt.shape)
            if "l" in parts:
                tgt = geo.lower(host["l"], i)
                tgt[...] = _as_tensor(m.blocks[i + 1, i], "cpu").reshape(tgt.shape)
    pin = torch.cuda.is_available()
    return tuple(
        (host[k].pin_memory() if pin else host[k]).to(device, non_blocking=True) if k in host else None
        for k in "dul"
    )


def _obc_batch(lst, sl, geo: Geometry, E: int, device):
    out = []
    for i, blk in enumerate(lst):
        if blk is None:
            out.append(None)
            continue
        b = geo.block_sizes[i]
        out.append(_as_tensor(blk[sl], device).reshape(E, b, b).contiguous())
    return out


def _outputs(m, geo: Geometry, E: int, device, parts: str):
    """Device output buffers: the container's own storage when possible."""
    torch = _torch()
    if _is_device_container(m) and m.dtype == torch.complex128:
        d, u, l = m.packed()
        if all(t.is_contiguous() for t in (d, u, l)):
            return (d if "d" in parts else None, u if "u" in parts else None,
                    l if "l" in parts else None), True
    sizes = {"d": geo.diag_size, "u": geo.offdiag_size, "l": geo.offdiag_size}
    return tuple(
        torch.empty((E, sizes[k]), dtype=torch.complex128, device=device) if k in parts else None
        for k in "dul"
    ), False


def _scatter(m, geo: Geometry, bufs, lead, blocks=("d", "u", "l")):
    """Write packed device results into a generic container view ``m``."""
    if isinstance(m, BlockTridiagonal) and not m.symmetry:
        # packed storage: one copy per part straight into it (a view is
        # required: a reshaped copy would not reach the container)
        try:
            dsts = [t.view(-1, t.shape[-1]) for t in m.packed(flatten=False)]
        except RuntimeError:
            dsts = None
        if dsts is not None:
            for key, dst, src in zip("dul", dsts, bufs):
                if key in blocks and src is not None:
                    dst.copy_(src.reshape(dst.shape))
            return
    d, u, l = (t.cpu() if t is not None else None for t in bufs)
    is_torch = isinstance(m, BlockTridiagonal)

    def conv(t):
        return t if is_torch else t.numpy()

    for i in range(geo.n_blocks):
        if d is not None and "d" in blocks:
            m.blocks[i, i] = conv(geo.diag(d, i).reshape(*lead, *geo.diag(d, i).shape[1:]))
        if i + 1 < geo.n_blocks:
            if u is not None and "u" in blocks:
                blk = geo.upper(u, i)
                m.blocks[i, i + 1] = conv(blk.reshape(*lead, *blk.shape[1:]))
            if l is not None and "l" in blocks:
                blk = geo.lower(l, i)
                m.blocks[i + 1, i] = conv(blk.reshape(*lead, *blk.shape[1:]))


def _is_dsdb(m) -> bool:
    """A block-sparse stack in the DSDBCOO (rows, cols) or DSDBCSR (rowptr_map,
    cols) layout, ``data[*stack, nnz]`` (dsdbcoo.py:23-120, dsdbcsr.py), or a
    substack view of one: the reference's ``_DStackView``
    (dsdbsparse.py:1118-1205), which ``ElectronSolver.solve`` passes for
    Sigma^<, Sigma^> and ``out`` (electron/solver.py:821-874), keeps the
    structure in its parent (``_dsdbsparse``) and exposes the substack's
    ``data``."""
    if not (hasattr(m, "data") and hasattr(m, "block_sizes")):
        return False
    s = _structure_of(m)
    return hasattr(s, "cols") and (hasattr(s, "rows") or isinstance(getattr(s, "rowptr_map", None), dict))


def _zeros_current(like_container, shape):
    torch = _torch()
    if isinstance(like_container, BlockTridiagonal):
        return torch.zeros(shape, dtype=torch.complex128, device=like_container.device)
    if _is_dsdb(like_container) and isinstance(like_container.data, torch.Tensor):
        return torch.zeros(shape, dtype=torch.complex128, device=like_container.data.device)
    return np.zeros(shape, dtype=np.complex128)


# Input bytes per transfer chunk of the host-resident pipeline: large enough
# that each chunk's solve keeps the GPU busy, small enough that the staging
# of all slots stays modest (4 energies of 32 x 256: ~0.9 GB in, ~1.2 GB out).
_HOST_CHUNK_BYTES = 1 << 30
# Fewest energies per chunk when block-sparse operands take part.  Their
# chunks also carry the (de)densification launches and packed buffers, and
# a solve of a few energies costs nearly as much as one of eight (the block
# inversions' panel eliminations take one workgroup per matrix), so small
# chunks leave such a pipeline bound by the solves instead of by PCIe: at
# 32 blocks x 256, 64 energies, 4 energies per chunk ran 62 energies/s, 8
# ran 80.7.  Block-tridiagonal stacks keep the byte-sized chunks (4 energies
# there: 88 energies/s, against 83.6 with 8).
_HOST_MIN_CHUNK = 8
# Staging slots: chunks in flight between the host copies, the transfers and
# the solves.
_HOST_SLOTS = 3


def _host_packed(m) -> bool:
    """A host-resident BlockTridiagonal whose packed storage can be copied
    part by part (non-symmetric, complex128, one stack axis)."""
    torch = _torch()
    return (isinstance(m, BlockTridiagonal) and m.device.type == "cpu" and not m.symmetry
            and m.dtype == torch.complex128 and len(m.stack_shape) == 1
            and all(t.is_contiguous() for t in m.packed(flatten=False)))


def _host_dsdb(m) -> bool:
    """A block-sparse stack whose stored entries are a host array ``(energies,
    nnz)`` of complex128 in C order, numpy or torch (a contiguous substack
    view of one included)."""
    if not _is_dsdb(m):
        return False
    d, torch = m.data, _torch()
    if isinstance(d, torch.Tensor):
        return (d.device.type == "cpu" and d.dtype == torch.complex128 and d.dim() == 2
                and d.is_contiguous())
    return (isinstance(d, np.ndarray) and d.dtype == np.complex128 and d.ndim == 2
            and d.flags.c_contiguous)


def _host_operand(m) -> bool:
    return _host_packed(m) or _host_dsdb(m)


class _HostPipe:
    """Staging of one selected-solve batch of host-resident stacks in energy
    chunks, so that the host copies, the PCIe transfers and the solves of
    consecutive chunks overlap.

    Per chunk c (slot s = c % slots): the host copies the inputs from the
    caller's pageable tensors into pinned buffers s; a copy stream moves them
    to device buffers s; the solve runs on the context stream into device
    outputs s; a second copy stream moves those to pinned outputs s; a worker
    thread copies them into the caller's outputs while the main thread stages
    the next chunks' inputs.  Events order the reuse of every buffer `slots`
    chunks later.  Results are those of one solve over the batch: energies
    are independent.
    """

    def __init__(self, device, in_sizes, out_sizes, chunk, slots):
        torch = _torch()
        z = torch.complex128
        self.chunk, self.slots = chunk, slots
        self.in_sizes, self.out_sizes = in_sizes, out_sizes
        S = range(slots)
        self.pin_in = [[torch.empty((chunk, n), dtype=z, pin_memory=True) for n in in_sizes] for _ in S]
        self.dev_in = [[torch.empty((chunk, n), dtype=z, device=device) for n in in_sizes] for _ in S]
        self.pin_out = [[torch.empty((chunk, n), dtype=z, pin_memory=True) for n in out_sizes] for _ in S]
        self.dev_out = [[torch.empty((chunk, n), dtype=z, device=device) for n in out_sizes] for _ in S]
        self.copy_in = torch.cuda.Stream(device)
        self.copy_out = torch.cuda.Stream(device)
        self.ev_in = [torch.cuda.Event() for _ in S]
        self.ev_solve = [torch.cuda.Event() for _ in S]
        self.ev_out = [torch.cuda.Event() for _ in S]
        self.pool = ThreadPoolExecutor(max_workers=1)

    def __del__(self):
        pool = getattr(self, "pool", None)
        if pool is not None:
            pool.shutdown(wait=False)

    def matches(self, in_sizes, out_sizes, chunk, slots) -> bool:
        return (self.in_sizes, self.out_sizes, self.chunk, self.slots) == (in_sizes, out_sizes, chunk, slots)

    def run(self, host_in, host_out, E, solve, stream):
        """host_in / host_out: lists of (E, n) host tensors in the order of
        in_sizes / out_sizes; solve(c0, c1, dev_in, dev_out) enqueues the
        solve of energies [c0, c1) on ``stream`` with device buffers holding
        them in their first c1 - c0 rows."""
        torch = _torch()
        C, S = self.chunk, self.slots
        nch = (E + C - 1) // C
        rng = [(c * C, min(E, (c + 1) * C)) for c in range(nch)]

        def stage_in(c):
            s, (c0, c1) = c % S, rng[c]
            if c >= S:
                self.ev_in[s].synchronize()  # pinned inputs s no longer read
            for pin, src in zip(self.pin_in[s], host_in):
                pin[: c1 - c0].copy_(src[c0:c1])
            with torch.cuda.stream(self.copy_in):
                if c >= S:
                    self.copy_in.wait_event(self.ev_solve[s])  # device inputs s no longer read
                for dev, pin in zip(self.dev_in[s], self.pin_in[s]):
                    dev[: c1 - c0].copy_(pin[: c1 - c0], non_blocking=True)
                self.ev_in[s].record(self.copy_in)

        def solve_chunk(c):
            s, (c0, c1) = c % S, rng[c]
            stream.wait_event(self.ev_in[s])
            if c >= S:
                stream.wait_event(self.ev_out[s])  # device outputs s copied out
            solve(c0, c1, [t[: c1 - c0] for t in self.dev_in[s]], [t[: c1 - c0] for t in self.dev_out[s]])
            self.ev_solve[s].record(stream)
            with torch.cuda.stream(self.copy_out):
                self.copy_out.wait_event(self.ev_solve[s])
                for pin, dev in zip(self.pin_out[s], self.dev_out[s]):
                    pin[: c1 - c0].copy_(dev[: c1 - c0], non_blocking=True)
                self.ev_out[s].record(self.copy_out)

        def stage_out(c):
            s, (c0, c1) = c % S, rng[c]
            self.ev_out[s].synchronize()
            for dst, pin in zip(host_out, self.pin_out[s]):
                dst[c0:c1].copy_(pin[: c1 - c0])

        # the copies out of pinned memory run on the worker thread beside the
        # copies into it (torch releases the GIL in copy_); pinned outputs s
        # are free for chunk c's transfer once chunk c - S's copy is done
        pending = [None] * S
        for c in range(min(S - 1, nch)):
            stage_in(c)
        for c in range(nch):
            if pending[c % S] is not None:
                pending[c % S].result()
                pending[c % S] = None
            solve_chunk(c)
            if c + S - 1 < nch:
                stage_in(c + S - 1)
            if c >= 1:
                pending[(c - 1) % S] = self.pool.submit(stage_out, c - 1)
        stage_out(nch - 1)
        for f in pending:
            if f is not None:
                f.result()


class RGF(GFSolver):
    """Selected inversion / selected solve by the recursive Green's function
    method on MI355X.

    Parameters
    ----------
    max_batch_size : int, optional
        Energy batch cap, by default 100; batches follow ``get_batches``
        (they may exceed the cap, e.g. 10 energies with cap 3 -> [4, 3, 3]).
    device : optional
        Torch device for the solves (default: the current device).
    """

    def __init__(self, max_batch_size: int = 100, device=None) -> None:
        self.max_batch_size = max_batch_size
        self.device = device

    def _dev(self):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("RGF needs a ROCm device (no CPU fallback)")
        if self.device is not None:
            return torch.device(self.device)
        return torch.device("cuda", torch.cuda.current_device())

    # ------------------------------------------------------------------
    def selected_inv(self, a, obc_blocks: OBCBlocks | None = None, out=None):
        device = self._dev()
        ctx = _solve.default_context(device)
        geo = _geometry(a)
        if obc_blocks is None:
            obc_blocks = OBCBlocks(num_blocks=a.num_blocks)
        _, offs = get_batches(a.shape[0], self.max_batch_size)
        x = out if out is not None else a.__class__.zeros_like(a)
        host = _host_operand(a) and _host_operand(x)
        for bidx in range(len(offs) - 1):
            sl = slice(int(offs[bidx]), int(offs[bidx + 1]))
            if host and self._inv_host_batch(geo, ctx, device, sl, a, x, obc_blocks):
                continue
            if _is_dsdb(a):
                A, lead = _gather_dsdb(a, sl, geo, "dul", device, ctx)
            else:
                a_ = a.stack[sl]
                A = _gather(a_, geo, "dul", device)
                lead = tuple(a_.shape[:-2])
            E = int(A[0].shape[0])
            x_ = None if _is_dsdb(x) else x.stack[sl]
            direct = x_ is not None and _is_device_container(x_) and not x_.symmetry
            X, inplace = _outputs(x_, geo, E, device, "dul") if direct else _outputs(None, geo, E, device, "dul")
            obr = _obc_batch(obc_blocks.retarded, sl, geo, E, device)
            _solve.selected_inv(geo, A, x=X, obc_retarded=obr if any(t is not None for t in obr) else None,
                                ctx=ctx)
            ctx.status()
            if x_ is None:
                _scatter_dsdb(x, sl, geo, X, "du" if x.symmetry else "dul", device, ctx)
            elif not inplace:
                _scatter(x_, geo, X, lead, ("d", "u") if x_.symmetry else ("d", "u", "l"))
        if out is None:
            return x
        return None

    # ------------------------------------------------------------------
    def _pipe_for(self, device, in_sizes, out_sizes, E, min_chunk=1):
        """The staging of this solver for a batch of E host-resident energies
        (None when the batch is too small to split)."""
        chunk = max(1, min((E + 1) // 2, max(min_chunk, _HOST_CHUNK_BYTES // (16 * sum(in_sizes)))))
        if E < 2:
            return None
        pipe = getattr(self, "_pipe", None)
        if pipe is None or not pipe.matches(in_sizes, out_sizes, chunk, _HOST_SLOTS):
            self._pipe = None  # release the old staging first
            pipe = self._pipe = _HostPipe(device, in_sizes, out_sizes, chunk, _HOST_SLOTS)
        return pipe

    def _host_batch(self, geo, ctx, device, sl, ins, outs, solve) -> bool:
        """One energy batch of host-resident operands through ``_HostPipe``;
        False (nothing done) when the batch is too small to split.

        ins: ``(stack, parts)`` pairs; outs: ``(stack, parts, written)``
        triples (``written``: the parts stored back, a subset of ``parts``).
        A host BlockTridiagonal moves its packed parts; a host block-sparse
        stack moves its stored entries, densified into packed
        buffers (inputs) or sparsified from them (outputs) on the device per
        chunk.  A block-sparse output whose written parts do not hold every
        stored entry is staged in as well, so that the entries outside them
        keep their values.  ``solve(c0, c1, packed_in, packed_out)`` enqueues
        the solve of energies [c0, c1) on the context stream, given
        ``(d, u, l)`` device tuples (None for parts not listed) of ins and
        outs.
        """
        torch = _torch()
        sizes = {"d": geo.diag_size, "u": geo.offdiag_size, "l": geo.offdiag_size}
        host_in, host_out, spec_in, spec_out, keep = [], [], [], [], []

        def host_parts(m, parts):
            if _is_dsdb(m):
                d = m.data[sl]
                return [d if isinstance(d, torch.Tensor) else torch.from_numpy(d)], True
            p = m.stack[sl].packed()
            return [p["dul".index(k)] for k in parts], False

        for m, parts in ins:
            ts, sparse = host_parts(m, parts)
            spec_in.append((m, parts, parts, sparse, len(host_in)))
            host_in += ts
        for m, parts, written in outs:
            ts, sparse = host_parts(m, parts)
            spec_out.append((m, parts, written, sparse, len(host_out)))
            if sparse and not _dsdb_writes_all(m, geo, written, int(ts[0].shape[1]), device):
                keep.append((len(host_in), len(host_out)))
                host_in.append(ts[0])
            host_out += ts
        E = int(host_in[0].shape[0])
        any_sparse = any(sp[3] for sp in spec_in + spec_out)
        pipe = self._pipe_for(device, tuple(int(t.shape[1]) for t in host_in),
                              tuple(int(t.shape[1]) for t in host_out), E,
                              _HOST_MIN_CHUNK if any_sparse else 1)
        if pipe is None:
            return False

        def packed(spec, bufs, n, inputs):
            res, post = [], []
            for m, parts, written, sparse, k in spec:
                if not sparse:
                    res.append(tuple(bufs[k + parts.index(c)] if c in parts else None for c in "dul"))
                elif inputs:
                    res.append(_gather_dsdb(m, None, geo, parts, device, ctx, data=bufs[k])[0])
                else:
                    P = tuple(torch.empty((n, sizes[c]), dtype=torch.complex128, device=device)
                              if c in parts else None for c in "dul")
                    res.append(P)
                    post.append((m, P, written, bufs[k]))
            return res, post

        def chunk(c0, c1, din, dout):
            for i, j in keep:  # the stored entries outside the written blocks
                dout[j].copy_(din[i])
            pin, _ = packed(spec_in, din, c1 - c0, True)
            pout, post = packed(spec_out, dout, c1 - c0, False)
            solve(c0, c1, pin, pout)
            for m, P, written, d in post:
                _scatter_dsdb(m, None, geo, P, written, device, ctx, data=d)

        pipe.run(host_in, host_out, E, chunk, ctx.stream)
        ctx.status()
        return True

    def _inv_host_batch(self, geo, ctx, device, sl, a, x, obc_blocks) -> bool:
        has_obc = any(t is not None for t in obc_blocks.retarded)

        def solve(c0, c1, pin, pout):
            obr = None
            if has_obc:
                obr = _obc_batch(obc_blocks.retarded, slice(sl.start + c0, sl.start + c1), geo, c1 - c0,
                                 device)
            _solve.selected_inv(geo, pin[0], x=pout[0], obc_retarded=obr, ctx=ctx)

        return self._host_batch(geo, ctx, device, sl, [(a, "dul")],
                                [(x, "dul", "du" if x.symmetry else "dul")], solve)

    def _solve_host_batch(self, geo, ctx, device, sl, a, sigma_lesser, sigma_greater, xl, xg, xr,
                          obc_blocks, current) -> bool:
        has_obc = any(t is not None for t in obc_blocks.retarded + obc_blocks.lesser + obc_blocks.greater)
        torch = _torch()
        E = sl.stop - sl.start
        cur = (torch.empty((E, geo.n_blocks + 1), dtype=torch.complex128, device=device)
               if current is not None else None)

        def solve(c0, c1, pin, pout):
            n = c1 - c0
            obc = None
            if has_obc:
                csl = slice(sl.start + c0, sl.start + c1)
                obc = (_obc_batch(obc_blocks.retarded, csl, geo, n, device),
                       _obc_batch(obc_blocks.lesser, csl, geo, n, device),
                       _obc_batch(obc_blocks.greater, csl, geo, n, device))
            _solve.selected_solve(geo, pin[0], pin[1][:2], pin[2][:2], g_lesser=pout[0],
                                  g_greater=pout[1], g_retarded=pout[2] if xr is not None else None,
                                  obc=obc, current=cur[c0:c1] if cur is not None else None, ctx=ctx)

        part = lambda m, sym: (sym, sym) if m.symmetry else ("dul", "dul")  # noqa: E731
        outs = [(xl, *part(xl, "du")), (xg, *part(xg, "du"))]
        if xr is not None:
            outs.append((xr, *part(xr, "d")))
        if not self._host_batch(geo, ctx, device, sl, [(a, "dul"), (sigma_lesser, "du"),
                                                       (sigma_greater, "du")], outs, solve):
            return False
        if cur is not None:
            current[sl] = cur if not isinstance(current, np.ndarray) else cur.cpu().numpy()
        return True

    def selected_solve(self, a, sigma_lesser, sigma_greater, obc_blocks: OBCBlocks | None = None,
                       out=None, return_retarded: bool = False, return_current: bool = False):
        device = self._dev()
        ctx = _solve.default_context(device)
        nb = sigma_lesser.num_blocks
        geo = _geometry(sigma_lesser)
        if obc_blocks is None:
            obc_blocks = OBCBlocks(num_blocks=nb)
        stack_shape = tuple(sigma_lesser.shape[:-2])
        if return_current:
            if (obc_blocks.lesser[0] is None or obc_blocks.greater[0] is None
                    or obc_blocks.lesser[-1] is None or obc_blocks.greater[-1] is None):
                raise TypeError("return_current needs lesser and greater OBC blocks at both ends")
            current = _zeros_current(sigma_lesser, (*stack_shape, nb + 1))
        _, offs = get_batches(stack_shape[0], self.max_batch_size)
        if out is not None:
            xl, xg, *xr = out
            if return_retarded:
                if len(xr) != 1:
                    raise ValueError("Invalid number of output matrices.")
                xr = xr[0]
        else:
            xl = a.__class__.zeros_like(a)
            xg = a.__class__.zeros_like(a)
            if return_retarded:
                xr = a.__class__.zeros_like(a)

        host = (_host_operand(a) and _host_operand(sigma_lesser) and _host_operand(sigma_greater)
                and _host_operand(xl) and _host_operand(xg) and (not return_retarded or _host_operand(xr)))
        for bidx in range(len(offs) - 1):
            sl = slice(int(offs[bidx]), int(offs[bidx + 1]))
            if host and self._solve_host_batch(geo, ctx, device, sl, a, sigma_lesser, sigma_greater,
                                               xl, xg, xr if return_retarded else None, obc_blocks,
                                               current if return_current else None):
                continue

            def inp(m, parts):
                if _is_dsdb(m):
                    return _gather_dsdb(m, sl, geo, parts, device, ctx)
                m_ = m.stack[sl]
                return _gather(m_, geo, parts, device), tuple(m_.shape[:-2])

            A, lead = inp(a, "dul")
            SL = inp(sigma_lesser, "du")[0][:2]
            SG = inp(sigma_greater, "du")[0][:2]
            E = int(A[0].shape[0])
            view = lambda m: None if _is_dsdb(m) else m.stack[sl]  # noqa: E731
            xl_, xg_ = view(xl), view(xg)
            xr_ = view(xr) if return_retarded else None

            gl_parts = "du" if xl.symmetry else "dul"
            gg_parts = "du" if xg.symmetry else "dul"
            GL, gl_inplace = _outputs(xl_ if _is_device_container(xl_) else None, geo, E, device, gl_parts)
            GG, gg_inplace = _outputs(xg_ if _is_device_container(xg_) else None, geo, E, device, gg_parts)
            GR, gr_inplace = None, True
            if return_retarded:
                gr_parts = "d" if xr.symmetry else "dul"
                GR, gr_inplace = _outputs(xr_ if _is_device_container(xr_) else None, geo, E, device,
                                          gr_parts)
            obc = None
            if any(t is not None for t in obc_blocks.retarded + obc_blocks.lesser + obc_blocks.greater):
                obc = (_obc_batch(obc_blocks.retarded, sl, geo, E, device),
                       _obc_batch(obc_blocks.lesser, sl, geo, E, device),
                       _obc_batch(obc_blocks.greater, sl, geo, E, device))
            cur = None
            if return_current:
                torch = _torch()
                cur = torch.empty((E, nb + 1), dtype=torch.complex128, device=device)
            _solve.selected_solve(geo, A, SL, SG, g_lesser=GL, g_greater=GG, g_retarded=GR, obc=obc,
                                  current=cur, ctx=ctx)
            ctx.status()
            for m, m_, bufs, parts, inplace in ((xl, xl_, GL, gl_parts, gl_inplace),
                                                 (xg, xg_, GG, gg_parts, gg_inplace)) + (
                    ((xr, xr_, GR, gr_parts, gr_inplace),) if return_retarded else ()):
                if m_ is None:
                    _scatter_dsdb(m, sl, geo, bufs, parts, device, ctx)
                elif not inplace:
                    _scatter(m_, geo, bufs, lead, tuple(parts))
            if return_current:
                c = cur.reshape(*lead, nb + 1)
                current[sl] = c if not isinstance(current, np.ndarray) else c.cpu().numpy()

        if out is None:
            res = (xl, xg) + ((xr,) if return_retarded else ())
            return res + ((current,) if return_current else ())
        if return_current:
            return current
        return None
