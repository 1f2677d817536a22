"""Block-sparse stacks in the reference's DSDBCOO / DSDBCSR layout, on the device.

A single-rank counterpart of ``DSDBCOO`` (``src/qttools/datastructures/
dsdbcoo.py``) and ``DSDBCSR`` (``dsdbcsr.py``): ``data[*stack, nnz]`` with
block-sorted global indices (COO: ``rows``, ``cols``; CSR: ``rowptr_map``,
``cols``), ``from_sparray`` (``dsdbcoo.py:885-977``, ``dsdbcsr.py:800-863``),
``.stack[...]`` views over the first stack dimension, ``.blocks[i, j]``
get/set (``dsdbcoo.py:333-521``, ``dsdbcsr.py:285-472``) with the symmetric
storage rules, ``zeros_like`` and ``to_dense``.  ``data`` is a torch tensor;
block access runs the HIP densify / sparsify kernels (``rgf_densify_z`` /
``rgf_sparsify_z``), so these containers need a ROCm device, like the solver.

The RGF drop-in recognises these containers - and the reference's own, by
their ``data``/``rows``/``cols`` or ``rowptr_map`` attributes - and moves the
tridiagonal blocks of a whole energy batch with one densify per input and one
sparsify per output.
"""

from __future__ import annotations

import ctypes

import numpy as np

from quatrex_amd import _lib
from quatrex_amd import sparse_blocks as SB
from quatrex_amd.layout import Geometry


def _torch():
    import torch

    return torch


def _block_sort(rows: np.ndarray, cols: np.ndarray, block_sizes) -> np.ndarray:
    """Order that sorts entries by block row, block column, then row, column."""
    off = np.concatenate(([0], np.cumsum(np.asarray(block_sizes, np.int64))))
    br = np.searchsorted(off, rows, side="right") - 1
    bc = np.searchsorted(off, cols, side="right") - 1
    return np.lexsort((cols, rows, bc, br))


def _canonical(arr, symmetry: bool):
    from scipy import sparse

    coo = sparse.coo_matrix(arr).copy()
    coo.sum_duplicates()
    if symmetry:
        coo = sparse.triu(coo, format="coo")
        coo.sum_duplicates()
    return coo


class _DSDB:
    """Shared part of the two layouts: data, stack views, block access."""

    def __init__(self, data, block_sizes, global_stack_shape, symmetry=False, symmetry_op=np.conj):
        torch = _torch()
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(np.ascontiguousarray(data))
        self.data = data
        self.block_sizes = np.asarray(block_sizes, dtype=np.int64)
        self.block_offsets = np.concatenate(([0], np.cumsum(self.block_sizes)))
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        self.global_stack_shape = tuple(int(s) for s in global_stack_shape)
        self.symmetry = bool(symmetry)
        self.symmetry_op = symmetry_op
        self.geometry = Geometry(tuple(int(b) for b in self.block_sizes))
        self._maps = {}

    # --- reference-like properties -----------------------------------------
    @property
    def num_blocks(self) -> int:
        return int(self.block_sizes.size)

    @property
    def nnz(self) -> int:
        return int(self.data.shape[-1])

    @property
    def shape(self) -> tuple:
        n = int(self.block_offsets[-1])
        return (*self.data.shape[:-1], n, n)

    @property
    def stack(self) -> "_Stack":
        return _Stack(self)

    @property
    def blocks(self) -> "_Blocks":
        return _Blocks(self)

    @property
    def dtype(self):
        return self.data.dtype

    def _structure_view(self, data):
        new = object.__new__(self.__class__)
        new.__dict__.update(self.__dict__)
        new.data = data
        return new

    @classmethod
    def zeros_like(cls, other: "_DSDB") -> "_DSDB":
        return other._structure_view(_torch().zeros_like(other.data))

    # --- device maps ---------------------------------------------------------
    def tridiagonal_map(self, device) -> SB.DeviceMap:
        key = ("tri", str(device))
        if key not in self._maps:
            self._maps[key] = SB.DeviceMap(SB.map_of(self, self.geometry), device,
                                           _part_sizes(self.geometry))
        return self._maps[key]

    def _coordinates(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(rows, cols, positions) of every stored entry."""
        raise NotImplementedError

    def _block_map(self, i: int, j: int, device):
        """Map and geometry for dense access to block (i, j) alone."""
        key = ("blk", i, j, str(device))
        if key not in self._maps:
            rows, cols, pos = self._coordinates()
            off = self.block_offsets
            bs = self.block_sizes
            si, sj = (j, i) if (self.symmetry and j < i) else (i, j)
            m = ((rows >= off[si]) & (rows < off[si + 1]) & (cols >= off[sj]) & (cols < off[sj + 1]))
            r, c, p = rows[m] - off[si], cols[m] - off[sj], pos[m]
            if si == sj:
                geo = Geometry((int(bs[si]),))
                part = "d"
                hm = SB.map_from_coordinates(geo, r, c, p, self.symmetry, self.symmetry_op)
            else:
                # stored block (si, sj) as the upper block of a two-block geometry;
                # a mirrored request (symmetric, j < i) is its lower block
                geo = Geometry((int(bs[si]), int(bs[sj])))
                part = "l" if (si, sj) != (i, j) else "u"
                hm = SB.map_from_coordinates(geo, r, c + int(bs[si]), p, self.symmetry, self.symmetry_op)
            self._maps[key] = (SB.DeviceMap(hm, device, _part_sizes(geo)), geo, part)
        return self._maps[key]

    def _ctx(self):
        from quatrex_amd import solve as S

        return S.default_context(self.data.device)

    def get_block(self, data, i: int, j: int):
        """Dense block (i, j) of the stack ``data`` (``(*lead, nnz)``)."""
        torch = _torch()
        dev = data.device
        if dev.type != "cuda":
            raise RuntimeError("block access runs on the ROCm device: move data there first")
        dmap, geo, part = self._block_map(i, j, dev)
        lead = tuple(data.shape[:-1])
        flat = data.reshape(-1, self.nnz).contiguous()
        E = int(flat.shape[0])
        b_i, b_j = int(self.block_sizes[i]), int(self.block_sizes[j])
        out = torch.empty((E, b_i * b_j), dtype=torch.complex128, device=dev)
        bt = _lib.Bt_t(out.data_ptr() if part == "d" else None, out.data_ptr() if part == "u" else None,
                       out.data_ptr() if part == "l" else None)
        g, keep = _lib.geometry_struct(geo.block_sizes)
        ctx = self._ctx()
        rc = _lib.load().rgf_densify_z(ctx.handle, ctypes.byref(g), E, ctypes.c_void_p(flat.data_ptr()),
                                       self.nnz, ctypes.byref(dmap.restricted(part)), ctypes.byref(bt))
        _lib.check(rc, ctx.handle)
        del keep
        return out.reshape(*lead, b_i, b_j)

    def set_block(self, data, i: int, j: int, value) -> None:
        """Write block (i, j) into the stack ``data`` (stored entries only)."""
        torch = _torch()
        dev = data.device
        if dev.type != "cuda":
            raise RuntimeError("block access runs on the ROCm device: move data there first")
        if self.symmetry and j < i:
            # the stored block is (j, i) = op(value^T), as _set_block does
            flags = SB.symmetry_flags(self.symmetry_op)
            v = torch.as_tensor(value, device=dev).to(torch.complex128).transpose(-1, -2)
            if not flags & SB.OP_NOCONJ:
                v = v.conj()
            if flags & SB.OP_NEGATE:
                v = -v
            return self.set_block(data, j, i, v.contiguous())
        dmap, geo, part = self._block_map(i, j, dev)
        lead = tuple(data.shape[:-1])
        E = int(np.prod(lead)) if lead else 1
        b_i, b_j = int(self.block_sizes[i]), int(self.block_sizes[j])
        val = torch.as_tensor(value, device=dev).to(torch.complex128)
        val = val.expand(*lead, b_i, b_j).reshape(E, b_i * b_j).contiguous()
        if not data.is_contiguous():
            raise ValueError("block writes need a contiguous data view")
        bt = _lib.BtConst_t(val.data_ptr() if part == "d" else None, val.data_ptr() if part == "u" else None,
                            val.data_ptr() if part == "l" else None)
        g, keep = _lib.geometry_struct(geo.block_sizes)
        ctx = self._ctx()
        rc = _lib.load().rgf_sparsify_z(ctx.handle, ctypes.byref(g), E, ctypes.byref(bt),
                                        ctypes.byref(dmap.restricted(part)), ctypes.c_void_p(data.data_ptr()),
                                        self.nnz)
        _lib.check(rc, ctx.handle)
        del keep

    def to_dense(self):
        """Dense ``(*stack, n, n)`` copy (tests and debugging)."""
        torch = _torch()
        n = int(self.block_offsets[-1])
        out = torch.zeros((*self.data.shape[:-1], n, n), dtype=torch.complex128, device=self.data.device)
        rows, cols, pos = self._coordinates()
        if rows.size:
            vals = self.data[..., torch.as_tensor(pos, device=self.data.device)]
            out[..., rows, cols] = vals
            if self.symmetry:
                flags = SB.symmetry_flags(self.symmetry_op)
                mirror = vals if flags & SB.OP_NOCONJ else vals.conj()
                mirror = -mirror if flags & SB.OP_NEGATE else mirror
                off = rows != cols
                out[..., cols[off], rows[off]] = mirror[..., torch.as_tensor(np.nonzero(off)[0], device=vals.device)]
                on = rows == cols
                if on.any():
                    d = vals[..., torch.as_tensor(np.nonzero(on)[0], device=vals.device)]
                    m_ = mirror[..., torch.as_tensor(np.nonzero(on)[0], device=vals.device)]
                    out[..., rows[on], cols[on]] = 0.5 * (d + m_)
        return out


class DSDBCOO(_DSDB):
    """``data[*stack, nnz]`` with block-sorted global ``rows`` and ``cols``."""

    def __init__(self, data, rows, cols, block_sizes, global_stack_shape, symmetry=False,
                 symmetry_op=np.conj):
        super().__init__(data, block_sizes, global_stack_shape, symmetry, symmetry_op)
        self.rows = np.asarray(rows, dtype=np.int64)
        self.cols = np.asarray(cols, dtype=np.int64)

    def _coordinates(self):
        return self.rows, self.cols, np.arange(self.rows.size, dtype=np.int64)

    @classmethod
    def from_sparray(cls, arr, block_sizes, global_stack_shape, symmetry=False, symmetry_op=np.conj,
                     device="cuda"):
        torch = _torch()
        coo = _canonical(arr, symmetry)
        order = _block_sort(coo.row.astype(np.int64), coo.col.astype(np.int64), block_sizes)
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        vals = torch.as_tensor(coo.data[order].astype(np.complex128), device=device)
        data = vals.expand(*global_stack_shape, vals.numel()).contiguous()
        return cls(data, coo.row[order], coo.col[order], block_sizes, global_stack_shape, symmetry,
                   symmetry_op)


class DSDBCSR(_DSDB):
    """``data[*stack, nnz]`` with a ``rowptr`` per stored block and global ``cols``."""

    def __init__(self, data, cols, rowptr_map, block_sizes, global_stack_shape, symmetry=False,
                 symmetry_op=np.conj):
        super().__init__(data, block_sizes, global_stack_shape, symmetry, symmetry_op)
        self.cols = np.asarray(cols, dtype=np.int64)
        self.rowptr_map = {k: np.asarray(v, dtype=np.int64) for k, v in rowptr_map.items()}

    def _coordinates(self):
        rows_l, pos_l = [], []
        for (bi, _bj), rp in self.rowptr_map.items():
            rows_l.append(self.block_offsets[bi] + np.repeat(np.arange(rp.size - 1), np.diff(rp)))
            pos_l.append(np.arange(rp[0], rp[-1], dtype=np.int64))
        if not pos_l:
            z = np.zeros(0, np.int64)
            return z, z, z
        pos = np.concatenate(pos_l)
        return np.concatenate(rows_l).astype(np.int64), self.cols[pos], pos

    @classmethod
    def from_sparray(cls, arr, block_sizes, global_stack_shape, symmetry=False, symmetry_op=np.conj,
                     device="cuda"):
        torch = _torch()
        coo = _canonical(arr, symmetry)
        rows = coo.row.astype(np.int64)
        cols = coo.col.astype(np.int64)
        order = _block_sort(rows, cols, block_sizes)
        rows, cols = rows[order], cols[order]
        off = np.concatenate(([0], np.cumsum(np.asarray(block_sizes, np.int64))))
        br = np.searchsorted(off, rows, side="right") - 1
        bc = np.searchsorted(off, cols, side="right") - 1
        rowptr_map = {}
        start = 0
        while start < rows.size:
            i, j = int(br[start]), int(bc[start])
            stop = start
            while stop < rows.size and br[stop] == i and bc[stop] == j:
                stop += 1
            local = rows[start:stop] - off[i]
            counts = np.bincount(local, minlength=int(off[i + 1] - off[i]))
            rowptr_map[(i, j)] = start + np.concatenate(([0], np.cumsum(counts)))
            start = stop
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        vals = torch.as_tensor(coo.data[order].astype(np.complex128), device=device)
        data = vals.expand(*global_stack_shape, vals.numel()).contiguous()
        return cls(data, cols, rowptr_map, block_sizes, global_stack_shape, symmetry, symmetry_op)


def _part_sizes(geo: Geometry) -> tuple:
    return (geo.diag_size, geo.offdiag_size, geo.offdiag_size)


def structure_of(m):
    """The object that holds the sparsity structure of ``m``: ``m`` itself, or
    the parent stack of a substack view (the reference's ``_DStackView`` and
    :class:`StackView` keep it in ``_dsdbsparse``)."""
    return getattr(m, "_dsdbsparse", m)


def device_map(m, geo: Geometry, device) -> SB.DeviceMap:
    """The tridiagonal gather map of ``m`` on ``device`` (cached on the object
    that holds the structure)."""
    m = structure_of(m)
    if hasattr(m, "tridiagonal_map"):
        return m.tridiagonal_map(device)
    cache = m.__dict__.setdefault("_qx_tridiagonal_maps", {})
    key = str(device)
    if key not in cache:
        cache[key] = SB.DeviceMap(SB.map_of(m, geo), device, _part_sizes(geo))
    return cache[key]


def writes_every_entry(m, geo: Geometry, parts: str, nnz: int, device) -> bool:
    """Whether writing the packed ``parts`` into ``m`` (rgf_sparsify_z)
    overwrites every one of its ``nnz`` stored entries, so that an output's
    previous values need not be read (cached per structure and parts)."""
    torch = _torch()
    dmap = device_map(m, geo, device)
    cache = dmap.__dict__.setdefault("_writes_all", {})
    key = (parts, nnz)
    if key not in cache:
        seen = torch.zeros(nnz, dtype=torch.bool, device=device)
        for k in parts:
            src, _, op = dmap.tensors[k]
            if src.numel():
                seen[src[(op & SB.OP_MIRROR) == 0]] = True
        cache[key] = bool(seen.all())
    return cache[key]


def _batch_data(m, sl, device):
    """The batch's data as a contiguous ``(E, nnz)`` complex128 device tensor."""
    torch = _torch()
    x = m.data[sl]
    lead = tuple(x.shape[:-1])
    nnz = int(x.shape[-1])
    if isinstance(x, torch.Tensor):
        t = x.to(device=device, dtype=torch.complex128)
    else:
        host = np.ascontiguousarray(SB._host(x), dtype=np.complex128)
        t = torch.from_numpy(host)
        t = (t.pin_memory() if torch.cuda.is_available() else t).to(device, non_blocking=True)
    return t.reshape(-1, nnz).contiguous(), lead


def densify_batch(m, sl, geo: Geometry, parts: str, device, ctx, data=None):
    """Packed ``(diag, upper, lower)`` of a batch of a block-sparse stack: one
    densify launch per part for all energies (rgf_densify_z).  ``data``: the
    batch's stored entries already on the device (``(E, nnz)``), else they are
    read from ``m.data[sl]``."""
    torch = _torch()
    if data is None:
        data, lead = _batch_data(m, sl, device)
    else:
        lead = (int(data.shape[0]),)
    E = int(data.shape[0])
    sizes = {"d": geo.diag_size, "u": geo.offdiag_size, "l": geo.offdiag_size}
    out = {k: torch.empty((E, sizes[k]), dtype=torch.complex128, device=device) for k in parts}
    dmap = device_map(m, geo, device)
    bt = _lib.Bt_t(*(out[k].data_ptr() if k in out else None for k in "dul"))
    g, keep = _lib.geometry_struct(geo.block_sizes)
    rc = _lib.load().rgf_densify_z(ctx.handle, ctypes.byref(g), E, ctypes.c_void_p(data.data_ptr()),
                                   int(data.shape[1]), ctypes.byref(dmap.restricted(parts)),
                                   ctypes.byref(bt))
    _lib.check(rc, ctx.handle)
    del keep
    return tuple(out.get(k) for k in "dul"), lead


def sparsify_batch(m, sl, geo: Geometry, bufs, parts: str, device, ctx, data=None) -> None:
    """Write packed results into the stored entries of a block-sparse stack
    batch (rgf_sparsify_z); entries outside the written blocks keep their
    values.  ``data``: a device ``(E, nnz)`` tensor holding the batch's
    stored entries, written in place instead of ``m.data[sl]``."""
    torch = _torch()
    if data is not None:
        view, direct = data, True
    else:
        view = m.data[sl]
        direct = (isinstance(view, torch.Tensor) and view.device == torch.device(device)
                  and view.dtype == torch.complex128 and view.is_contiguous())
    if direct:
        data = view.reshape(-1, view.shape[-1])
    else:
        data, _ = _batch_data(m, sl, device)
    E = int(data.shape[0])
    dmap = device_map(m, geo, device)
    use = "".join(k for k, t in zip("dul", bufs) if t is not None and k in parts)
    bt = _lib.BtConst_t(*(t.data_ptr() if (t is not None and k in use) else None for k, t in zip("dul", bufs)))
    g, keep = _lib.geometry_struct(geo.block_sizes)
    rc = _lib.load().rgf_sparsify_z(ctx.handle, ctypes.byref(g), E, ctypes.byref(bt),
                                    ctypes.byref(dmap.restricted(use)), ctypes.c_void_p(data.data_ptr()),
                                    int(data.shape[1]))
    _lib.check(rc, ctx.handle)
    del keep
    if not direct:
        res = data.reshape(view.shape)
        if isinstance(view, torch.Tensor):
            view.copy_(res.to(view.device, view.dtype))
        else:
            view[...] = res.cpu().numpy().astype(view.dtype, copy=False)



def _compose_first(index, sub, n: int):
    """Index on the first stack axis equal to applying ``index`` then ``sub``
    (both a slice with positive step or an int)."""
    r = range(n)[index]
    if isinstance(r, int):
        raise IndexError("a stack view that selects one entry has no stack axis to index")
    r = r[sub]
    if isinstance(r, int):
        return r
    return slice(r.start, r.start + len(r) * r.step, r.step)


class StackView:
    """A substack of a block-sparse stack, shaped like the reference's
    ``_DStackView`` (``dsdbsparse.py:1118-1205``): it holds the parent stack
    (``_dsdbsparse``) and an index on the first stack axis (``_stack_index``),
    exposes ``data`` as the parent's data at that index (writes reach the
    parent) and, like the reference's view, carries no ``rows`` / ``cols`` /
    ``rowptr_map`` of its own."""

    def __init__(self, parent: _DSDB, index):
        self._dsdbsparse = parent
        self._stack_index = (index,)
        self.symmetry = parent.symmetry

    @property
    def data(self):
        return self._dsdbsparse.data[self._stack_index]

    @data.setter
    def data(self, value):
        self._dsdbsparse.data[self._stack_index] = value

    @property
    def dtype(self):
        return self._dsdbsparse.dtype

    @property
    def block_sizes(self):
        return self._dsdbsparse.block_sizes

    @property
    def num_blocks(self) -> int:
        return self._dsdbsparse.num_blocks

    @property
    def nnz(self) -> int:
        return self._dsdbsparse.nnz

    @property
    def shape(self) -> tuple:
        n = int(self._dsdbsparse.block_offsets[-1])
        return (*self.data.shape[:-1], n, n)

    @property
    def stack(self) -> "_Stack":
        return _Stack(self)

    @property
    def blocks(self) -> "_Blocks":
        return _Blocks(self)

    def get_block(self, data, i: int, j: int):
        return self._dsdbsparse.get_block(data, i, j)

    def set_block(self, data, i: int, j: int, value) -> None:
        self._dsdbsparse.set_block(data, i, j, value)

    def to_dense(self):
        return self._dsdbsparse._structure_view(self.data).to_dense()


class _Stack:
    """``.stack[index]``: a :class:`StackView` over the first stack axis
    (views of views compose into one view of the parent)."""

    def __init__(self, m):
        self.m = m

    def __getitem__(self, sl):
        if isinstance(sl, tuple):
            if len(sl) != 1:
                raise IndexError("stack views index the first stack axis only")
            sl = sl[0]
        m = self.m
        if isinstance(m, StackView):
            n = int(m._dsdbsparse.data.shape[0])
            return StackView(m._dsdbsparse, _compose_first(m._stack_index[0], sl, n))
        return StackView(m, sl)


class _Blocks:
    def __init__(self, m):
        self.m = m

    def _ij(self, ij):
        i, j = ij
        n = self.m.num_blocks
        i, j = (i + n if i < 0 else i), (j + n if j < 0 else j)
        if not (0 <= i < n and 0 <= j < n):
            raise IndexError("block index out of range")
        return i, j

    def __getitem__(self, ij):
        i, j = self._ij(ij)
        return self.m.get_block(self.m.data, i, j)

    def __setitem__(self, ij, value):
        i, j = self._ij(ij)
        self.m.set_block(self.m.data, i, j, value)
