"""Dense-block tridiagonal stack with the ``DSDBSparse`` block interface.

Implements the part of ``DSDBSparse`` that the GF solvers use
(``src/qttools/datastructures/dsdbsparse.py``): ``.blocks[i, j]`` get/set
(``:1285-1363``), ``.stack[...]`` views over the first stack dimension
(``:1118-1283``), ``zeros_like`` (``dsdbcoo.py:849-883``), ``from_sparray``
(``dsdbcoo.py:885-977``), ``to_dense``, ``num_blocks``, ``block_sizes``,
``block_offsets``, ``shape``, ``symmetry``/``symmetry_op``.  Storage is the
packed layout of :mod:`quatrex_amd.layout` in torch tensors (host or device),
so device-resident stacks reach the HIP library without any copy.  Blocks
outside the tridiagonal pattern are not stored (writes to them are dropped,
as ``_set_block`` does for blocks absent from the sparsity pattern,
``dsdbcoo.py:511-514``).
"""

from __future__ import annotations

import numpy as np

from quatrex_amd.layout import Geometry


def _torch():
    import torch

    return torch


def _conj(x):
    return x.conj()


class BlockTridiagonal:
    def __init__(self, block_sizes, global_stack_shape, dtype=None, device="cpu", symmetry=False,
                 symmetry_op=None, _storage=None):
        torch = _torch()
        self.geometry = Geometry(tuple(int(b) for b in block_sizes))
        if isinstance(global_stack_shape, int):
            global_stack_shape = (global_stack_shape,)
        self.stack_shape = tuple(int(s) for s in global_stack_shape)
        self.dtype = dtype or torch.complex128
        self.symmetry = bool(symmetry)
        self.symmetry_op = symmetry_op or _conj
        g = self.geometry
        if _storage is None:
            dev = torch.device(device)
            _storage = (
                torch.zeros((*self.stack_shape, g.diag_size), dtype=self.dtype, device=dev),
                torch.zeros((*self.stack_shape, g.offdiag_size), dtype=self.dtype, device=dev),
                torch.zeros((*self.stack_shape, g.offdiag_size), dtype=self.dtype, device=dev),
            )
        self.diag, self.upper, self.lower = _storage

    # --- reference-like properties ------------------------------------------
    @property
    def num_blocks(self) -> int:
        return self.geometry.n_blocks

    @property
    def block_sizes(self) -> np.ndarray:
        return np.asarray(self.geometry.block_sizes, dtype=np.int64)

    @property
    def block_offsets(self) -> np.ndarray:
        return self.geometry.block_offsets

    @property
    def shape(self) -> tuple:
        n = self.geometry.n
        return (*self.stack_shape, n, n)

    @property
    def device(self):
        return self.diag.device

    @property
    def blocks(self) -> "_Blocks":
        return _Blocks(self)

    @property
    def stack(self) -> "_Stack":
        return _Stack(self)

    # --- constructors --------------------------------------------------------
    @classmethod
    def zeros_like(cls, other: "BlockTridiagonal") -> "BlockTridiagonal":
        return cls(other.geometry.block_sizes, other.stack_shape, other.dtype, other.device,
                   other.symmetry, other.symmetry_op)

    @classmethod
    def from_sparray(cls, arr, block_sizes, global_stack_shape, symmetry=False, device="cpu",
                     dtype=None):
        """Build from one (sparse or dense) matrix, broadcast over the stack."""
        torch = _torch()
        dense = arr.toarray() if hasattr(arr, "toarray") else np.asarray(arr)
        obj = cls(block_sizes, global_stack_shape, dtype, device, symmetry)
        d, u, l = obj.geometry.from_dense(dense.astype(np.complex128))
        obj.diag.copy_(torch.as_tensor(d).expand_as(obj.diag))
        obj.upper.copy_(torch.as_tensor(u).expand_as(obj.upper))
        obj.lower.copy_(torch.as_tensor(l).expand_as(obj.lower))
        return obj

    @classmethod
    def from_packed(cls, block_sizes, diag, upper, lower, symmetry=False):
        stack = tuple(diag.shape[:-1])
        return cls(block_sizes, stack, diag.dtype, diag.device, symmetry,
                   _storage=(diag, upper, lower))

    def to_dense(self):
        """Dense ``(*stack, n, n)`` tensor on the container's device."""
        torch = _torch()
        g = self.geometry
        bo = g.block_offsets
        out = torch.zeros(self.shape, dtype=self.dtype, device=self.device)
        for i in range(g.n_blocks):
            for j in range(max(0, i - 1), min(g.n_blocks, i + 2)):
                out[..., bo[i]:bo[i + 1], bo[j]:bo[j + 1]] = self.blocks[i, j]
        return out

    def packed(self, flatten: bool = True):
        """``(diag, upper, lower)`` storage, stack dims flattened to one axis."""
        if not flatten:
            return self.diag, self.upper, self.lower
        return tuple(t.reshape(-1, t.shape[-1]) for t in (self.diag, self.upper, self.lower))


class _Stack:
    def __init__(self, m: BlockTridiagonal):
        self._m = m

    def __getitem__(self, sl):
        m = self._m
        if not isinstance(sl, slice):
            sl = slice(sl, sl + 1)
        return BlockTridiagonal(m.geometry.block_sizes, (len(range(*sl.indices(m.stack_shape[0]))),
                                                         *m.stack_shape[1:]),
                                m.dtype, m.device, m.symmetry, m.symmetry_op,
                                _storage=(m.diag[sl], m.upper[sl], m.lower[sl]))


class _Blocks:
    def __init__(self, m: BlockTridiagonal):
        self._m = m

    def _view(self, i, j):
        m = self._m
        g = m.geometry
        if i == j:
            return g.diag(m.diag, i)
        if j == i + 1:
            return g.upper(m.upper, i)
        if i == j + 1:
            return g.lower(m.lower, j)
        return None

    def __getitem__(self, ij):
        torch = _torch()
        i, j = (int(v) for v in ij)
        m = self._m
        if m.symmetry and j < i:
            return m.symmetry_op(self[j, i].transpose(-1, -2)).contiguous()
        v = self._view(i, j)
        if v is None:
            b = m.geometry.block_sizes
            return torch.zeros((*m.stack_shape, b[i], b[j]), dtype=m.dtype, device=m.device)
        return v.clone()

    def __setitem__(self, ij, value):
        torch = _torch()
        i, j = (int(v) for v in ij)
        m = self._m
        value = torch.as_tensor(value, device=m.device)
        if m.symmetry and j < i:
            self[j, i] = m.symmetry_op(value.transpose(-1, -2))
            return
        v = self._view(i, j)
        if v is not None:
            v.copy_(value.to(m.dtype).expand_as(v))
