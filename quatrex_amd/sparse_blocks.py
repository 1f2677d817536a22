"""Gather maps between block-sparse stacks and the packed tridiagonal layout.

The reference stores each energy's matrix as ``data[*stack, nnz]`` with
block-sorted indices: DSDBCOO keeps global ``rows``/``cols`` per stored entry
(``src/qttools/datastructures/dsdbcoo.py:23-120``, sort order from
``kernels/datastructure/numba/dsdbcoo.py:175-235``), DSDBCSR a ``rowptr`` per
block and global ``cols`` (``dsdbcsr.py:285-472``).  RGF reads and writes
only the tridiagonal blocks ``(i,i)``, ``(i,i+1)``, ``(i+1,i)``
(``greens_function_solver/rgf.py``).

A map lists, for each packed part (diag, upper, lower), the stored entries
that land in it: ``src`` (position in ``data``), ``dst`` (offset in the part,
packed as in ``quatrex_amd.layout``) and ``op`` flags.  With symmetric
storage (``symmetry=True``: only the upper triangle is stored) the map also
carries the mirror images that ``_get_block`` produces
(``dsdbcoo.py:366-368, 405-408``): a lower block is ``op(upper^T)`` and a
diagonal block is ``X + op(X^T)`` with its diagonal halved.  The HIP library
applies a map to all energies of a batch in one launch per part
(``rgf_densify_z`` / ``rgf_sparsify_z``).
"""

from __future__ import annotations

import numpy as np

from quatrex_amd.layout import Geometry

OP_MIRROR = 1  # image of a stored entry (densify only): op(v)
OP_DIAGSYM = 2  # stored diagonal element of a symmetric diagonal block: (v + op(v)) / 2
OP_NEGATE = 4  # op(v) = -...
OP_NOCONJ = 8  # op(v) = +-v instead of +-conj(v)

PARTS = ("d", "u", "l")


def symmetry_flags(symmetry_op) -> int:
    """Flags of ``symmetry_op`` (one of conj, -conj, identity, negation)."""
    probe = np.asarray(symmetry_op(np.array([1.0 + 2.0j])), dtype=np.complex128)[0]
    table = {complex(1, -2): 0, complex(-1, 2): OP_NEGATE, complex(1, 2): OP_NOCONJ,
             complex(-1, -2): OP_NEGATE | OP_NOCONJ}
    for value, flags in table.items():
        if abs(probe - value) < 1e-12:
            return flags
    raise ValueError("symmetry_op must be conj, -conj, identity or negation")


def _empty():
    return (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.uint8))


def map_from_coordinates(geo: Geometry, rows, cols, positions=None, symmetry: bool = False,
                         symmetry_op=np.conj) -> dict:
    """Map of the stored entries ``(rows[k], cols[k])`` at ``data[..., positions[k]]``.

    ``rows``/``cols`` are global indices; ``positions`` defaults to
    ``arange(len(rows))`` (COO: entry k is data position k).
    """
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    pos = np.arange(rows.size, dtype=np.int64) if positions is None else np.asarray(positions, np.int64)
    bs = np.asarray(geo.block_sizes, dtype=np.int64)
    off = np.concatenate(([0], np.cumsum(bs)))
    n = int(off[-1])
    if rows.size and (rows.min() < 0 or cols.min() < 0 or rows.max() >= n or cols.max() >= n):
        raise ValueError("indices outside the matrix")
    doff = np.concatenate(([0], np.cumsum(bs * bs)))
    uoff = np.concatenate(([0], np.cumsum(bs[:-1] * bs[1:]))) if bs.size > 1 else np.zeros(1, np.int64)
    br = np.searchsorted(off, rows, side="right") - 1
    bc = np.searchsorted(off, cols, side="right") - 1
    lr = rows - off[br]
    lc = cols - off[bc]
    flags = symmetry_flags(symmetry_op) if symmetry else 0
    if symmetry and np.any(rows > cols):
        raise ValueError("symmetric storage holds the upper triangle only")

    out = {}
    # diagonal blocks
    m = br == bc
    i, r, c, p = br[m], lr[m], lc[m], pos[m]
    dst = doff[i] + r * bs[i] + c
    op = np.zeros(dst.size, np.uint8)
    src = p
    if symmetry:
        on_diag = r == c
        op[on_diag] = OP_DIAGSYM | flags
        off_d = ~on_diag
        src = np.concatenate((p, p[off_d]))
        dst = np.concatenate((dst, doff[i[off_d]] + c[off_d] * bs[i[off_d]] + r[off_d]))
        op = np.concatenate((op, np.full(int(off_d.sum()), OP_MIRROR | flags, np.uint8)))
    out["d"] = (src.astype(np.int64), dst.astype(np.int64), op)
    # upper blocks (i, i+1): b_i x b_{i+1}
    m = bc == br + 1
    i, r, c, p = br[m], lr[m], lc[m], pos[m]
    out["u"] = (p.astype(np.int64), (uoff[i] + r * bs[i + 1] + c).astype(np.int64),
                np.zeros(p.size, np.uint8))
    # lower blocks (i+1, i): b_{i+1} x b_i, stored or mirrored from the upper ones
    if symmetry:
        out["l"] = (p.astype(np.int64), (uoff[i] + c * bs[i] + r).astype(np.int64),
                    np.full(p.size, OP_MIRROR | flags, np.uint8))
    else:
        m = br == bc + 1
        i, r, c, p = bc[m], lr[m], lc[m], pos[m]
        out["l"] = (p.astype(np.int64), (uoff[i] + r * bs[i] + c).astype(np.int64),
                    np.zeros(p.size, np.uint8))
    return out


def map_from_csr(geo: Geometry, rowptr_map: dict, cols, symmetry: bool = False,
                 symmetry_op=np.conj) -> dict:
    """Map of a DSDBCSR structure: ``rowptr_map[(i, j)]`` holds the absolute
    positions of block (i, j)'s local rows, ``cols`` the global columns."""
    bs = np.asarray(geo.block_sizes, dtype=np.int64)
    off = np.concatenate(([0], np.cumsum(bs)))
    rows_l, pos_l = [], []
    for (bi, bj), rowptr in rowptr_map.items():
        if abs(int(bi) - int(bj)) > 1:
            continue
        rp = np.asarray(rowptr, dtype=np.int64)
        # a block's entries are contiguous: positions rowptr[0] .. rowptr[-1]
        pos = np.arange(rp[0], rp[-1], dtype=np.int64)
        rows_l.append(off[int(bi)] + np.repeat(np.arange(rp.size - 1, dtype=np.int64), np.diff(rp)))
        pos_l.append(pos)
    if not pos_l:
        return {k: _empty() for k in PARTS}
    pos = np.concatenate(pos_l)
    rows = np.concatenate(rows_l)
    return map_from_coordinates(geo, rows, np.asarray(cols, np.int64)[pos], positions=pos,
                                symmetry=symmetry, symmetry_op=symmetry_op)


def map_of(container, geo: Geometry) -> dict:
    """The map of a DSDBCOO-like (``rows``, ``cols``) or DSDBCSR-like
    (``rowptr_map``, ``cols``) container."""
    symmetry = bool(getattr(container, "symmetry", False))
    op = getattr(container, "symmetry_op", np.conj)
    if hasattr(container, "rows"):
        return map_from_coordinates(geo, _host(container.rows), _host(container.cols), None, symmetry, op)
    rowptr_map = getattr(container, "rowptr_map", None)
    if isinstance(rowptr_map, dict):
        return map_from_csr(geo, rowptr_map, _host(container.cols), symmetry, op)
    raise TypeError("not a block-sparse stack: needs rows/cols or a rowptr_map dict and cols")


def _host(x):
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    if hasattr(x, "get"):  # cupy-like
        return x.get()
    return np.asarray(x)


class DeviceMap:
    """A map resident on one device, with the ctypes struct the library takes."""

    def __init__(self, host_map: dict, device, part_sizes=None):
        import torch

        from quatrex_amd import _lib

        self.tensors = {}
        self.struct = _lib.BlockMap_t()
        for k, key in enumerate(PARTS):
            src, dst, op = host_map[key]
            # every destination appears at most once, so a full count is full coverage
            covers = part_sizes is not None and dst.size == part_sizes[k]
            t = (torch.as_tensor(src, device=device), torch.as_tensor(dst, device=device),
                 torch.as_tensor(op, device=device))
            self.tensors[key] = t
            part = self.struct.part[k]
            part.n = int(src.size)
            part.src, part.dst, part.op = (x.data_ptr() if x.numel() else None for x in t)
            part.covers_part = int(covers)

    def restricted(self, parts: str):
        """The ctypes struct with only ``parts`` (others empty)."""
        from quatrex_amd import _lib

        s = _lib.BlockMap_t()
        for k, key in enumerate(PARTS):
            if key in parts:
                s.part[k] = self.struct.part[k]
        return s
