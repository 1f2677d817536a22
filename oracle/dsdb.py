"""ORACLE (test infrastructure only): block access of the reference's
block-sparse stacks, restated in numpy.

Only tests import this module.  It restates, with plain loops:

* ``DSDBCOO.from_sparray`` (``src/qttools/datastructures/dsdbcoo.py:885-977``):
  canonical COO (``sum_duplicates``), upper triangle only for symmetric
  storage (``sparse.triu``), entries sorted by block row, block column
  (``kernels/datastructure/numba/dsdbcoo.py:175-235``), then row and column
  as in canonical COO; ``data`` repeats the values for every stack element.
* ``DSDBCSR.from_sparray`` (``dsdbcsr.py:800-863``): the same order, with a
  ``rowptr`` per stored block holding absolute positions.
* ``_get_block`` (``dsdbcoo.py:333-412``, ``dsdbcsr.py:285-345``): zeros,
  then the stored entries of the block (``densify_block``,
  ``numba/dsdbcoo.py:106-150``); with symmetric storage a block below the
  diagonal is ``symmetry_op(block(j, i)^T)`` and a diagonal block becomes
  ``X + symmetry_op(X^T)`` with its diagonal halved.
* ``_set_block`` (``dsdbcoo.py:466-521``, ``dsdbcsr.py:420-472``): the
  stored entries of the block take the dense values (``sparsify_block``);
  entries outside the pattern are dropped; with symmetric storage a block
  below the diagonal writes ``symmetry_op(value^T)`` into block (j, i).
"""

from __future__ import annotations

import numpy as np
from scipy import sparse


class OracleDSDB:
    """Local (single-rank) block-sparse stack: ``data[*stack, nnz]``."""

    def __init__(self, data, rows, cols, block_sizes, symmetry=False, symmetry_op=np.conj):
        self.data = np.array(data, dtype=np.complex128)
        self.rows = np.asarray(rows, dtype=np.int64)
        self.cols = np.asarray(cols, dtype=np.int64)
        self.block_sizes = np.asarray(block_sizes, dtype=np.int64)
        self.block_offsets = np.concatenate(([0], np.cumsum(self.block_sizes)))
        self.symmetry = symmetry
        self.symmetry_op = symmetry_op

    @property
    def num_blocks(self) -> int:
        return int(self.block_sizes.size)

    @property
    def shape(self) -> tuple:
        n = int(self.block_offsets[-1])
        return (*self.data.shape[:-1], n, n)

    @classmethod
    def from_sparray(cls, arr, block_sizes, stack_shape, symmetry=False, symmetry_op=np.conj):
        coo = sparse.coo_matrix(arr).copy()
        if symmetry:
            coo = sparse.triu(coo, format="coo")
        coo.sum_duplicates()
        off = np.concatenate(([0], np.cumsum(block_sizes)))
        nb = len(block_sizes)
        order = []
        # block row, then block column; inside a block, canonical (row, col) order
        for bi in range(nb):
            for bj in range(nb):
                sel = [k for k in range(coo.nnz)
                       if off[bi] <= coo.row[k] < off[bi + 1] and off[bj] <= coo.col[k] < off[bj + 1]]
                sel.sort(key=lambda k: (coo.row[k], coo.col[k]))
                order.extend(sel)
        order = np.asarray(order, dtype=np.int64)
        data = np.broadcast_to(coo.data[order], tuple(stack_shape) + (order.size,)).copy()
        return cls(data, coo.row[order], coo.col[order], block_sizes, symmetry, symmetry_op)

    def rowptr_map(self) -> dict:
        """The DSDBCSR view of the same storage: rowptr per stored block."""
        off = self.block_offsets
        out = {}
        k = 0
        n = self.rows.size
        while k < n:
            bi = int(np.searchsorted(off, self.rows[k], side="right") - 1)
            bj = int(np.searchsorted(off, self.cols[k], side="right") - 1)
            rp = [k]
            for r in range(off[bi], off[bi + 1]):
                while (k < n and self.rows[k] == r and off[bj] <= self.cols[k] < off[bj + 1]):
                    k += 1
                rp.append(k)
            out[(bi, bj)] = np.asarray(rp, dtype=np.int64)
        return out

    def _block_entries(self, i, j):
        off = self.block_offsets
        return [k for k in range(self.rows.size)
                if off[i] <= self.rows[k] < off[i + 1] and off[j] <= self.cols[k] < off[j + 1]]

    def get_block(self, i, j):
        if self.symmetry and j < i:
            return np.ascontiguousarray(self.symmetry_op(np.swapaxes(self.get_block(j, i), -1, -2)))
        off = self.block_offsets
        lead = self.data.shape[:-1]
        block = np.zeros(lead + (int(self.block_sizes[i]), int(self.block_sizes[j])), np.complex128)
        for k in self._block_entries(i, j):
            block[..., self.rows[k] - off[i], self.cols[k] - off[j]] = self.data[..., k]
        if self.symmetry and i == j:
            block = block + self.symmetry_op(np.swapaxes(block, -1, -2))
            idx = np.arange(block.shape[-1])
            block[..., idx, idx] /= 2
        return block

    def set_block(self, i, j, value):
        value = np.asarray(value, dtype=np.complex128)
        if self.symmetry and j < i:
            self.set_block(j, i, self.symmetry_op(np.swapaxes(value, -1, -2)))
            return
        off = self.block_offsets
        value = np.broadcast_to(value, self.data.shape[:-1] + value.shape[-2:])
        for k in self._block_entries(i, j):
            self.data[..., k] = value[..., self.rows[k] - off[i], self.cols[k] - off[j]]
