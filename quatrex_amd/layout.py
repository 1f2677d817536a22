"""Packed block-tridiagonal layout shared by the Python layer and the C-ABI.

A block-tridiagonal matrix with ``N`` diagonal blocks of sizes ``b_0..b_{N-1}``
is stored per stack entry (energy point) as three flat complex arrays:

* ``diag``  -- block ``(i, i)`` (``b_i x b_i``, row-major) at element offset
  ``diag_offsets[i] = sum_{k<i} b_k**2``;
* ``upper`` -- block ``(i, i+1)`` (``b_i x b_{i+1}``) at
  ``offdiag_offsets[i] = sum_{k<i} b_k * b_{k+1}``;
* ``lower`` -- block ``(i+1, i)`` (``b_{i+1} x b_i``) at the same offset.

A stack of ``E`` energies is ``(E, diag_size)`` / ``(E, offdiag_size)``.  For a
constant block size this is the natural ``(E, N, b, b)`` / ``(E, N-1, b, b)``
array, so contiguous numpy/torch tensors pass to the library without copies.

These are the dense blocks that ``DSDBSparse.blocks[i, j]`` returns in the
reference (``src/qttools/datastructures/dsdbsparse.py:1285-1363``,
``dsdbcoo.py:333-412``); only the tridiagonal pattern used by ``rgf.py`` is kept.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class Geometry:
    """Block partition of one device (``n_blocks`` blocks of ``block_sizes``)."""

    block_sizes: tuple[int, ...]
    diag_offsets: np.ndarray = field(init=False, repr=False, compare=False)
    offdiag_offsets: np.ndarray = field(init=False, repr=False, compare=False)
    block_offsets: np.ndarray = field(init=False, repr=False, compare=False)

    def __post_init__(self) -> None:
        bs = tuple(int(b) for b in self.block_sizes)
        if len(bs) < 1 or min(bs) < 1:
            raise ValueError(f"invalid block sizes {bs}")
        object.__setattr__(self, "block_sizes", bs)
        b = np.asarray(bs, dtype=np.int64)
        object.__setattr__(self, "diag_offsets", np.concatenate(([0], np.cumsum(b * b))))
        object.__setattr__(
            self, "offdiag_offsets", np.concatenate(([0], np.cumsum(b[:-1] * b[1:])))
        )
        object.__setattr__(self, "block_offsets", np.concatenate(([0], np.cumsum(b))))

    @classmethod
    def uniform(cls, n_blocks: int, block_size: int) -> "Geometry":
        return cls((int(block_size),) * int(n_blocks))

    @property
    def n_blocks(self) -> int:
        return len(self.block_sizes)

    @property
    def diag_size(self) -> int:
        return int(self.diag_offsets[-1])

    @property
    def offdiag_size(self) -> int:
        return int(self.offdiag_offsets[-1])

    @property
    def n(self) -> int:
        """Total matrix dimension."""
        return int(self.block_offsets[-1])

    @property
    def max_block(self) -> int:
        return max(self.block_sizes)

    @property
    def is_uniform(self) -> bool:
        return len(set(self.block_sizes)) == 1

    # --- block views into packed (..., size) arrays -------------------------
    def diag(self, arr, i: int):
        b = self.block_sizes[i]
        o = int(self.diag_offsets[i])
        return arr[..., o : o + b * b].reshape(*arr.shape[:-1], b, b)

    def upper(self, arr, i: int):
        bi, bj = self.block_sizes[i], self.block_sizes[i + 1]
        o = int(self.offdiag_offsets[i])
        return arr[..., o : o + bi * bj].reshape(*arr.shape[:-1], bi, bj)

    def lower(self, arr, i: int):
        bi, bj = self.block_sizes[i], self.block_sizes[i + 1]
        o = int(self.offdiag_offsets[i])
        return arr[..., o : o + bi * bj].reshape(*arr.shape[:-1], bj, bi)

    # --- dense conversion (tests, small sizes only) -------------------------
    def to_dense(self, diag, upper=None, lower=None) -> np.ndarray:
        """Assemble the dense ``(..., n, n)`` matrix from packed arrays (numpy)."""
        diag = np.asarray(diag)
        out = np.zeros((*diag.shape[:-1], self.n, self.n), dtype=diag.dtype)
        bo = self.block_offsets
        for i in range(self.n_blocks):
            out[..., bo[i] : bo[i + 1], bo[i] : bo[i + 1]] = self.diag(diag, i)
            if i + 1 < self.n_blocks:
                if upper is not None:
                    out[..., bo[i] : bo[i + 1], bo[i + 1] : bo[i + 2]] = self.upper(
                        np.asarray(upper), i
                    )
                if lower is not None:
                    out[..., bo[i + 1] : bo[i + 2], bo[i] : bo[i + 1]] = self.lower(
                        np.asarray(lower), i
                    )
        return out

    def from_dense(self, dense):
        """Extract packed ``(diag, upper, lower)`` from a dense ``(..., n, n)`` array."""
        dense = np.asarray(dense)
        lead = dense.shape[:-2]
        diag = np.empty((*lead, self.diag_size), dtype=dense.dtype)
        upper = np.empty((*lead, self.offdiag_size), dtype=dense.dtype)
        lower = np.empty((*lead, self.offdiag_size), dtype=dense.dtype)
        bo = self.block_offsets
        for i in range(self.n_blocks):
            self.diag(diag, i)[...] = dense[..., bo[i] : bo[i + 1], bo[i] : bo[i + 1]]
            if i + 1 < self.n_blocks:
                self.upper(upper, i)[...] = dense[..., bo[i] : bo[i + 1], bo[i + 1] : bo[i + 2]]
                self.lower(lower, i)[...] = dense[..., bo[i + 1] : bo[i + 2], bo[i] : bo[i + 1]]
        return diag, upper, lower
