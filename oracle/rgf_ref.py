"""Numpy restatement of the reference RGF solver (test infrastructure only).

Restates ``RGF.selected_inv`` (``src/qttools/greens_function_solver/rgf.py:25-113``)
and ``RGF.selected_solve`` (``rgf.py:115-438``) with the same operation order,
on :class:`BlockTri`, a minimal stand-in for the ``DSDBSparse`` block interface
(``.blocks[i, j]``, ``.stack[slice]``, ``zeros_like``, ``symmetry``) restricted
to the tridiagonal pattern.

Reproduced reference conventions:

* OBC: ``A_jj - B^R_j`` and ``Sigma^<_jj + B^<_j`` (``rgf.py:216-262``);
* only the upper blocks ``Sigma_{i,i+1}`` are read (skew-Hermitian input);
* stored diagonal G^<, G^> are ``(X - X^H)/2``; the carried X is unprojected
  (``rgf.py:305-310, 362-364, 384-386``);
* ``G^<_{j,i} = -G^<_{i,j}^H`` written only for non-symmetric storage
  (``rgf.py:357-359, 379-381``);
* ``selected_solve`` writes only diagonal G^R (``rgf.py:311-312, 410-412``)
  unless ``offdiag_retarded=True`` (then ``G_ij = -g_i A_ij G_jj`` and
  ``G_ji = -G_jj A_ji g_i``, the ``selected_inv`` formulas ``rgf.py:105-107``);
* current (``rgf.py:388-428``) from the left-connected ``g``, ``g^<``, ``g^>``.
  Deviation: the boundary currents ``J_0``, ``J_N`` are evaluated per batch;
  the reference evaluates them once after the batch loop with the last
  batch's blocks, which only lines up when there is a single batch.
"""

from __future__ import annotations

import numpy as np

from oracle.batching import batches


def _dag(x):
    return x.conj().swapaxes(-2, -1)


class BlockTri:
    """Dense tridiagonal blocks over a stack, indexed like ``DSDBSparse``."""

    def __init__(self, block_sizes, stack_shape, dtype=np.complex128, symmetry=False,
                 symmetry_op=np.conj, _data=None, _stack_index=None):
        self.block_sizes = np.asarray(block_sizes, dtype=np.int64)
        self.num_blocks = len(self.block_sizes)
        self.block_offsets = np.concatenate(([0], np.cumsum(self.block_sizes)))
        self.stack_shape = tuple(stack_shape)
        self.dtype = np.dtype(dtype)
        self.symmetry = symmetry
        self.symmetry_op = symmetry_op
        n = int(self.block_offsets[-1])
        self.shape = (*self.stack_shape, n, n)
        if _data is None:
            _data = {}
            for i in range(self.num_blocks):
                for j in range(max(0, i - 1), min(self.num_blocks, i + 2)):
                    if symmetry and j < i:
                        continue
                    _data[(i, j)] = np.zeros(
                        (*self.stack_shape, self.block_sizes[i], self.block_sizes[j]), dtype
                    )
        self._data = _data
        self._stack_index = _stack_index if _stack_index is not None else slice(None)

    # --- reference-like API -------------------------------------------------
    @classmethod
    def zeros_like(cls, other: "BlockTri") -> "BlockTri":
        return cls(other.block_sizes, other.stack_shape, other.dtype, other.symmetry,
                   other.symmetry_op)

    @classmethod
    def from_dense(cls, dense, block_sizes, stack_shape, symmetry=False):
        obj = cls(block_sizes, stack_shape, np.asarray(dense).dtype, symmetry)
        bo = obj.block_offsets
        for (i, j), arr in obj._data.items():
            arr[...] = dense[..., bo[i]:bo[i + 1], bo[j]:bo[j + 1]]
        return obj

    def to_dense(self):
        out = np.zeros(self.shape, self.dtype)
        bo = self.block_offsets
        for i in range(self.num_blocks):
            for j in range(max(0, i - 1), min(self.num_blocks, i + 2)):
                out[..., bo[i]:bo[i + 1], bo[j]:bo[j + 1]] = self.blocks[i, j]
        return out

    @property
    def stack(self):
        return _StackIndexer(self)

    @property
    def blocks(self):
        return _BlockIndexer(self)


class _StackIndexer:
    def __init__(self, mat):
        self._m = mat

    def __getitem__(self, sl):
        m = self._m
        view = BlockTri.__new__(BlockTri)
        view.__dict__.update(m.__dict__)
        view._data = {k: v[sl] for k, v in m._data.items()}
        n0 = len(range(*sl.indices(m.stack_shape[0])))
        view.stack_shape = (n0, *m.stack_shape[1:])
        view.shape = (*view.stack_shape, *m.shape[-2:])
        return view


class _BlockIndexer:
    def __init__(self, mat):
        self._m = mat

    def __getitem__(self, ij):
        i, j = ij
        m = self._m
        if m.symmetry and j < i:
            return np.ascontiguousarray(m.symmetry_op(m._data[(j, i)].swapaxes(-1, -2)))
        return m._data[(i, j)].copy()

    def __setitem__(self, ij, val):
        i, j = ij
        m = self._m
        if m.symmetry and j < i:
            m._data[(j, i)][...] = m.symmetry_op(np.asarray(val).swapaxes(-1, -2))
            return
        if (i, j) in m._data:  # blocks outside the pattern are dropped
            m._data[(i, j)][...] = val


def selected_inv(a: BlockTri, obc_retarded=None, out=None, max_batch_size=100):
    """Restatement of ``RGF.selected_inv`` (``rgf.py:25-113``)."""
    nb = a.num_blocks
    obc_retarded = obc_retarded if obc_retarded is not None else [None] * nb
    x = out if out is not None else BlockTri.zeros_like(a)
    _, offs = batches(a.shape[0], max_batch_size)
    for bidx in range(len(offs) - 1):
        sl = slice(int(offs[bidx]), int(offs[bidx + 1]))
        a_, x_ = a.stack[sl], x.stack[sl]
        g = [None] * nb

        def a_tilde(j):
            blk = a_.blocks[j, j]
            return blk if obc_retarded[j] is None else blk - obc_retarded[j][sl]

        g[0] = np.linalg.inv(a_tilde(0))
        for i in range(nb - 1):
            j = i + 1
            g[j] = np.linalg.inv(a_tilde(j) - a_.blocks[j, i] @ g[i] @ a_.blocks[i, j])
        x_.blocks[nb - 1, nb - 1] = g[-1]
        for i in range(nb - 2, -1, -1):
            j = i + 1
            gii, xjj, aij = g[i], g[j], a_.blocks[i, j]
            xji = -xjj @ a_.blocks[j, i] @ gii
            x_.blocks[j, i] = xji
            x_.blocks[i, j] = -gii @ aij @ xjj
            g[i] = gii - gii @ aij @ xji
            x_.blocks[i, i] = g[i]
    if out is None:
        return x
    return None


def selected_solve(a: BlockTri, sigma_lesser: BlockTri, sigma_greater: BlockTri,
                   obc=None, out=None, return_retarded=False, return_current=False,
                   max_batch_size=100, offdiag_retarded=False):
    """Restatement of ``RGF.selected_solve`` (``rgf.py:115-438``).

    ``obc`` is ``None`` or a triple of per-block lists ``(retarded, lesser,
    greater)`` like ``OBCBlocks``.  Returns like the reference: the new
    ``(xl, xg[, xr][, current])`` when ``out`` is None, else ``None`` or the
    current.
    """
    nb = sigma_lesser.num_blocks
    if obc is None:
        obc = ([None] * nb, [None] * nb, [None] * nb)
    obc_r, obc_l, obc_g = obc
    if return_current:
        current = np.zeros((*sigma_lesser.shape[:-2], nb + 1), dtype=sigma_lesser.dtype)
    _, offs = batches(sigma_lesser.shape[0], max_batch_size)
    if out is not None:
        xl, xg, *rest = out
        if return_retarded:
            if len(rest) != 1:
                raise ValueError("Invalid number of output matrices.")
            xr = rest[0]
    else:
        xl, xg = BlockTri.zeros_like(a), BlockTri.zeros_like(a)
        if return_retarded:
            xr = BlockTri.zeros_like(a)

    for bidx in range(len(offs) - 1):
        sl = slice(int(offs[bidx]), int(offs[bidx + 1]))
        a_, sl_, sg_ = a.stack[sl], sigma_lesser.stack[sl], sigma_greater.stack[sl]
        xl_, xg_ = xl.stack[sl], xg.stack[sl]
        xr_ = xr.stack[sl] if return_retarded else None

        def diag_terms(j):
            aj = a_.blocks[j, j] if obc_r[j] is None else a_.blocks[j, j] - obc_r[j][sl]
            lj = sl_.blocks[j, j] if obc_l[j] is None else sl_.blocks[j, j] + obc_l[j][sl]
            gj = sg_.blocks[j, j] if obc_g[j] is None else sg_.blocks[j, j] + obc_g[j][sl]
            return aj, lj, gj

        gr, gl, gg = [None] * nb, [None] * nb, [None] * nb
        a00, s_l, s_g = diag_terms(0)
        gr[0] = np.linalg.inv(a00)
        gl[0] = gr[0] @ s_l @ _dag(gr[0])
        gg[0] = gr[0] @ s_g @ _dag(gr[0])

        # Forward sweep (rgf.py:241-302).
        for i in range(nb - 1):
            j = i + 1
            ajj, s_l, s_g = diag_terms(j)
            aji = a_.blocks[j, i]
            aji_h = _dag(aji)
            p = aji @ gr[i]
            gr[j] = np.linalg.inv(ajj - p @ a_.blocks[i, j])
            gj_h = _dag(gr[j])
            q = p @ sl_.blocks[i, j]
            gl[j] = gr[j] @ (s_l + aji @ gl[i] @ aji_h + _dag(q) - q) @ gj_h
            q = p @ sg_.blocks[i, j]
            gg[j] = gr[j] @ (s_g + aji @ gg[i] @ aji_h + _dag(q) - q) @ gj_h

        last = nb - 1
        xl_.blocks[last, last] = 0.5 * (gl[-1] - _dag(gl[-1]))
        xg_.blocks[last, last] = 0.5 * (gg[-1] - _dag(gg[-1]))
        if return_retarded:
            xr_.blocks[last, last] = gr[-1]

        # Backward sweep (rgf.py:315-412).
        for i in range(nb - 2, -1, -1):
            j = i + 1
            g_i, G_jj = gr[i], gr[j]
            aij, aji = a_.blocks[i, j], a_.blocks[j, i]
            l_i, L_jj, r_i, R_jj = gl[i], gl[j], gg[i], gg[j]
            s_lij, s_gij = sl_.blocks[i, j], sg_.blocks[i, j]

            G_jj_h = _dag(G_jj)
            u = g_i @ aij                      # g_i A_ij
            u_h = _dag(u)
            r = G_jj @ aji                     # G_jj A_ji
            r_h = _dag(r)
            m = u @ G_jj                       # g_i A_ij G_jj
            m_h = _dag(m)
            k = u @ r                          # g_i A_ij G_jj A_ji

            new = {}
            for key, lc, Lc, s_ij, target in (("l", l_i, L_jj, s_lij, xl_),
                                              ("g", r_i, R_jj, s_gij, xg_)):
                t1 = k @ lc - g_i @ s_ij @ m_h
                t1 = t1 - _dag(t1)
                t2 = u @ Lc
                x_ij = -t2 - lc @ r_h + g_i @ s_ij @ G_jj_h
                target.blocks[i, j] = x_ij
                if not target.symmetry:
                    target.blocks[j, i] = -_dag(x_ij)
                new[key] = lc + t2 @ u_h + t1
                target.blocks[i, i] = 0.5 * (new[key] - _dag(new[key]))

            if return_current:
                p = aji @ g_i
                q = p @ s_lij
                st_l = aji @ l_i @ _dag(aji) + _dag(q) - q
                q = p @ s_gij
                st_g = aji @ r_i @ _dag(aji) + _dag(q) - q
                current[sl, ..., j] = np.trace(st_g @ gl[j] - gg[j] @ st_l,
                                               axis1=-2, axis2=-1)

            gl[i], gg[i] = new["l"], new["g"]
            if return_retarded and offdiag_retarded:
                xr_.blocks[i, j] = -m
                xr_.blocks[j, i] = -r @ g_i
            gr[i] = g_i + k @ g_i
            if return_retarded:
                xr_.blocks[i, i] = gr[i]

        if return_current:
            if obc_l[0] is None or obc_g[0] is None or obc_l[-1] is None or obc_g[-1] is None:
                raise TypeError("the current needs lesser/greater OBC blocks at both ends")
            current[sl, ..., 0] = np.trace(obc_g[0][sl] @ gl[0] - gg[0] @ obc_l[0][sl],
                                           axis1=-2, axis2=-1)
            current[sl, ..., -1] = -np.trace(obc_g[-1][sl] @ gl[-1] - gg[-1] @ obc_l[-1][sl],
                                             axis1=-2, axis2=-1)

    if out is None:
        res = (xl, xg) + ((xr,) if return_retarded else ())
        return res + ((current,) if return_current else ())
    if return_current:
        return current
    return None
