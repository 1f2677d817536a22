"""Dense oracles (test infrastructure only).

* :func:`dense_solve` -- ``G^R = inv(A - B^R)`` and ``G^<> = G^R (Sigma^<> + B^<>) G^R^H``
  by ``numpy.linalg.inv`` of the full matrix, the method of the reference
  tests (``tests/qttools/greens_function_solver/test_selected_inv.py:47-69``,
  ``test_selected_solve.py:56-64``) and of ``Inv`` (``inv.py:97-227``).  The OBC
  signs follow ``rgf.py:216-232`` (``Inv`` subtracts B^<>, ``inv.py:196-202``;
  the RGF convention is the one the drop-in reproduces).
* :func:`dense_current` -- Meir-Wingreen layer currents computed by brute force:
  the left-connected ``g_j``, ``g^<>_j`` are the last diagonal block of dense
  inverses of the leading ``j+1`` blocks, the boundary self-energy seen by
  block ``j`` is ``Sigma~_j = g_j^{-1} g^<>_j g_j^{-H} - S^<>_jj``, and
  ``J_j = tr(Sigma~>_j G<_jj - G>_jj Sigma~<_j)`` with the full dense G.  This
  uses none of the RGF recurrences, so it checks them (``rgf.py:388-428``).
"""

from __future__ import annotations

import numpy as np


def _dag(x):
    return x.conj().swapaxes(-2, -1)


def place_obc(dense, block_offsets, obc_blocks, sign):
    """Add ``sign * obc[i]`` onto diagonal block ``i`` of a dense stack (copy)."""
    out = np.array(dense, copy=True)
    if obc_blocks is None:
        return out
    for i, blk in enumerate(obc_blocks):
        if blk is None:
            continue
        s = slice(block_offsets[i], block_offsets[i + 1])
        out[..., s, s] += sign * blk
    return out


def dense_solve(a, sl, sg, block_offsets=None, obc=None):
    """Dense ``(G^R, G^<, G^>)`` for stacked dense inputs ``(..., n, n)``."""
    if obc is not None:
        a = place_obc(a, block_offsets, obc[0], -1.0)
        sl = place_obc(sl, block_offsets, obc[1], +1.0)
        sg = place_obc(sg, block_offsets, obc[2], +1.0)
    gr = np.linalg.inv(a)
    return gr, gr @ sl @ _dag(gr), gr @ sg @ _dag(gr)


def dense_inv(a, block_offsets=None, obc_retarded=None):
    if obc_retarded is not None:
        a = place_obc(a, block_offsets, obc_retarded, -1.0)
    return np.linalg.inv(a)


def left_connected(a, s, block_offsets, j):
    """Left-connected ``(g_j, g^<>_j)`` from dense inverses of blocks ``0..j``."""
    e = block_offsets[j + 1]
    s0 = block_offsets[j]
    g = np.linalg.inv(a[..., :e, :e])
    gs = g @ s[..., :e, :e] @ _dag(g)
    return g[..., s0:e, s0:e], gs[..., s0:e, s0:e]


def dense_current(a, sl, sg, block_offsets, obc):
    """Brute-force Meir-Wingreen currents ``(..., N+1)`` (tiny inputs only).

    ``a``, ``sl``, ``sg`` are dense without OBC; ``obc`` is the triple of
    per-block lists ``(retarded, lesser, greater)``.
    """
    nb = len(block_offsets) - 1
    at = place_obc(a, block_offsets, obc[0], -1.0)
    slt = place_obc(sl, block_offsets, obc[1], +1.0)
    sgt = place_obc(sg, block_offsets, obc[2], +1.0)
    gr = np.linalg.inv(at)
    gl = gr @ slt @ _dag(gr)
    gg = gr @ sgt @ _dag(gr)
    cur = np.zeros((*a.shape[:-2], nb + 1), dtype=np.complex128)

    def blk(m, i):
        s = slice(block_offsets[i], block_offsets[i + 1])
        return m[..., s, s]

    for j in range(1, nb):
        g, g_l = left_connected(at, slt, block_offsets, j)
        _, g_g = left_connected(at, sgt, block_offsets, j)
        ginv = np.linalg.inv(g)
        st_l = ginv @ g_l @ _dag(ginv) - blk(slt, j)
        st_g = ginv @ g_g @ _dag(ginv) - blk(sgt, j)
        cur[..., j] = np.trace(st_g @ blk(gl, j) - blk(gg, j) @ st_l, axis1=-2, axis2=-1)
    cur[..., 0] = np.trace(obc[2][0] @ blk(gl, 0) - blk(gg, 0) @ obc[1][0], axis1=-2, axis2=-1)
    cur[..., nb] = -np.trace(obc[2][-1] @ blk(gl, nb - 1) - blk(gg, nb - 1) @ obc[1][-1],
                             axis1=-2, axis2=-1)
    return cur


def rel_fro(x, ref):
    """Relative Frobenius error over the last two axes (max over the stack)."""
    num = np.linalg.norm(np.asarray(x) - ref, axis=(-2, -1))
    den = np.linalg.norm(ref, axis=(-2, -1))
    return float(np.max(num / np.maximum(den, np.finfo(float).tiny)))
