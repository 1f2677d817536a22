"""ORACLE (test infrastructure only): how many bits per product operand the
selected solve needs, by the reference recurrences with a pluggable product.

Only tests import this module.  ``solve_lesser`` runs the forward and
backward recurrences of ``RGF.selected_solve`` (``src/qttools/
greens_function_solver/rgf.py:216-412``, G^R diagonal and G^< diagonal and
upper blocks, no OBC, no current) for one energy with every block product
replaced by ``mm(A, B)``.  Two products model reduced precision:

* ``fp32_product``: operands and result rounded to complex64 (what the FP32
  matrix cores compute, up to the accumulation order);
* ``truncated_product(k)``: each row of A and each column of B scaled by a
  power of two so its largest component has ``k`` bits, rounded to integers,
  then multiplied exactly - the input truncation of an integer-emulated FP64
  product (Ozaki scheme II, Ozaki, Uchino and Imamura 2025), whose modular
  reconstruction is exact once the moduli cover ``2k + log2(K) + 1`` bits.

The inversions stay in FP64 (``numpy.linalg.inv``), as in the mixed path.
"""

from __future__ import annotations

import numpy as np


def fp32_product(A, B):
    return (A.astype(np.complex64) @ B.astype(np.complex64)).astype(np.complex128)


def _trunc_rows(A, k):
    m = np.maximum(np.abs(A.real).max(axis=-1), np.abs(A.imag).max(axis=-1))
    s = np.exp2(np.ceil(np.log2(np.maximum(m, 1e-300))) - k)[..., None]
    return (np.round(A.real / s) + 1j * np.round(A.imag / s)) * s


def truncated_product(k):
    return lambda A, B: _trunc_rows(A, k) @ _trunc_rows(B.T, k).T


def exact_product(A, B):
    return A @ B


def solve_lesser(a_d, a_u, a_l, s_d, s_u, mm):
    """``(GR_diag, GL_diag, GL_upper)`` lists for one energy from the lists of
    diagonal, upper and lower blocks of A and diagonal / upper blocks of
    Sigma^<."""
    N = len(a_d)
    inv = np.linalg.inv
    dag = lambda x: x.conj().T  # noqa: E731
    g = [None] * N
    gl = [None] * N
    g[0] = inv(a_d[0])
    gl[0] = mm(mm(g[0], s_d[0]), dag(g[0]))
    for j in range(1, N):
        i = j - 1
        P = mm(a_l[i], g[i])
        g[j] = inv(a_d[j] - mm(P, a_u[i]))
        Y = mm(a_l[i], gl[i])
        Q = mm(P, s_u[i])
        W = mm(Y, dag(a_l[i])) + dag(Q) - Q + s_d[j]
        gl[j] = mm(mm(g[j], W), dag(g[j]))
    GR, GL = [None] * N, [None] * N
    GR[N - 1], GL[N - 1] = g[N - 1], gl[N - 1]
    GLu = [None] * (N - 1)
    for i in range(N - 2, -1, -1):
        j = i + 1
        U = mm(g[i], a_u[i])
        R = mm(GR[j], a_l[i])
        K = mm(U, R)
        GR[i] = g[i] + mm(K, g[i])
        B = mm(mm(g[i], s_u[i]), dag(GR[j]))
        T2 = mm(U, GL[j])
        GLu[i] = -mm(gl[i], dag(R)) + B - T2
        T1 = mm(K, gl[i]) - mm(B, dag(U))
        GL[i] = gl[i] + mm(T2, dag(U)) + T1 - dag(T1)
    return GR, GL, GLu


def max_block_error(res, ref):
    return max(np.linalg.norm(x - r) / np.linalg.norm(r) for X, R in zip(res, ref) for x, r in zip(X, R))
