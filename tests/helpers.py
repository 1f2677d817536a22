"""Seeded inputs shaped like the reference's GF-solver tests.

``bt_dense`` follows ``tests/qttools/greens_function_solver/conftest.py:58-104``:
50%-dense random complex blocks (``scipy.sparse.random``), identity added on
the diagonal blocks, then made diagonally dominant; here with a seeded
generator.  The reference draws them unseeded.
"""

import numpy as np
import scipy.sparse as sps

BLOCK_SIZES = {
    "constant-block-size": [2] * 10,
    "mixed-block-size": [2] * 3 + [4] * 2 + [2] * 3,
}
BATCHING = [1, 2, 100]
STACK_SHAPES = [(10,), (7, 2), (9, 2, 4)]


def random_block(m, n, rng):
    coo = sps.random(m, n, density=0.5, format="coo", random_state=rng).astype(np.complex128)
    coo.data += 1j * rng.uniform(size=coo.nnz)
    return coo.toarray()


def bt_dense(block_sizes, seed=0):
    rng = np.random.default_rng(seed)
    bo = np.concatenate(([0], np.cumsum(block_sizes)))
    nb = len(block_sizes)
    arr = np.zeros((bo[-1], bo[-1]), dtype=np.complex128)
    for i in range(nb):
        arr[bo[i]:bo[i + 1], bo[i]:bo[i + 1]] = random_block(block_sizes[i], block_sizes[i], rng) + np.eye(block_sizes[i])
        if i > 0:
            arr[bo[i]:bo[i + 1], bo[i - 1]:bo[i]] = random_block(block_sizes[i], block_sizes[i - 1], rng)
            arr[bo[i - 1]:bo[i], bo[i]:bo[i + 1]] = random_block(block_sizes[i - 1], block_sizes[i], rng)
    for i in range(arr.shape[0]):
        arr[i, i] = (1 + 1j) + arr[i, :].sum()
    return arr


def tridiagonal_mask(block_sizes):
    bo = np.concatenate(([0], np.cumsum(block_sizes)))
    nb = len(block_sizes)
    mask = np.zeros((bo[-1], bo[-1]), dtype=bool)
    for i in range(nb):
        for j in range(max(0, i - 1), min(nb, i + 2)):
            mask[bo[i]:bo[i + 1], bo[j]:bo[j + 1]] = True
    return mask


def block_rel_errors(x_dense, ref_dense, block_sizes):
    """Max relative Frobenius error over tridiagonal blocks (and the stack)."""
    bo = np.concatenate(([0], np.cumsum(block_sizes)))
    nb = len(block_sizes)
    worst = 0.0
    for i in range(nb):
        for j in range(max(0, i - 1), min(nb, i + 2)):
            x = x_dense[..., bo[i]:bo[i + 1], bo[j]:bo[j + 1]]
            r = ref_dense[..., bo[i]:bo[i + 1], bo[j]:bo[j + 1]]
            num = np.linalg.norm(x - r, axis=(-2, -1))
            den = np.linalg.norm(r, axis=(-2, -1))
            worst = max(worst, float(np.max(num / np.maximum(den, 1e-300))))
    return worst


def skew_tridiagonal(block_sizes, rng, scale=0.1):
    """Random skew-Hermitian tridiagonal matrix (Sigma^<> with off-diagonal blocks)."""
    n = int(np.sum(block_sizes))
    x = scale * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    x = x * tridiagonal_mask(block_sizes)
    return x - x.conj().T


def edge_case(bs, seed, E=2):
    """(geometry, A, Sigma^<, Sigma^>, OBC lists) for edge geometries (one block,
    1 x 1 blocks): a Hermitian tridiagonal H shifted off the real axis,
    skew-Hermitian Sigma with off-diagonal blocks, retarded / lesser / greater
    boundary blocks of both contacts (summed when there is one block)."""
    from quatrex_amd.layout import Geometry

    rng = np.random.default_rng(seed)
    g = Geometry(tuple(bs))
    n = g.n
    a = np.empty((E, n, n), np.complex128)
    for e in range(E):
        h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        a[e] = (0.3 * e + 4.0 + 1e-2j) * np.eye(n) - 0.25 * (h + h.conj().T) * tridiagonal_mask(bs)
    A = g.from_dense(a)
    sl = np.broadcast_to(skew_tridiagonal(bs, rng), (E, n, n))
    sg = np.broadcast_to(skew_tridiagonal(bs, rng), (E, n, n))

    def one(b, f):
        br = -0.5j * np.eye(b) + 0.05 * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b)))
        gam = 1j * (br - br.conj().T)
        rep = lambda x: np.broadcast_to(x, (E, b, b)).copy()  # noqa
        return rep(br), rep(1j * gam * f), rep(1j * gam * (f - 1))

    N = len(bs)
    L, R = one(bs[0], 0.6), one(bs[-1], 0.1)
    if N == 1:  # both contacts attach to the one block
        obc = [[L[k] + R[k]] for k in range(3)]
    else:
        obc = [[L[k]] + [None] * (N - 2) + [R[k]] for k in range(3)]
    return g, A, g.from_dense(sl)[:2], g.from_dense(sg)[:2], obc
