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
