"""ORACLE (test infrastructure only): the electron solver's system-matrix
assembly restated in numpy on dense stacks.

Only tests import this module.  It restates
``ElectronSolver._assemble_system_matrix`` (``src/quatrex/electron/
solver.py:668-709``) in its operation order: start from zero, add the overlap
(``_add_overlap``, ``:541-565``) or fill the diagonal with 1, scale by
``E + i eta`` (``scale_stack``, ``src/qttools/utils/stack_utils.py:6-30``),
subtract the potential (``diags(v)`` without an overlap, else
``(S_ab v_a + S_ab v_b) / 2`` per block, ``_apply_potential``, ``:567-609``),
then subtract ``Sigma^R_herm + 0.5 (Sigma^> - Sigma^<) + H`` block by block on
the tridiagonal (``_subtract_hamiltonian_and_self_energy``, ``:611-665``).
Every input is a dense ``(E, n, n)`` stack (H and S may be ``(n, n)``); the
result keeps only the tridiagonal blocks of ``block_sizes``.

Independent checks (``tests/test_assembly_oracle.py``): without overlap,
potential and self-energies it equals the synthetic generator's
``(E + i eta) I - H``; an explicit identity overlap gives the same matrix as
no overlap (the two potential branches agree); for real energies the
Hermitian part is ``E S - P - H - Sigma^R_herm`` and the anti-Hermitian part
``-(Sigma^> - Sigma^<) / 2``.
"""

from __future__ import annotations

import numpy as np


def tridiagonal_mask(block_sizes):
    off = np.concatenate(([0], np.cumsum(block_sizes)))
    n = int(off[-1])
    m = np.zeros((n, n), dtype=bool)
    for i in range(len(block_sizes)):
        for j in range(max(0, i - 1), min(len(block_sizes), i + 2)):
            m[off[i]:off[i + 1], off[j]:off[j + 1]] = True
    return m


def assemble(block_sizes, energies, eta, hamiltonian, overlap=None, potential=None,
             sigma_retarded_herm=None, sigma_lesser=None, sigma_greater=None):
    energies = np.asarray(energies, dtype=np.float64)
    E = energies.size
    mask = tridiagonal_mask(block_sizes)
    n = mask.shape[0]
    stack = lambda x: np.broadcast_to(x, (E, n, n)) if x is not None else None  # noqa: E731
    m = np.zeros((E, n, n), dtype=np.complex128)
    if overlap is None:
        m[:, np.arange(n), np.arange(n)] = 1.0
    else:
        m += np.where(mask, stack(overlap), 0)
    m *= (energies + 1j * eta)[:, None, None]
    if potential is not None:
        v = np.asarray(potential, dtype=np.float64)
        if overlap is None:
            m[:, np.arange(n), np.arange(n)] -= v
        else:
            s = np.where(mask, stack(overlap), 0)
            m -= (s * v[:, None] + s * v[None, :]) / 2
    zero = np.zeros((E, n, n), dtype=np.complex128)
    sr = stack(sigma_retarded_herm) if sigma_retarded_herm is not None else zero
    sl = stack(sigma_lesser) if sigma_lesser is not None else zero
    sg = stack(sigma_greater) if sigma_greater is not None else zero
    sub = sr + 0.5 * (sg - sl) + stack(hamiltonian)
    m -= np.where(mask, sub, 0)
    return np.where(mask, m, 0)
