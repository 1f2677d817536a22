"""ORACLE (test infrastructure only): open boundary conditions restated in numpy.

Only tests import this module.  It restates:

* ``SanchoRubio.__call__`` (``src/qttools/boundary_conditions/obc/
  sancho_rubio.py:33-85``): the decimation loop with the batch-wide stopping
  rule (max over the stack of ``||(|alpha| + |beta|)||_F / 2``), the warning
  when no iteration met it, ``x_ii = inv(epsilon_s)``.
* ``OBCSystem._fixed_point_step`` / ``_get_starting_guess`` /
  ``_get_residuals`` (``obc.py:102-202``) and the memoising wrapper
  ``BaseBoundarySystem._solve`` / ``_memoized_solve`` / ``__call__``
  (``boundary_system.py:196-373``) for one process (stack communicator of
  size 1) and the identity system reducer.
* ``ElectronSolver._compute_contact_obc`` (``src/quatrex/electron/
  solver.py:420-495``) for ``block_sections = 1``, with or without an
  overlap matrix (whose blocks the reference adds unreversed, also for the
  right contact):
  ``sigma = m_10 g m_01``, ``gamma = i (sigma - sigma^H)``,
  ``lesser = i f gamma``, ``greater = i (f - 1) gamma``; the right contact
  works on the block-reversed boundary blocks (``order="reverse"``).

Independent checks (``tests/test_obc_oracle.py``): the closed-form surface
Green's function of a one-dimensional chain, the last diagonal block of the
inverse of a long finite chain (brute force), and the fixed-point identity
``x = inv(a_ii - a_ji x a_ij)`` at convergence.
"""

from __future__ import annotations

import warnings

import numpy as np


def sancho_rubio(a_ii, a_ij, a_ji, max_iterations: int = 100, convergence_tol: float = 1e-6):
    """Returns ``(x_ii, iterations, delta)``; warns when not converged."""
    epsilon = np.array(a_ii, dtype=np.complex128)
    epsilon_s = epsilon.copy()
    alpha = np.array(a_ji, dtype=np.complex128)
    beta = np.array(a_ij, dtype=np.complex128)
    delta = np.nan
    iterations = 0
    converged = False
    for _ in range(max_iterations):
        inverse = np.linalg.inv(epsilon)
        epsilon = epsilon - alpha @ inverse @ beta - beta @ inverse @ alpha
        epsilon_s = epsilon_s - alpha @ inverse @ beta
        alpha = alpha @ inverse @ alpha
        beta = beta @ inverse @ beta
        delta = np.linalg.norm(np.abs(alpha) + np.abs(beta), axis=(-2, -1)).max() / 2
        iterations += 1
        if delta < convergence_tol:
            converged = True
            break
    if not converged:
        warnings.warn("Surface Green's function did not converge.", RuntimeWarning)
    return np.linalg.inv(epsilon_s), iterations, delta


def fixed_point_step(a_ii, a_ij, a_ji, x):
    return np.linalg.inv(a_ii - a_ji @ x @ a_ij)


def residuals(a_ii, a_ij, a_ji, x):
    """``(rel, abs, x_ref)`` with ``x_ref`` one fixed-point step from ``x``."""
    x_ref = fixed_point_step(a_ii, a_ij, a_ji, x)
    abs_res = np.linalg.norm(x_ref - x, axis=(-2, -1))
    return abs_res / np.linalg.norm(x_ref, axis=(-2, -1)), abs_res, x_ref


class OracleOBCSystem:
    """The memoising boundary-system wrapper for one process."""

    def __init__(self, max_iterations=100, convergence_tol=1e-6, num_ref_iterations=2,
                 relative_tol=2e-1, absolute_tol=1e-6, warning_threshold=1e-1,
                 memoization_mode="auto", agreement_threshold=0.999):
        self.max_iterations = max_iterations
        self.convergence_tol = convergence_tol
        self.num_ref_iterations = max(2, num_ref_iterations)
        self.relative_tol = relative_tol
        self.absolute_tol = absolute_tol
        self.warning_threshold = warning_threshold
        self.memoization_mode = memoization_mode
        self.agreement_threshold = agreement_threshold
        self._cache = {}
        self.events = []  # "solve" / "memo" per call, for comparison with the product

    def _solve(self, system, contact):
        x, _, _ = sancho_rubio(*system, self.max_iterations, self.convergence_tol)
        self.events.append("solve")
        if self.memoization_mode != "off":
            self._cache[contact] = x.copy()
        return x

    def _memoized_solve(self, system, contact):
        if self.memoization_mode == "off":
            return self._solve(system, contact)
        solution = self._cache.get(contact)
        if solution is None:
            if self.memoization_mode in ("auto", "force-after-first"):
                return self._solve(system, contact)
            if self.memoization_mode == "force":
                solution = np.linalg.inv(system[0])
            else:
                raise RuntimeError(f"Invalid memoizing mode: {self.memoization_mode}")
        rel, abs_res, solution = residuals(*system, solution)
        if self.memoization_mode == "auto":
            # count of converged stack entries over the communicator size (1)
            count = int(np.sum((abs_res < self.absolute_tol) | (rel < self.relative_tol)))
            if count / 1 > self.agreement_threshold:
                return self._solve(system, contact)
        self.events.append("memo")
        for _ in range(self.num_ref_iterations - 2):
            solution = fixed_point_step(*system, solution)
        return solution

    def __call__(self, system, contact):
        solution = self._memoized_solve(system, contact)
        rel, abs_res, solution = residuals(*system, solution)
        if np.any((rel > self.warning_threshold) & (abs_res > self.absolute_tol)):
            warnings.warn(f"High error for {contact}", RuntimeWarning)
        if self.memoization_mode != "off":
            self._cache[contact] = solution.copy()
        return solution, rel, abs_res


def contact_obc(m_10, m_00, m_01, occupancies, eta_obc, obc, contact="left", overlap=None):
    """``(sigma, lesser, greater)`` of one contact; ``obc(system, contact)``
    returns ``(x, rel, abs)`` like :class:`OracleOBCSystem`.  ``overlap``:
    ``None`` or ``(s_10, s_00, s_01)``, added without reversal (as the
    reference does for both contacts)."""
    flip = (lambda x: x[..., ::-1, ::-1].copy()) if contact == "right" else (lambda x: x)
    m_10, m_00, m_01 = flip(m_10), flip(m_00), flip(m_01)
    if overlap is None:
        s_10 = np.zeros_like(m_10)
        s_00 = 1j * eta_obc * np.eye(m_00.shape[-1], dtype=np.complex128)
        s_01 = np.zeros_like(m_01)
    else:
        s_10, s_00, s_01 = (1j * eta_obc * np.asarray(s) for s in overlap)
    g_00, *_ = obc((m_00 + s_00, m_01 + s_01, m_10 + s_10), contact)
    sigma = m_10 @ g_00 @ m_01
    gamma = 1j * (sigma - sigma.conj().swapaxes(-2, -1))
    f = np.asarray(occupancies, dtype=np.float64)
    lesser = 1j * (gamma * f[:, None, None])
    greater = 1j * (gamma * (f - 1)[:, None, None])
    return flip(sigma), flip(lesser), flip(greater)


def boundary_blocks(diag_blocks, upper_blocks, lower_blocks, contact):
    """``(m_10, m_00, m_01)`` of a contact from lists of ``(E, b, b)`` blocks
    (the unflipped blocks the reference reads before ``order_block``)."""
    if contact == "left":
        return lower_blocks[0], diag_blocks[0], upper_blocks[0]
    n = len(diag_blocks) - 1
    # right: a_ji = A[n-1, n], a_ii = A[n, n], a_ij = A[n, n-1]
    return upper_blocks[n - 1], diag_blocks[n], lower_blocks[n - 1]


def chain_surface_gf(a_ii, a_ij, a_ji, length: int):
    """Brute force: the last diagonal block of the inverse of a finite chain of
    ``length`` cells (diagonal a_ii, upper a_ij, lower a_ji), one matrix."""
    b = a_ii.shape[-1]
    n = b * length
    m = np.zeros((n, n), dtype=np.complex128)
    for k in range(length):
        m[k * b:(k + 1) * b, k * b:(k + 1) * b] = a_ii
        if k + 1 < length:
            m[k * b:(k + 1) * b, (k + 1) * b:(k + 2) * b] = a_ij
            m[(k + 1) * b:(k + 2) * b, k * b:(k + 1) * b] = a_ji
    return np.linalg.inv(m)[-b:, -b:]


def chain_surface_gf_closed_form(z, eps0, t):
    """Surface Green's function of the 1-D chain with on-site ``eps0`` and
    hopping ``t`` at complex energy ``z`` (Im z > 0): the root of
    ``t^2 x^2 - (z - eps0) x + 1 = 0`` with ``|t x| < 1``."""
    a = z - eps0
    r = np.sqrt(a * a - 4 * t * t + 0j)
    x1 = (a + r) / (2 * t * t)
    x2 = (a - r) / (2 * t * t)
    return x1 if abs(t * x1) < abs(t * x2) else x2
