"""ORACLE (test infrastructure only): the spectral open-boundary solver restated
in numpy.

Only tests import this module.  It restates, for ``block_sections = 1``:

* ``Beyn`` (``src/qttools/nevp/beyn.py:14-379``): quadrature on the two
  circles of radii ``r_o`` and ``r_i`` (``beyn.py:183-238``), the moments
  ``P_0 = q_0 Y``, ``P_1 = q_1 Y`` (``beyn.py:307-318``), the SVD projection
  with the rank cut ``s > s_max * m_0 * eps`` (``beyn.py:101-145``) and the
  reduced eigenproblem, whose unused columns keep ``Y`` with eigenvalue 0
  (``beyn.py:240-282``).  ``Y`` is drawn from ``default_rng(seed)``; the
  reference's generator is ``default_rng(42)`` advanced by every call, so
  only quantities that do not depend on ``Y`` (the surface Green's function,
  the set of eigenvalues inside the annulus) are compared with it.
* ``Spectral.__call__`` (``src/qttools/boundary_conditions/obc/
  spectral.py:482-629``): unit-norm eigenvectors (``:353-357``), the
  reflected-mode masks (``:221-327``), ``x_ii`` from the Moore-Penrose
  pseudo-inverse of the reflected modes (``:386-436``), then
  ``num_ref_iterations`` fixed-point steps ``x = inv(a_ii - a_ji x a_ij)``
  (``:561-571, 620-629``).

Independent checks (``tests/test_spectral_oracle.py``): the surface Green's
function agrees with converged Sancho-Rubio decimation and with the closed
form of a one-dimensional chain, every returned eigenpair satisfies the
quadratic eigenvalue equation, and the eigenvalues inside the annulus are
exactly those of the linearised ``2b x 2b`` pencil (QZ, ``scipy.linalg.eigvals``).
"""

from __future__ import annotations

import numpy as np


def contour_points(r_o: float, r_i: float, num_quad_points: int):
    """Quadrature nodes on both circles and the (equal) weights (beyn.py:196-201)."""
    zeta = np.arange(num_quad_points) + 1
    phase = np.exp(2j * np.pi * (zeta + 0.5) / num_quad_points)
    return r_o * phase, r_i * phase, np.ones(num_quad_points) / num_quad_points


def beyn_moments(a_xx, r_o, r_i, num_quad_points):
    """``q_0``, ``q_1`` of beyn.py:183-238 for the coefficient blocks ``a_xx``
    (lowest to highest order, ``len(a_xx) = 2p + 1``), batched ``(E, d, d)``."""
    p = len(a_xx) // 2
    z_o, z_i, w = contour_points(r_o, r_i, num_quad_points)
    q0 = np.zeros_like(a_xx[0])
    q1 = np.zeros_like(a_xx[0])
    for k in range(num_quad_points):
        for z, sgn in ((z_o[k], 1.0), (z_i[k], -1.0)):
            op = sum(z ** n * a for a, n in zip(a_xx, range(-p, p + 1)))
            inv = np.linalg.inv(op)
            q0 += sgn * w[k] * z * inv
            q1 += sgn * w[k] * z ** 2 * inv
    return q0, q1


def beyn(a_xx, r_o, r_i, m_0, num_quad_points, seed=42):
    """``(ws, vs, n_valid)``: eigenvalues ``(E, m_0)``, eigenvectors
    ``(E, d, m_0)`` and the number of computed pairs per energy."""
    a_xx = tuple(np.asarray(a, dtype=np.complex128) for a in a_xx)
    E, d, _ = a_xx[0].shape
    rng = np.random.default_rng(seed)
    Y = rng.random((E, d, m_0)) + 1j * rng.random((E, d, m_0))
    q0, q1 = beyn_moments(a_xx, r_o, r_i, num_quad_points)
    P0 = q0 @ Y
    P1 = q1 @ Y
    u, s, vh = np.linalg.svd(P0, full_matrices=False)
    ws = np.zeros((E, m_0), dtype=np.complex128)
    vs = Y.copy()
    n_valid = np.zeros(E, dtype=np.int64)
    for e in range(E):
        eps_svd = s[e].max() * m_0 * np.finfo(np.float64).eps
        keep = np.where(s[e] > eps_svd)[0]
        ui, si, vhi = u[e][:, keep], s[e][keep], vh[e][keep, :]
        a = ui.conj().T @ P1[e] @ vhi.conj().T / si
        w, v = np.linalg.eig(a)
        ws[e, : len(w)] = w
        vs[e, :, : len(w)] = ui @ v
        n_valid[e] = len(w)
    return ws, vs, n_valid


def reflected_modes(ws, vs, a_xx, min_decay=1e-3, max_decay=6.9, min_propagation=0.01,
                    residual_tolerance=1e-3, residual_normalization=True, eta_decay=1e-14):
    """The reflected-mode mask of spectral.py:221-327 (``find_injected`` off)."""
    p = len(a_xx) // 2
    with np.errstate(all="ignore"):
        products = sum(a @ vs * ws[..., np.newaxis, :] ** (i - p) for i, a in enumerate(a_xx))
        residuals = np.linalg.norm(products, axis=-2) / np.linalg.norm(vs, axis=-2)
        if residual_normalization:
            residuals /= np.abs(ws)
        dEk_dk = -sum((1j * n) * np.diagonal(vs.conj().swapaxes(-1, -2) @ a @ vs, axis1=-2, axis2=-1)
                      * ws ** n for a, n in zip(a_xx, range(-p, p + 1)))
        ks = -1j * np.log(ws)
    dEk_dk = np.nan_to_num(dEk_dk, nan=0, posinf=0, neginf=0)
    ks = np.nan_to_num(ks, nan=0, posinf=0, neginf=0)
    eta = np.finfo(dEk_dk.dtype).eps
    ratio = np.abs(dEk_dk.real) / (np.abs(dEk_dk.imag) + eta)
    propagating = (np.abs(ks.imag) < min_decay) & (min_propagation < ratio) & (dEk_dk.real < 0)
    decaying = ks.imag < -min_decay
    decaying |= (min_propagation >= ratio) & (ks.imag < -eta_decay)
    decaying &= ks.imag > -max_decay
    return (propagating | decaying) & (residuals < residual_tolerance)


def spectral_x_ii(a_ii, a_ij, a_ji, r_o=10.0, r_i=0.8, m_0=10, num_quad_points=20, seed=42,
                  num_ref_iterations=2, **mode_kw):
    """Surface Green's function of ``Spectral(Beyn(...))`` (spectral.py:482-629),
    batched ``(E, b, b)``; returns ``(x_ii, mask, ws, vs)``."""
    a_ii, a_ij, a_ji = (np.asarray(x, dtype=np.complex128) for x in (a_ii, a_ij, a_ji))
    layer = (a_ji, a_ii, a_ij)
    ws, vs, _ = beyn(layer, r_o, r_i, m_0, num_quad_points, seed=seed)
    with np.errstate(all="ignore"):
        vs = vs / np.linalg.norm(vs, axis=-2, keepdims=True)
    mask = reflected_modes(ws, vs, layer, **mode_kw)
    x_ij = np.zeros_like(a_ij)
    for e in range(a_ii.shape[0]):
        vr = vs[e][:, mask[e]]
        w = ws[e][mask[e]]
        v_inv = np.linalg.inv(vr.conj().T @ vr) @ vr.conj().T
        x_ij[e] = vr / w @ v_inv
    x = np.linalg.inv(a_ii + a_ji @ x_ij)
    for _ in range(num_ref_iterations - 1):
        x = np.linalg.inv(a_ii - a_ji @ x @ a_ij)
    x = np.linalg.inv(a_ii - a_ji @ x @ a_ij)
    return x, mask, ws, vs


def companion_eigenvalues(a_ji, a_ii, a_ij):
    """All finite eigenvalues w of ``(a_ji / w + a_ii + a_ij w) v = 0`` for one
    energy, from the linearisation ``[[0, I], [-a_ji, -a_ii]] u =
    w [[I, 0], [0, a_ij]] u`` by the QZ algorithm (``scipy.linalg.eigvals``;
    a singular ``a_ij`` gives infinite eigenvalues, which are dropped)."""
    import scipy.linalg

    d = a_ii.shape[-1]
    A = np.zeros((2 * d, 2 * d), dtype=np.complex128)
    B = np.zeros((2 * d, 2 * d), dtype=np.complex128)
    A[:d, d:] = np.eye(d)
    A[d:, :d] = -a_ji
    A[d:, d:] = -a_ii
    B[:d, :d] = np.eye(d)
    B[d:, d:] = a_ij
    w = scipy.linalg.eigvals(A, B)
    return w[np.isfinite(w)]
