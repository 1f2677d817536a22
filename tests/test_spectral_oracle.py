"""The spectral open-boundary oracle (Beyn + reflected modes) against closed
forms, brute force, an independent eigen-solver and converged decimation."""

import numpy as np
import pytest

from oracle import obc as O
from oracle import spectral as SP


def low_rank_lead(b, r, energies, eta, seed, coupling=0.6):
    """A lead whose cells couple through r orbitals only (last r of a cell to
    the first r of the next), as tight-binding leads do: 2r finite non-zero
    eigenvalues, the rest at 0 and infinity, so Beyn's subspace (at most b
    modes) holds every mode inside the annulus."""
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b))) / np.sqrt(2 * b)
    h00 = (x + x.conj().T) / 2
    h01 = np.zeros((b, b), dtype=np.complex128)
    h01[b - r:, :r] = coupling * (rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r))) / np.sqrt(2 * r)
    energies = np.asarray(energies, dtype=np.float64)
    a_ii = (energies[:, None, None] + 1j * eta) * np.eye(b) - h00
    a_ij = np.broadcast_to(-h01, a_ii.shape).copy()
    a_ji = np.broadcast_to(-h01.conj().T, a_ii.shape).copy()
    return a_ii, a_ij, a_ji


def contour_radii(a_ii, a_ij, a_ji):
    """Radii (r_i, r_o) of Beyn's annulus holding every finite non-zero
    eigenvalue of every energy with a factor 2 to spare on both sides (from the
    QZ eigenvalues of the linearised pencil), so that no decaying mode is
    left outside and the trapezoidal rule converges like 2^-num_quad_points.
    Physical leads are run with radii chosen the same way by hand."""
    mods = np.concatenate([np.abs(SP.companion_eigenvalues(a_ji[e], a_ii[e], a_ij[e]))
                           for e in range(a_ii.shape[0])])
    mods = mods[mods > 1e-12]
    return 0.5 * float(mods.min()), 2.0 * float(mods.max())


def chain_cell(cell, z, eps0, t):
    """A 1-D chain (on-site eps0, hopping t) cut into cells of `cell` sites."""
    h00 = eps0 * np.eye(cell) + t * (np.eye(cell, k=1) + np.eye(cell, k=-1))
    h01 = np.zeros((cell, cell))
    h01[cell - 1, 0] = t
    a_ii = (z * np.eye(cell) - h00)[None]
    a_ij = (-h01 + 0j)[None]
    a_ji = (-h01.T + 0j)[None]
    return a_ii, a_ij, a_ji


@pytest.mark.parametrize("b,r", [(4, 1), (8, 2), (16, 3), (24, 4)])
def test_spectral_equals_converged_decimation(b, r):
    a_ii, a_ij, a_ji = low_rank_lead(b, r, [-0.5, 0.1, 0.7], 1e-6, seed=b)
    r_i, r_o = contour_radii(a_ii, a_ij, a_ji)
    x, mask, ws, _ = SP.spectral_x_ii(a_ii, a_ij, a_ji, r_o=r_o, r_i=r_i, m_0=b, num_quad_points=64,
                                      max_decay=1.5 * np.log(r_o))
    ref, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 400, 1e-13)
    err = np.linalg.norm(x - ref, axis=(1, 2)) / np.linalg.norm(ref, axis=(1, 2))
    assert err.max() <= 1e-9, err
    assert mask.sum(axis=1).min() >= 1


@pytest.mark.parametrize("z", [0.3 + 1e-4j, -0.9 + 1e-4j, 1.6 + 1e-3j])
def test_spectral_chain_closed_form(z):
    """Cells of four chain sites: the surface site of the last cell carries the
    closed-form surface Green's function of the 1-D chain."""
    t, eps0 = 0.7, 0.1
    a_ii, a_ij, a_ji = chain_cell(4, z, eps0, t)
    x, _, _, _ = SP.spectral_x_ii(a_ii, a_ij, a_ji, m_0=4, num_quad_points=32)
    ref = O.chain_surface_gf_closed_form(z, eps0, t)
    assert abs(x[0, 3, 3] - ref) <= 1e-9 * abs(ref), (x[0, 3, 3], ref)


def test_spectral_brute_force_chain():
    """Against the last diagonal block of the inverse of a long finite lead."""
    a_ii, a_ij, a_ji = low_rank_lead(6, 2, [0.2, -0.4], 0.05, seed=3)
    x, _, _, _ = SP.spectral_x_ii(a_ii, a_ij, a_ji, m_0=6, num_quad_points=24)
    for e in range(2):
        ref = O.chain_surface_gf(a_ii[e], a_ij[e], a_ji[e], 300)
        assert np.linalg.norm(x[e] - ref) <= 1e-8 * np.linalg.norm(ref)


def test_beyn_eigenvalues_match_companion_linearisation():
    """Every eigenvalue of the linearised pencil inside the annulus is found
    to 1e-9, and every Beyn eigenpair that solves the quadratic problem is one
    of them.  Pairs that do not solve it (rounding noise above the rank cut
    when the moments are large) are allowed, as in the reference's own NEVP
    test, which filters such outliers; the reflected-mode mask removes them
    by their residual."""
    a_ii, a_ij, a_ji = low_rank_lead(8, 2, [0.15], 1e-3, seed=9)
    r_i, r_o = contour_radii(a_ii, a_ij, a_ji)
    ws, vs, nv = SP.beyn((a_ji, a_ii, a_ij), r_o, r_i, 8, 64)
    check_beyn_pairs(ws[0, : nv[0]], vs[0, :, : nv[0]], a_ji[0], a_ii[0], a_ij[0], r_i, r_o)


def check_beyn_pairs(ws, vs, a_ji, a_ii, a_ij, r_i, r_o, tol=1e-9):
    comp = SP.companion_eigenvalues(a_ji, a_ii, a_ij)
    inside = comp[(np.abs(comp) > r_i) & (np.abs(comp) < r_o)]
    scale = np.abs(inside).max()
    for w in inside:
        assert np.abs(ws - w).min() <= tol * scale, (w, ws)
    good = 0
    for k in range(len(ws)):
        w, v = ws[k], vs[:, k]
        res = np.linalg.norm((a_ji / w + a_ii + a_ij * w) @ v) / np.linalg.norm(v)
        if res <= 1e-6:
            good += 1
            assert np.abs(inside - w).min() <= tol * scale
    assert good == len(inside)


def test_reflected_mask_rejects_padding_and_growing_modes():
    """Padded columns (eigenvalue 0) are never reflected, and for a lead at an
    energy outside its band exactly the decaying solution (|w| > 1) is kept
    (per cell |w| = exp(4 acosh((E - eps0) / 2t)) = 4.5 and its inverse, both
    inside the annulus)."""
    a_ii, a_ij, a_ji = chain_cell(4, 1.6 + 1e-3j, 0.1, 0.7)
    ws, vs, nv = SP.beyn((a_ji, a_ii, a_ij), 10.0, 0.1, 4, 48)
    with np.errstate(all="ignore"):
        vs = vs / np.linalg.norm(vs, axis=-2, keepdims=True)
    mask = SP.reflected_modes(ws, vs, (a_ji, a_ii, a_ij))
    assert not mask[0, nv[0]:].any()
    kept = ws[0][mask[0]]
    assert nv[0] == 2
    assert len(kept) == 1 and abs(kept[0]) > 1


def test_swapped_couplings_fail_the_check():
    """A plausible mistake (a_ij and a_ji swapped) gives the other lead's
    surface and fails the decimation comparison."""
    a_ii, a_ij, a_ji = low_rank_lead(8, 2, [0.3], 1e-6, seed=21)
    x, _, _, _ = SP.spectral_x_ii(a_ii, a_ji, a_ij, m_0=8, num_quad_points=20)
    ref, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 400, 1e-13)
    assert np.linalg.norm(x - ref) > 1e-3 * np.linalg.norm(ref)
