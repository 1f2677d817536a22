"""The system-matrix assembly oracle against the synthetic generator, its own
two potential branches and the Hermitian / anti-Hermitian split."""

import numpy as np

from oracle import assembly as AS
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry


def _herm(rng, n, scale=1.0):
    x = scale * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return 0.5 * (x + x.conj().T)


def test_plain_assembly_is_the_generator():
    """No overlap, potential or self-energy: (E + i eta) I - H, as the
    synthetic generator builds it (boundary term not folded)."""
    g = Geometry((6, 4, 5))
    model = W.make_device(g, 123)
    en = np.array([-0.3, 0.25])
    ref = g.to_dense(*W.system_matrix(model, en, 1e-3, fold_boundary=False))
    h = g.to_dense(model.h_diag[None], model.h_upper[None], model.h_lower[None])[0]
    got = AS.assemble(g.block_sizes, en, 1e-3, h)
    np.testing.assert_array_equal(got, ref)


def test_identity_overlap_equals_no_overlap():
    """S = I: the overlap branch (S_ab (v_a + v_b) / 2 on every block) and the
    diagonal branch (diag(v)) give the same matrix."""
    rng = np.random.default_rng(4)
    bs = [3, 5, 4]
    n = sum(bs)
    h = _herm(rng, n)
    v = rng.standard_normal(n)
    sl = 1j * _herm(rng, n)
    sg = 1j * _herm(rng, n)
    sr = _herm(rng, n)
    en = np.array([0.1, -0.7, 0.4])
    a = AS.assemble(bs, en, 1e-2, h, None, v, sr, sl, sg)
    b = AS.assemble(bs, en, 1e-2, h, np.eye(n), v, sr, sl, sg)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_hermitian_split_for_real_energies():
    """eta = 0: (M + M^H)/2 = E S - P - H - Sigma^R_herm and
    (M - M^H)/2 = -(Sigma^> - Sigma^<)/2, with P = (S v + v S)/2 on the
    tridiagonal pattern.  A wrong sign or a dropped term fails one of them."""
    rng = np.random.default_rng(5)
    bs = [4, 4, 3, 5]
    n = sum(bs)
    mask = AS.tridiagonal_mask(bs)
    s = np.where(mask, _herm(rng, n, 0.1) + np.eye(n), 0)
    h = np.where(mask, _herm(rng, n), 0)
    v = rng.standard_normal(n)
    sr = np.where(mask, _herm(rng, n, 0.3), 0)
    sl = np.where(mask, 1j * _herm(rng, n, 0.2), 0)
    sg = np.where(mask, 1j * _herm(rng, n, 0.2), 0)
    en = np.array([0.3, -0.45])
    m = AS.assemble(bs, en, 0.0, h, s, v, sr, sl, sg)
    p = (s * v[:, None] + s * v[None, :]) / 2
    for e, E in enumerate(en):
        herm = 0.5 * (m[e] + m[e].conj().T)
        anti = 0.5 * (m[e] - m[e].conj().T)
        np.testing.assert_allclose(herm, E * s - p - h - sr, rtol=0, atol=1e-13)
        np.testing.assert_allclose(anti, -(sg - sl) / 2, rtol=0, atol=1e-13)
