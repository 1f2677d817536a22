"""GPU parity of the spectral open-boundary path (Beyn's contour integral, the
reflected-mode mask, the surface Green's function) against the oracle, the
converged Sancho-Rubio decimation, the closed form of a 1-D chain, the QZ
eigenvalues of the linearised pencil, and the Landauer-Caroli current of a
ballistic device whose contacts come from the spectral solver.

Bars: surface Green's functions within 1e-8 relative Frobenius error per
energy (the decimation and the spectral construction converge to the same
matrix from different algorithms; the spectral one carries the rounding of
quadrature sums over circles up to 100 times the mode moduli); Beyn
eigenvalues within 1e-9 of the QZ ones; layer currents within 1e-8 of
-(f_L - f_R) T.  The annulus of each case holds every finite non-zero
eigenvalue with a factor 2 to spare (contour_radii).
"""

import numpy as np
import pytest

from oracle import obc as O
from oracle import spectral as SP
from quatrex_amd.boundary_conditions import OBCSystem, Spectral, compute_obc
from quatrex_amd.nevp import Beyn
from tests.test_spectral_oracle import chain_cell, check_beyn_pairs, contour_radii, low_rank_lead

pytestmark = pytest.mark.gpu

TOL = 1e-8


def _t(x, device):
    import torch

    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.complex128, device=device)


def _rel(x, ref):
    x = np.asarray(x)
    num = np.linalg.norm((x - ref).reshape(len(ref), -1), axis=1)
    return float(np.max(num / np.linalg.norm(ref.reshape(len(ref), -1), axis=1)))


@pytest.mark.parametrize("b,r,m0,E", [(4, 1, 4, 3), (8, 2, 8, 5), (24, 4, 16, 4), (64, 8, 32, 6),
                                      (256, 12, 32, 3)])
def test_spectral_equals_decimation_and_oracle(device, b, r, m0, E):
    energies = np.linspace(-0.6, 0.7, E)
    a_ii, a_ij, a_ji = low_rank_lead(b, r, energies, 1e-6, seed=b)
    r_i, r_o = contour_radii(a_ii, a_ij, a_ji)
    print(f"spectral b={b} r={r} m0={m0} E={E} r_i={r_i:.3g} r_o={r_o:.3g}", flush=True)
    kw = dict(max_decay=1.5 * np.log(r_o))
    sp = Spectral(Beyn(r_o=r_o, r_i=r_i, m_0=m0, num_quad_points=64), **kw)
    x = sp(_t(a_ii, device), _t(a_ij, device), _t(a_ji, device), contact="left").cpu().numpy()
    ref, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 400, 1e-13)
    assert _rel(x, ref) <= TOL
    xo, mask, _, _ = SP.spectral_x_ii(a_ii, a_ij, a_ji, r_o=r_o, r_i=r_i, m_0=m0, num_quad_points=64, **kw)
    assert _rel(x, xo) <= TOL
    np.testing.assert_array_equal(sp.n_reflected.cpu().numpy(), mask.sum(axis=1))


def test_spectral_numpy_in_numpy_out(device):
    a_ii, a_ij, a_ji = low_rank_lead(8, 2, [0.1, 0.4], 1e-6, seed=7)
    sp = Spectral(Beyn(10.0, 0.8, 8, 20))
    x = sp(a_ii, a_ij, a_ji, contact="left")
    assert isinstance(x, np.ndarray) and x.shape == a_ii.shape
    x2 = sp(a_ii[0], a_ij[0], a_ji[0], contact="left")
    assert x2.shape == (8, 8)


@pytest.mark.parametrize("z", [0.3 + 1e-4j, -0.9 + 1e-4j, 1.6 + 1e-3j])
def test_spectral_chain_closed_form(device, z):
    t, eps0 = 0.7, 0.1
    a_ii, a_ij, a_ji = chain_cell(4, z, eps0, t)
    x = Spectral(Beyn(10.0, 0.1, 4, 32))(a_ii, a_ij, a_ji, contact="left")
    ref = O.chain_surface_gf_closed_form(z, eps0, t)
    assert abs(x[0, 3, 3] - ref) <= TOL * abs(ref), (x[0, 3, 3], ref)


def test_spectral_brute_force_finite_lead(device):
    a_ii, a_ij, a_ji = low_rank_lead(6, 2, [0.2, -0.4], 0.05, seed=3)
    r_i, r_o = contour_radii(a_ii, a_ij, a_ji)
    x = Spectral(Beyn(r_o, r_i, 6, 64), max_decay=1.5 * np.log(r_o))(a_ii, a_ij, a_ji, contact="left")
    for e in range(2):
        ref = O.chain_surface_gf(a_ii[e], a_ij[e], a_ji[e], 300)
        assert np.linalg.norm(x[e] - ref) <= 1e-8 * np.linalg.norm(ref)


@pytest.mark.parametrize("b,r", [(8, 2), (32, 5)])
def test_beyn_eigenpairs_vs_qz(device, b, r):
    """Device Beyn: every QZ eigenvalue of the linearised pencil inside the
    annulus is found, every eigenpair that solves the quadratic problem is one
    of them, eigenvectors have unit norm and columns beyond the rank carry
    zeros."""
    a_ii, a_ij, a_ji = low_rank_lead(b, r, [0.15, -0.35], 1e-3, seed=9 + b)
    r_i, r_o = contour_radii(a_ii, a_ij, a_ji)
    ws, vs = Beyn(r_o, r_i, b, 64)((a_ji, a_ii, a_ij))
    for e in range(2):
        valid = np.abs(ws[e]) > 0
        check_beyn_pairs(ws[e][valid], vs[e][:, valid], a_ji[e], a_ii[e], a_ij[e], r_i, r_o)
        for k in np.flatnonzero(valid):
            assert abs(np.linalg.norm(vs[e, :, k]) - 1.0) <= 1e-12
        assert np.all(vs[e][:, ~valid] == 0)


def test_spectral_through_obc_system_memoisation(device):
    """OBCSystem around the spectral solver: the first call solves, the next
    (memoised, a slightly perturbed system) refines the stored solution by
    fixed-point steps without calling the solver; both agree with decimation
    on their own inputs."""

    class Counting(Spectral):
        calls = 0

        def __call__(self, *args, **kwargs):
            Counting.calls += 1
            return super().__call__(*args, **kwargs)

    a_ii, a_ij, a_ji = low_rank_lead(16, 3, [0.1, -0.2, 0.5], 1e-3, seed=44)
    r_i, r_o = contour_radii(a_ii, a_ij, a_ji)
    osys = OBCSystem(Counting(Beyn(r_o, r_i, 16, 64), max_decay=1.5 * np.log(r_o)),
                     memoization_mode="force-after-first")
    x1 = osys((_t(a_ii, device), _t(a_ij, device), _t(a_ji, device)), contact="left")[0]
    ref1, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 400, 1e-13)
    assert _rel(x1.cpu().numpy(), ref1) <= TOL
    s = 1 + 1e-7
    x2 = osys((_t(a_ii * s, device), _t(a_ij * s, device), _t(a_ji * s, device)), contact="left")[0]
    ref2, _, _ = O.sancho_rubio(a_ii * s, a_ij * s, a_ji * s, 400, 1e-13)
    assert _rel(x2.cpu().numpy(), ref2) <= 1e-4
    assert Counting.calls == 1


def low_rank_periodic_device(b, r, N, energies, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b))) / np.sqrt(2 * b)
    h00 = (x + x.conj().T) / 2
    h01 = np.zeros((b, b), dtype=np.complex128)
    h01[b - r:, :r] = 0.6 * (rng.standard_normal((r, r)) + 1j * rng.standard_normal((r, r))) / np.sqrt(2 * r)
    n = b * N
    H = np.zeros((n, n), np.complex128)
    for k in range(N):
        H[k * b:(k + 1) * b, k * b:(k + 1) * b] = h00
        if k + 1 < N:
            H[k * b:(k + 1) * b, (k + 1) * b:(k + 2) * b] = h01
            H[(k + 1) * b:(k + 2) * b, k * b:(k + 1) * b] = h01.conj().T
    A = np.stack([e * np.eye(n) - H for e in energies])
    return A, np.arange(N + 1) * b


def test_ballistic_current_with_spectral_contacts(device):
    """compute_obc with the spectral solver on a periodic device, then the
    drop-in RGF: the current through every layer is -(f_L - f_R) T with T the
    (integer) Caroli transmission, and the boundary blocks agree with the
    Sancho-Rubio ones."""
    import torch

    from oracle.dense import place_obc
    from quatrex_amd.boundary_conditions import SanchoRubio
    from quatrex_amd.datastructures import BlockTridiagonal
    from quatrex_amd.greens_function_solver import RGF

    b, r, N = 16, 3, 5
    energies = np.array([0.05, -0.2, 0.3, 0.6])
    E = len(energies)
    A, off = low_rank_periodic_device(b, r, N, energies, seed=61)
    sizes = [b] * N
    a_dev = BlockTridiagonal(sizes, (E,), device=device)
    for i in range(N):
        a_dev.blocks[i, i] = _t(A[:, off[i]:off[i + 1], off[i]:off[i + 1]], device)
        if i + 1 < N:
            a_dev.blocks[i, i + 1] = _t(A[:, off[i]:off[i + 1], off[i + 1]:off[i + 2]], device)
            a_dev.blocks[i + 1, i] = _t(A[:, off[i + 1]:off[i + 2], off[i]:off[i + 1]], device)
    fL = np.array([1.0, 0.9, 0.7, 0.6])
    fR = np.array([0.0, 0.05, 0.3, 0.55])
    blk = lambda i, j: A[:, off[i]:off[i + 1], off[j]:off[j + 1]]  # noqa: E731
    r_i, r_o = contour_radii(blk(0, 0), blk(0, 1), blk(1, 0))
    spec = OBCSystem(Spectral(Beyn(r_o, r_i, 16, 64), max_decay=1.5 * np.log(r_o)),
                     memoization_mode="off")
    obc = compute_obc(a_dev, fL, fR, 1e-8, spec)
    sr = OBCSystem(SanchoRubio(200, 1e-12), memoization_mode="off")
    obc_sr = compute_obc(a_dev, fL, fR, 1e-8, sr)
    # the retarded boundary self-energies (the lesser / greater ones follow
    # from them by the contact kernel; some are zero where f = 0): the two
    # solvers reach the same surface at eta = 1e-8 up to the decimation's
    # stopping accuracy
    for x, y in zip(obc.retarded, obc_sr.retarded):
        if x is None:
            assert y is None
            continue
        assert _rel(x.cpu().numpy(), y.cpu().numpy()) <= 1e-5
    zero = BlockTridiagonal(sizes, (E,), device=device)
    _, _, _, cur = RGF(max_batch_size=100).selected_solve(a_dev, zero, zero, obc_blocks=obc,
                                                          return_retarded=True, return_current=True)
    cur = cur.cpu().numpy() if isinstance(cur, torch.Tensor) else np.asarray(cur)
    obc_np = tuple([None if t is None else t.cpu().numpy() for t in lst]
                   for lst in (obc.retarded, obc.lesser, obc.greater))
    G = np.linalg.inv(place_obc(A, off, obc_np[0], -1.0))
    gl = 1j * (obc_np[0][0] - obc_np[0][0].conj().swapaxes(-2, -1))
    gr = 1j * (obc_np[0][-1] - obc_np[0][-1].conj().swapaxes(-2, -1))
    g0n = G[:, :b, -b:]
    T = np.trace(gl @ g0n @ gr @ g0n.conj().swapaxes(-2, -1), axis1=-2, axis2=-1).real
    print(f"spectral contacts: transmission {T}")
    np.testing.assert_allclose(T, np.round(T), atol=1e-6)
    assert np.round(T).max() >= 1
    for j in range(N + 1):
        np.testing.assert_allclose(cur[:, j], -(fL - fR) * T, rtol=0, atol=1e-8)


def test_spectral_rejects_what_is_not_built(device):
    a_ii, a_ij, a_ji = low_rank_lead(4, 1, [0.1], 1e-6, seed=1)
    with pytest.raises(NotImplementedError):
        Spectral(Beyn(10.0, 0.8, 4, 8), block_sections=2)(a_ii, a_ij, a_ji, contact="left")
    with pytest.raises(NotImplementedError):
        Spectral(Beyn(10.0, 0.8, 4, 8))(a_ii, a_ij, a_ji, contact="left", return_injected=True)
    with pytest.raises(ValueError):
        Spectral(Beyn(10.0, 0.8, 4, 8), num_ref_iterations=0)
    from quatrex_amd import _lib

    with pytest.raises(_lib.RgfError, match="m_0"):
        Spectral(Beyn(10.0, 0.8, 65, 8))(a_ii, a_ij, a_ji, contact="left")
