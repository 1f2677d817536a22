"""The open-boundary oracle against closed forms, brute force and invariants."""

import warnings

import numpy as np
import pytest

from oracle import obc as O


def _lead(b, E, eta, seed, coupling=0.3):
    """Boundary blocks of a periodic lead: a_ii = (E + i eta) I - H00,
    a_ij = -H01, a_ji = -H01^H, stacked over energies."""
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b))) / np.sqrt(2 * b)
    h00 = (x + x.conj().T) / 2
    h01 = coupling * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b))) / np.sqrt(2 * b)
    energies = rng.uniform(-1, 1, E)
    eye = np.eye(b)
    a_ii = np.stack([(e + 1j * eta) * eye - h00 for e in energies])
    a_ij = np.broadcast_to(-h01, (E, b, b)).copy()
    a_ji = np.broadcast_to(-h01.conj().T, (E, b, b)).copy()
    return a_ii, a_ij, a_ji


def test_chain_closed_form():
    """1-D chain: the decimation result is the analytic surface Green's function."""
    t, eps0 = 0.7, 0.1
    for z in (0.3 + 1e-3j, -1.1 + 1e-3j, 2.5 + 1e-2j, 0.0 + 0.05j):
        a_ii = np.array([[[z - eps0]]])
        a_ij = np.array([[[-t]]], dtype=np.complex128)
        x, it, delta = O.sancho_rubio(a_ii, a_ij, a_ij.copy(), 100, 1e-12)
        ref = O.chain_surface_gf_closed_form(z, eps0, t)
        assert abs(x[0, 0, 0] - ref) <= 1e-10 * abs(ref), (z, x[0, 0, 0], ref)
        assert delta < 1e-12 and it >= 1


def test_closed_form_is_the_decaying_root():
    """The closed form solves the recursion and its imaginary part is negative."""
    t, eps0 = 0.5, -0.2
    for z in (0.1 + 1e-3j, 1.5 + 1e-3j):
        x = O.chain_surface_gf_closed_form(z, eps0, t)
        assert abs(x - 1 / (z - eps0 - t * t * x)) < 1e-12
        assert x.imag < 0


def test_brute_force_long_chain():
    """The surface Green's function is the last diagonal block of the inverse
    of a long finite chain (damping eta makes the truncation error ~e^-20)."""
    a_ii, a_ij, a_ji = _lead(4, 3, 0.05, seed=11)
    x, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 100, 1e-12)
    for e in range(3):
        ref = O.chain_surface_gf(a_ii[e], a_ij[e], a_ji[e], 400)
        assert np.linalg.norm(x[e] - ref) <= 1e-8 * np.linalg.norm(ref)


def test_brute_force_catches_swapped_couplings():
    """Swapping a_ij and a_ji gives the other lead's surface: a different
    matrix, so the brute-force check distinguishes the two."""
    a_ii, a_ij, a_ji = _lead(4, 1, 0.05, seed=12)
    x_swapped, _, _ = O.sancho_rubio(a_ii, a_ji, a_ij, 100, 1e-12)
    ref = O.chain_surface_gf(a_ii[0], a_ij[0], a_ji[0], 400)
    assert np.linalg.norm(x_swapped[0] - ref) > 1e-3 * np.linalg.norm(ref)


def test_fixed_point_identity_at_convergence():
    a_ii, a_ij, a_ji = _lead(8, 4, 1e-2, seed=13)
    x, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 100, 1e-10)
    rel, abs_res, x_ref = O.residuals(a_ii, a_ij, a_ji, x)
    assert rel.max() < 1e-9
    assert abs_res.shape == (4,)
    np.testing.assert_allclose(x_ref, O.fixed_point_step(a_ii, a_ij, a_ji, x))


def test_stopping_rule():
    a_ii, a_ij, a_ji = _lead(4, 2, 1e-2, seed=14)
    zero = np.zeros_like(a_ij)
    # no coupling: alpha = beta = 0 after the first step -> one iteration, inv(a_ii)
    x, it, delta = O.sancho_rubio(a_ii, zero, zero, 100, 1e-6)
    assert it == 1 and delta == 0.0
    np.testing.assert_allclose(x, np.linalg.inv(a_ii), rtol=1e-14, atol=0)
    # no iteration: warning, inv(a_ii)
    with pytest.warns(RuntimeWarning, match="did not converge"):
        x, it, delta = O.sancho_rubio(a_ii, a_ij, a_ji, 0, 1e-6)
    assert it == 0 and np.isnan(delta)
    np.testing.assert_allclose(x, np.linalg.inv(a_ii), rtol=1e-14, atol=0)
    # too few iterations: warning
    with pytest.warns(RuntimeWarning, match="did not converge"):
        _, it, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 2, 1e-6)
    assert it == 2


def test_stack_stops_together():
    """The stopping rule takes the max over the stack: an energy that alone
    converges early runs as many steps as the slowest one."""
    fast = _lead(4, 1, 0.3, seed=15)
    slow = _lead(4, 1, 1e-4, seed=15)
    both = tuple(np.concatenate((f, s)) for f, s in zip(fast, slow))
    _, it_fast, _ = O.sancho_rubio(*fast, 100, 1e-8)
    _, it_slow, _ = O.sancho_rubio(*slow, 100, 1e-8)
    _, it_both, _ = O.sancho_rubio(*both, 100, 1e-8)
    assert it_fast < it_slow == it_both


def _obc_system_callable(tol=1e-10):
    def call(system, contact):
        x, _, _ = O.sancho_rubio(*system, 100, tol)
        return x, None, None
    return call


def test_right_contact_reversal_equivalence():
    """The reference evaluates the right contact on block-reversed blocks and
    reverses the result back; the same value comes from the unreversed blocks
    with the roles a_ij = A[n,n-1], a_ji = A[n-1,n]."""
    a_ii, a_ij, a_ji = _lead(6, 3, 1e-2, seed=16)
    occ = np.array([0.1, 0.5, 0.9])
    # right contact of a device whose last blocks are A[n,n] = a_ii,
    # A[n,n-1] = a_ij (coupling toward the lead), A[n-1,n] = a_ji
    m_10, m_00, m_01 = a_ji, a_ii, a_ij
    ref = O.contact_obc(m_10, m_00, m_01, occ, 1e-3, _obc_system_callable(), contact="right")
    direct = O.contact_obc(m_10, m_00, m_01, occ, 1e-3, _obc_system_callable(), contact="left")
    for r, d in zip(ref, direct):
        np.testing.assert_allclose(d, r, rtol=1e-10, atol=1e-12 * np.abs(r).max())


def test_gamma_is_hermitian_psd_and_lesser_greater_relations():
    a_ii, a_ij, a_ji = _lead(6, 4, 1e-2, seed=17)
    occ = np.array([0.0, 0.3, 0.7, 1.0])
    sigma, lesser, greater = O.contact_obc(a_ji, a_ii, a_ij, occ, 1e-3, _obc_system_callable())
    gamma = 1j * (sigma - sigma.conj().swapaxes(-2, -1))
    for e in range(4):
        w = np.linalg.eigvalsh(gamma[e])
        assert w.min() > -1e-10 * max(1.0, w.max())
        np.testing.assert_allclose(lesser[e], 1j * occ[e] * gamma[e], rtol=0, atol=1e-14)
        np.testing.assert_allclose(greater[e], 1j * (occ[e] - 1) * gamma[e], rtol=0, atol=1e-14)
    # f = 0: no lesser term; f = 1: no greater term
    assert np.abs(lesser[0]).max() == 0 and np.abs(greater[3]).max() == 0


def test_obc_system_memoisation_modes():
    a_ii, a_ij, a_ji = _lead(4, 3, 1e-2, seed=18)
    system = (a_ii, a_ij, a_ji)
    # off: always solves
    s = O.OracleOBCSystem(memoization_mode="off", convergence_tol=1e-10)
    s(system, "c")
    s(system, "c")
    assert s.events == ["solve", "solve"]
    # auto: the first call solves; a second call finds the cached solution
    # converged for every energy (count 3 > 0.999) and solves again
    s = O.OracleOBCSystem(memoization_mode="auto", convergence_tol=1e-10)
    s(system, "c")
    s(system, "c")
    assert s.events == ["solve", "solve"]
    # auto with a cache that converges nowhere keeps the refined cache value
    s._cache["c"] = np.zeros_like(a_ii) + 10.0
    s.events.clear()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        s(system, "c")
    assert s.events == ["memo"]
    # force: starts from inv(a_ii) without solving
    s = O.OracleOBCSystem(memoization_mode="force", num_ref_iterations=4, convergence_tol=1e-10)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        x, rel, _ = s(system, "c")
    assert s.events == ["memo"]
    # force-after-first: solves first, reuses afterwards
    s = O.OracleOBCSystem(memoization_mode="force-after-first", convergence_tol=1e-10)
    s(system, "c")
    s(system, "c")
    assert s.events == ["solve", "memo"]


def periodic_device(b, N, energies, seed, coupling=0.6):
    """A ballistic device made of N copies of one lead cell: A = E I - H
    (no damping inside), returned as a dense stack and its block offsets."""
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b))) / np.sqrt(2 * b)
    h00 = (x + x.conj().T) / 2
    h01 = coupling * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b))) / np.sqrt(2 * b)
    n = b * N
    H = np.zeros((n, n), np.complex128)
    for k in range(N):
        H[k * b:(k + 1) * b, k * b:(k + 1) * b] = h00
        if k + 1 < N:
            H[k * b:(k + 1) * b, (k + 1) * b:(k + 2) * b] = h01
            H[(k + 1) * b:(k + 2) * b, k * b:(k + 1) * b] = h01.conj().T
    A = np.stack([e * np.eye(n) - H for e in energies])
    return A, np.arange(N + 1) * b


def test_ballistic_current_is_landauer():
    """Contact self-energies of a periodic device: the current through every
    layer is -(f_L - f_R) T with T = tr(Gamma_L G^R_0N Gamma_R G^A_N0) the
    Caroli transmission, an integer (the number of open channels)."""
    from oracle.dense import dense_current, place_obc

    b, N = 4, 5
    E = np.array([0.05, -0.2, 0.3])
    A, off = periodic_device(b, N, E, seed=3)

    def blk(M, i, j):
        return M[:, off[i]:off[i + 1], off[j]:off[j + 1]]

    call = _obc_system_callable(1e-12)
    fL, fR = np.array([1.0, 0.8, 0.6]), np.array([0.0, 0.1, 0.5])
    left = O.contact_obc(*O.boundary_blocks([blk(A, i, i) for i in range(N)],
                                            [blk(A, i, i + 1) for i in range(N - 1)],
                                            [blk(A, i + 1, i) for i in range(N - 1)], "left"),
                         fL, 1e-8, call, "left")
    right = O.contact_obc(*O.boundary_blocks([blk(A, i, i) for i in range(N)],
                                             [blk(A, i, i + 1) for i in range(N - 1)],
                                             [blk(A, i + 1, i) for i in range(N - 1)], "right"),
                          fR, 1e-8, call, "right")
    obc = tuple([l] + [None] * (N - 2) + [r] for l, r in zip(left, right))
    zero = np.zeros_like(A)
    cur = dense_current(A, zero, zero, off, obc)
    gl = 1j * (left[0] - left[0].conj().swapaxes(-2, -1))
    gr = 1j * (right[0] - right[0].conj().swapaxes(-2, -1))
    G = np.linalg.inv(place_obc(A, off, obc[0], -1.0))
    g0n = G[:, :b, -b:]
    T = np.trace(gl @ g0n @ gr @ g0n.conj().swapaxes(-2, -1), axis1=-2, axis2=-1)
    assert np.abs(T.imag).max() < 1e-10
    np.testing.assert_allclose(T.real, np.round(T.real), atol=1e-8)
    assert np.round(T.real).min() >= 1
    for j in range(N + 1):
        np.testing.assert_allclose(cur[:, j], -(fL - fR) * T, rtol=0, atol=1e-9)
