"""GPU parity of the open-boundary path (Sancho-Rubio, boundary system,
contact self-energies) against the oracle, and the ballistic current of a
periodic device against the Landauer-Caroli value.

Bars: surface Green's function and self-energies within 1e-10 relative
Frobenius error per energy of the oracle with the same iteration count; the
current through every layer within 1e-8 of -(f_L - f_R) T.
"""

import warnings

import numpy as np
import pytest

from oracle import obc as O
from oracle.dense import dense_current, place_obc
from quatrex_amd import solve as S
from quatrex_amd.boundary_conditions import OBCSystem, SanchoRubio, compute_contact_obc, compute_obc
from quatrex_amd.datastructures import BlockTridiagonal
from quatrex_amd.greens_function_solver import RGF, OBCBlocks
from tests.test_obc_oracle import _lead, periodic_device

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _t(x, device):
    import torch

    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.complex128, device=device)


def _rel(x, ref):
    x = np.asarray(x)
    num = np.linalg.norm((x - ref).reshape(len(ref), -1), axis=1)
    return float(np.max(num / np.linalg.norm(ref.reshape(len(ref), -1), axis=1)))


@pytest.mark.parametrize("b,E,eta", [(1, 3, 1e-3), (4, 5, 1e-2), (20, 6, 1e-3), (45, 4, 5e-3),
                                     (64, 8, 1e-3), (128, 3, 1e-4)])
def test_sancho_rubio_vs_oracle(device, b, E, eta):
    a_ii, a_ij, a_ji = _lead(b, E, eta, seed=100 + b)
    ref, it_ref, d_ref = O.sancho_rubio(a_ii, a_ij, a_ji, 100, 1e-8)
    sr = SanchoRubio(100, 1e-8)
    x = sr(_t(a_ii, device), _t(a_ij, device), _t(a_ji, device), "left").cpu().numpy()
    print(f"b={b} E={E} eta={eta}: {sr.iterations} iterations (oracle {it_ref}), "
          f"delta {sr.delta:.3e} (oracle {d_ref:.3e})")
    assert sr.iterations == it_ref
    assert sr.delta == pytest.approx(d_ref, rel=1e-6, abs=1e-14)
    assert _rel(x, ref) <= TOL


def test_sancho_rubio_numpy_in_numpy_out_and_stack_shape(device):
    a_ii, a_ij, a_ji = _lead(8, 6, 1e-2, seed=7)
    shp = (2, 3, 8, 8)
    x = SanchoRubio()(a_ii.reshape(shp), a_ij.reshape(shp), a_ji.reshape(shp), "c")
    assert isinstance(x, np.ndarray) and x.shape == shp
    ref, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji)
    assert _rel(x.reshape(6, 8, 8), ref) <= TOL


def test_sancho_rubio_stopping_rule(device):
    a_ii, a_ij, a_ji = _lead(16, 3, 1e-2, seed=8)
    zero = np.zeros_like(a_ij)
    sr = SanchoRubio()
    x = sr(_t(a_ii, device), _t(zero, device), _t(zero, device), "c").cpu().numpy()
    assert sr.iterations == 1 and sr.delta == 0.0
    assert _rel(x, np.linalg.inv(a_ii)) <= 1e-13
    with pytest.warns(RuntimeWarning, match="did not converge"):
        x = SanchoRubio(max_iterations=0)(_t(a_ii, device), _t(a_ij, device), _t(a_ji, device), "c")
    assert _rel(x.cpu().numpy(), np.linalg.inv(a_ii)) <= 1e-13
    with pytest.warns(RuntimeWarning, match="did not converge"):
        sr = SanchoRubio(max_iterations=2)
        sr(_t(a_ii, device), _t(a_ij, device), _t(a_ji, device), "c")
    assert sr.iterations == 2


def test_sancho_rubio_stack_stops_together(device):
    fast = _lead(32, 2, 0.3, seed=9)
    slow = _lead(32, 2, 1e-4, seed=9)
    both = [np.concatenate((f, s)) for f, s in zip(fast, slow)]
    ref, it_ref, _ = O.sancho_rubio(*both, 100, 1e-8)
    sr = SanchoRubio(100, 1e-8)
    x = sr(*(_t(v, device) for v in both), "c").cpu().numpy()
    assert sr.iterations == it_ref
    assert _rel(x, ref) <= TOL


def test_sancho_rubio_chunked_matches_whole(device):
    """With the scratch split into chunks of 3 energies, every energy still
    runs the batch-wide iteration count and the results agree."""
    a_ii, a_ij, a_ji = _lead(48, 8, 1e-3, seed=10)
    args = [_t(v, device) for v in (a_ii, a_ij, a_ji)]
    whole = SanchoRubio(100, 1e-8)
    xw = whole(*args, "c").cpu().numpy()
    ctx = S.default_context(device)
    ctx.set_option(1, 3)  # RGF_OPT_MAX_CHUNK
    try:
        part = SanchoRubio(100, 1e-8)
        xp = part(*args, "c").cpu().numpy()
    finally:
        ctx.set_option(1, 0)
    assert part.iterations == whole.iterations
    assert _rel(xp, xw) <= 1e-13


def test_fixed_point_and_residuals_vs_oracle(device):
    a_ii, a_ij, a_ji = _lead(40, 5, 1e-2, seed=11)
    x0, _, _ = O.sancho_rubio(a_ii, a_ij, a_ji, 3, 1e-30)  # deliberately unconverged
    rel_ref, abs_ref, xref = O.residuals(a_ii, a_ij, a_ji, x0)
    system = tuple(_t(v, device) for v in (a_ii, a_ij, a_ji))
    osys = OBCSystem(SanchoRubio())
    rel, ab, xr = osys._get_residuals(system, _t(x0, device))
    assert _rel(xr.cpu().numpy(), xref) <= TOL
    np.testing.assert_allclose(ab.cpu().numpy(), abs_ref, rtol=1e-9)
    np.testing.assert_allclose(rel.cpu().numpy(), rel_ref, rtol=1e-9)
    step = osys._fixed_point_step(system, _t(x0, device)).cpu().numpy()
    assert _rel(step, xref) <= TOL


class _Counting(SanchoRubio):
    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.calls = 0

    def __call__(self, *a, **k):
        self.calls += 1
        return super().__call__(*a, **k)


@pytest.mark.parametrize("mode", ["off", "auto", "force", "force-after-first"])
def test_obc_system_memoisation_matches_oracle(device, mode):
    a_ii, a_ij, a_ji = _lead(24, 4, 1e-2, seed=12)
    system_np = (a_ii, a_ij, a_ji)
    system = tuple(_t(v, device) for v in system_np)
    solver = _Counting(100, 1e-10)
    osys = OBCSystem(solver, memoization_mode=mode, num_ref_iterations=3)
    ref = O.OracleOBCSystem(convergence_tol=1e-10, memoization_mode=mode, num_ref_iterations=3)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        for _ in range(3):
            x, rel, ab = osys(system, "left")
            xr, relr, abr = ref(system_np, "left")
            assert _rel(x.cpu().numpy(), xr) <= TOL
            # converged residuals are rounding noise (~1e-13 ||x||), which two
            # implementations do not share
            scale = float(np.linalg.norm(xr, axis=(-2, -1)).max())
            np.testing.assert_allclose(ab.cpu().numpy(), abr, rtol=1e-6, atol=1e-11 * scale)
    assert solver.calls == ref.events.count("solve")


def test_boundary_system_with_overlap(device):
    """a_ii = m00 + i eta s00 (and the couplings) with overlap blocks."""
    rng = np.random.default_rng(13)
    E, b, eta = 3, 33, 0.02
    m = [rng.standard_normal((E, b, b)) + 1j * rng.standard_normal((E, b, b)) for _ in range(3)]
    s = [rng.standard_normal((E, b, b)) + 1j * rng.standard_normal((E, b, b)) for _ in range(3)]
    seen = {}

    def grab(system, contact):
        seen["system"] = [v.cpu().numpy() for v in system]
        seen["contact"] = contact
        return _t(np.linalg.inv(seen["system"][0]), device), None, None

    compute_contact_obc(_t(m[0], device), _t(m[1], device), _t(m[2], device), np.full(E, 0.5), eta,
                        grab, contact="left-x", overlap_blocks=[_t(v, device) for v in s])
    a_ii, a_ij, a_ji = seen["system"]
    np.testing.assert_allclose(a_ii, m[1] + 1j * eta * s[1], rtol=0, atol=1e-15)
    np.testing.assert_allclose(a_ij, m[2] + 1j * eta * s[2], rtol=0, atol=1e-15)
    np.testing.assert_allclose(a_ji, m[0] + 1j * eta * s[0], rtol=0, atol=1e-15)
    assert seen["contact"] == "G: left-x"
    # without overlap: i eta on the diagonal only
    compute_contact_obc(_t(m[0], device), _t(m[1], device), _t(m[2], device), np.full(E, 0.5), eta,
                        grab)
    np.testing.assert_array_equal(seen["system"][0], m[1] + 1j * eta * np.eye(b))
    np.testing.assert_array_equal(seen["system"][1], m[2])


@pytest.mark.parametrize("b", [8, 37, 96])
def test_contact_self_energies_vs_oracle(device, b):
    E = 4
    a_ii, a_ij, a_ji = _lead(b, E, 0.0, seed=14 + b)
    occ = np.array([0.0, 0.25, 0.8, 1.0])
    osys = OBCSystem(SanchoRubio(100, 1e-10), memoization_mode="off")
    got = compute_contact_obc(_t(a_ji, device), _t(a_ii, device), _t(a_ij, device), occ, 1e-3, osys)
    ref = O.contact_obc(a_ji, a_ii, a_ij, occ, 1e-3,
                        O.OracleOBCSystem(convergence_tol=1e-10, memoization_mode="off"))
    for g, r, name in zip(got, ref, ("sigma", "lesser", "greater")):
        g = g.cpu().numpy()
        nz = np.linalg.norm(r.reshape(E, -1), axis=1) > 0
        assert _rel(g[nz], r[nz]) <= TOL, name
        assert np.abs(g[~nz]).max(initial=0.0) == 0.0, name


@pytest.mark.parametrize("b,N", [(16, 6), (40, 5)])
def test_ballistic_device_current_is_landauer(device, b, N):
    """compute_obc on a periodic device, then the drop-in RGF with zero
    scattering self-energies: the current through every layer equals
    -(f_L - f_R) T with T the Caroli transmission (an integer here), and G^R
    matches the dense inverse with the same boundary blocks."""
    import torch

    energies = np.array([0.05, -0.2, 0.3, 0.6])
    E = len(energies)
    A, off = periodic_device(b, N, energies, seed=20 + b)
    sizes = [b] * N
    sp = [A[e] for e in range(E)]
    a_dev = BlockTridiagonal(sizes, (E,), device=device)
    for i in range(N):
        a_dev.blocks[i, i] = _t(A[:, off[i]:off[i + 1], off[i]:off[i + 1]], device)
        if i + 1 < N:
            a_dev.blocks[i, i + 1] = _t(A[:, off[i]:off[i + 1], off[i + 1]:off[i + 2]], device)
            a_dev.blocks[i + 1, i] = _t(A[:, off[i + 1]:off[i + 2], off[i]:off[i + 1]], device)
    del sp
    fL = np.array([1.0, 0.9, 0.7, 0.6])
    fR = np.array([0.0, 0.05, 0.3, 0.55])
    osys = OBCSystem(SanchoRubio(100, 1e-12), memoization_mode="off")
    obc = compute_obc(a_dev, fL, fR, 1e-8, osys)
    assert isinstance(obc, OBCBlocks)
    zero = BlockTridiagonal(sizes, (E,), device=device)
    xl, xg, xr, cur = RGF(max_batch_size=100).selected_solve(a_dev, zero, zero, obc_blocks=obc,
                                                            return_retarded=True,
                                                            return_current=True)
    cur = cur.cpu().numpy() if isinstance(cur, torch.Tensor) else np.asarray(cur)
    obc_np = tuple([None if t is None else t.cpu().numpy() for t in lst]
                   for lst in (obc.retarded, obc.lesser, obc.greater))
    G = np.linalg.inv(place_obc(A, off, obc_np[0], -1.0))
    gl = 1j * (obc_np[0][0] - obc_np[0][0].conj().swapaxes(-2, -1))
    gr = 1j * (obc_np[0][-1] - obc_np[0][-1].conj().swapaxes(-2, -1))
    g0n = G[:, :b, -b:]
    T = np.trace(gl @ g0n @ gr @ g0n.conj().swapaxes(-2, -1), axis1=-2, axis2=-1).real
    print(f"b={b} N={N}: transmission {T}")
    np.testing.assert_allclose(T, np.round(T), atol=1e-6)
    for j in range(N + 1):
        np.testing.assert_allclose(cur[:, j], -(fL - fR) * T, rtol=0, atol=1e-8)
    ref_cur = dense_current(A, np.zeros_like(A), np.zeros_like(A), off, obc_np)
    np.testing.assert_allclose(cur, ref_cur, rtol=0, atol=1e-8)
    gr_dev = xr.to_dense().cpu().numpy() if hasattr(xr, "to_dense") else None
    if gr_dev is not None:
        for i in range(N):
            s = slice(off[i], off[i + 1])
            assert _rel(gr_dev[:, s, s], G[:, s, s]) <= 1e-9


def test_compute_obc_with_overlap_matches_oracle(device):
    """Both contacts with an overlap matrix, including the reference's
    unreversed overlap on the right contact."""
    rng = np.random.default_rng(31)
    b, N, E = 12, 4, 3
    A, off = periodic_device(b, N, np.array([0.1, -0.3, 0.45]), seed=31)
    Sm = np.zeros_like(A)
    n = b * N
    for e in range(E):
        x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        Sm[e] = np.eye(n) + 0.05 * (x + x.conj().T)
    a_dev = BlockTridiagonal([b] * N, (E,), device=device)
    s_dev = BlockTridiagonal([b] * N, (E,), device=device)

    def blk(M, i, j):
        return M[:, off[i]:off[i + 1], off[j]:off[j + 1]]

    for i in range(N):
        for j in (i - 1, i, i + 1):
            if 0 <= j < N:
                a_dev.blocks[i, j] = _t(blk(A, i, j), device)
                s_dev.blocks[i, j] = _t(blk(Sm, i, j), device)
    fL, fR = np.array([0.9, 0.5, 0.2]), np.array([0.1, 0.4, 0.6])
    eta = 0.05
    osys = OBCSystem(SanchoRubio(100, 1e-10), memoization_mode="off")
    obc = compute_obc(a_dev, fL, fR, eta, osys, overlap=s_dev)
    call = lambda system, contact: (O.sancho_rubio(*system, 100, 1e-10)[0], None, None)  # noqa: E731
    ref_l = O.contact_obc(blk(A, 1, 0), blk(A, 0, 0), blk(A, 0, 1), fL, eta, call, "left",
                          overlap=(blk(Sm, 1, 0), blk(Sm, 0, 0), blk(Sm, 0, 1)))
    m = N - 2
    ref_r = O.contact_obc(blk(A, m, N - 1), blk(A, N - 1, N - 1), blk(A, N - 1, m), fR, eta, call,
                          "right", overlap=(blk(Sm, m, N - 1), blk(Sm, N - 1, N - 1), blk(Sm, N - 1, m)))
    for k, lst in enumerate((obc.retarded, obc.lesser, obc.greater)):
        assert _rel(lst[0].cpu().numpy(), ref_l[k]) <= TOL
        assert _rel(lst[-1].cpu().numpy(), ref_r[k]) <= TOL


def test_obc_entry_points_reject_bad_arguments(device):
    """Invalid sizes and arguments come back as status codes (RgfError) or the
    reference's exception types; nothing is launched for them."""
    import ctypes

    import torch

    from quatrex_amd import _lib

    lib = _lib.load()
    ctx = S.default_context(device)
    x = torch.zeros((2, 8, 8), dtype=torch.complex128, device=device)
    it, d = ctypes.c_int32(), ctypes.c_double()
    p = ctypes.c_void_p(x.data_ptr())
    # block size out of range, negative energies, negative iterations, NULL buffers
    assert lib.rgf_sancho_rubio_z(ctx.handle, 0, 2, p, p, p, 10, 1e-6, p, ctypes.byref(it),
                                  ctypes.byref(d)) == _lib.RGF_EUNSUPPORTED
    assert lib.rgf_sancho_rubio_z(ctx.handle, 8, -1, p, p, p, 10, 1e-6, p, ctypes.byref(it),
                                  ctypes.byref(d)) == _lib.RGF_EINVAL
    assert lib.rgf_sancho_rubio_z(ctx.handle, 8, 2, p, p, p, -1, 1e-6, p, ctypes.byref(it),
                                  ctypes.byref(d)) == _lib.RGF_EINVAL
    assert lib.rgf_sancho_rubio_z(ctx.handle, 8, 2, None, p, p, 10, 1e-6, p, ctypes.byref(it),
                                  ctypes.byref(d)) == _lib.RGF_EINVAL
    r = torch.zeros(2, dtype=torch.float64, device=device)
    # x_ref aliasing x; one residual array without the other
    assert lib.rgf_obc_fixed_point_z(ctx.handle, 8, 2, p, p, p, p, p, None, None) == _lib.RGF_EINVAL
    x2 = torch.zeros_like(x)
    assert lib.rgf_obc_fixed_point_z(ctx.handle, 8, 2, p, p, p, p, ctypes.c_void_p(x2.data_ptr()),
                                     ctypes.c_void_p(r.data_ptr()), None) == _lib.RGF_EINVAL
    assert lib.rgf_contact_obc_z(ctx.handle, 8, 2, p, p, p, None, p, p, p) == _lib.RGF_EINVAL
    assert lib.rgf_boundary_system_z(ctx.handle, 8, 2, 0.1, None, p, p, None, None, None, p, None,
                                     None) == _lib.RGF_EINVAL
    # zero energies: nothing to do
    assert lib.rgf_sancho_rubio_z(ctx.handle, 8, 0, p, p, p, 10, 1e-6, p, ctypes.byref(it),
                                  ctypes.byref(d)) == _lib.RGF_OK
    assert it.value == 0
    # Python layer: mismatched shapes, non-square blocks, wrong occupancy count
    y = torch.zeros((2, 6, 6), dtype=torch.complex128, device=device)
    with pytest.raises(ValueError):
        SanchoRubio()(x, y, x, "c")
    with pytest.raises(ValueError):
        SanchoRubio()(torch.zeros((2, 8, 6), dtype=torch.complex128, device=device), x, x, "c")
    with pytest.raises(ValueError):
        compute_contact_obc(x, x, x, np.zeros(3), 0.0, OBCSystem(SanchoRubio(), memoization_mode="off"))
    with pytest.raises(ValueError):
        OBCSystem(SanchoRubio(), agreement_threshold=1.5)
    with pytest.raises(NotImplementedError):
        OBCSystem(SanchoRubio(), system_reducer=object())
    with pytest.warns(RuntimeWarning, match="at least 2"):
        OBCSystem(SanchoRubio(), num_ref_iterations=1)


@pytest.mark.parametrize("layout_name", ["coo", "csr"])
def test_boundary_conditions_and_solve_on_block_sparse_stacks(device, layout_name):
    """The electron-solver step on the reference's storage: boundary
    self-energies of both contacts read from a block-sparse system matrix
    (compute_obc through .blocks), then the drop-in RGF on block-sparse
    stacks with those OBC blocks and the current, against the dense oracle
    with the same boundary blocks."""
    from oracle.dense import dense_solve
    from quatrex_amd.dsdb import DSDBCOO, DSDBCSR
    from tests.helpers import block_rel_errors, tridiagonal_mask
    from tests.test_gpu_dsdb import _container

    layout = DSDBCOO if layout_name == "coo" else DSDBCSR
    b, N = 10, 5
    energies = np.array([0.05, -0.25, 0.4])
    E = len(energies)
    A, off = periodic_device(b, N, energies, seed=41)
    A = A + 1j * 1e-3 * np.eye(b * N)
    bs = [b] * N
    zero = np.zeros_like(A)
    skew = lambda x: -np.conj(x)  # noqa: E731
    a = _container(layout, A, bs, False, np.conj, device)
    sl = _container(layout, zero, bs, True, skew, device)
    sg = _container(layout, zero, bs, True, skew, device)
    fL, fR = np.array([0.9, 0.6, 0.3]), np.array([0.1, 0.4, 0.2])
    obc = compute_obc(a, fL, fR, 1e-3, OBCSystem(SanchoRubio(100, 1e-10), memoization_mode="off"))
    xl, xg, xr, cur = RGF(max_batch_size=2).selected_solve(a, sl, sg, obc_blocks=obc,
                                                          return_retarded=True, return_current=True)
    obc_np = tuple([None if t is None else t.cpu().numpy() for t in lst]
                   for lst in (obc.retarded, obc.lesser, obc.greater))
    ref = dense_solve(A, zero, zero, off, obc_np)
    mask = tridiagonal_mask(bs)
    for X, R, name in zip((xr, xl, xg), ref, ("G^R", "G^<", "G^>")):
        err = block_rel_errors(X.to_dense().cpu().numpy(), R * mask, bs)
        assert err <= 1e-10, (name, err)
    cur = cur.cpu().numpy() if hasattr(cur, "cpu") else np.asarray(cur)
    ref_cur = dense_current(A, zero, zero, off, obc_np)
    np.testing.assert_allclose(cur, ref_cur, rtol=0, atol=1e-8 * np.abs(ref_cur).max())
