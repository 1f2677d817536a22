"""Block-sparse stacks on the GPU: the HIP densify / sparsify against the
oracle's block access, and the drop-in RGF on block-sparse containers
against the dense oracle."""

import numpy as np
import pytest
from scipy import sparse

from oracle.dense import dense_solve
from oracle.dsdb import OracleDSDB
from quatrex_amd import workloads as W
from quatrex_amd.dsdb import DSDBCOO, DSDBCSR
from quatrex_amd.greens_function_solver import RGF
from quatrex_amd.layout import Geometry
from tests.helpers import block_rel_errors, tridiagonal_mask

pytestmark = pytest.mark.gpu

SYMMETRY = [(False, np.conj), (True, np.conj), (True, lambda x: -np.conj(x))]
LAYOUTS = [DSDBCOO, DSDBCSR]


def _matrix(bs, seed, symmetry, op):
    n = int(sum(bs))
    rng = np.random.default_rng(seed)
    m = sparse.random(n, n, density=0.3, random_state=rng, format="coo").astype(np.complex128)
    m.data += 1j * rng.uniform(size=m.nnz)
    m.setdiag(rng.uniform(size=n) + 1j * rng.uniform(size=n))
    if symmetry:
        mt = m.copy()
        mt.data = op(mt.data)
        m = (m + mt.T).tocoo()
    return m


def _fill_per_energy(container, oracle, rng):
    """Different values per stack element, same pattern, in both copies."""
    import torch

    vals = rng.standard_normal(oracle.data.shape) + 1j * rng.standard_normal(oracle.data.shape)
    oracle.data[...] = vals
    container.data.copy_(torch.as_tensor(vals, device=container.data.device))


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("symmetry,op", SYMMETRY)
def test_device_blocks_match_oracle(device, layout, symmetry, op):
    bs = [5] * 3 + [10] * 2 + [5] * 3
    m = _matrix(bs, 41, symmetry, op)
    o = OracleDSDB.from_sparray(m, bs, (3,), symmetry, op)
    c = layout.from_sparray(m, bs, (3,), symmetry, op, device=device)
    _fill_per_energy(c, o, np.random.default_rng(2))
    for i in range(len(bs)):
        for j in range(len(bs)):
            got = c.blocks[i, j].cpu().numpy()
            np.testing.assert_array_equal(got, o.get_block(i, j), err_msg=f"block {(i, j)}")


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("symmetry,op", SYMMETRY)
def test_device_set_block_matches_oracle(device, layout, symmetry, op):
    bs = [2] * 3 + [4] * 2 + [2] * 3
    m = _matrix(bs, 42, symmetry, op)
    o = OracleDSDB.from_sparray(m, bs, (2,), symmetry, op)
    c = layout.from_sparray(m, bs, (2,), symmetry, op, device=device)
    rng = np.random.default_rng(3)
    for (i, j) in [(0, 0), (3, 4), (4, 3), (6, 6), (2, 5)]:
        val = rng.standard_normal((2, bs[i], bs[j])) + 1j * rng.standard_normal((2, bs[i], bs[j]))
        o.set_block(i, j, val)
        c.blocks[i, j] = val
    np.testing.assert_array_equal(c.data.cpu().numpy(), o.data)


def _system(geo, E, eta, seed):
    model = W.make_device(geo, seed)
    en = W.draw_energies(E)
    A = W.system_matrix(model, en, eta)
    sld, slu, sgd, sgu = W.sigmas(model, E)
    z = np.zeros_like(slu)
    dA = geo.to_dense(*A)
    dL = geo.to_dense(sld, slu, z)
    dG = geo.to_dense(sgd, sgu, z)
    return dA, dL, dG


def _container(layout, dense, bs, symmetry, op, device):
    """Pattern of energy 0 (tridiagonal blocks), values of every energy."""
    import torch

    E = dense.shape[0]
    mask = tridiagonal_mask(bs)
    src = np.where(mask, dense[0], 0)
    if symmetry:
        src = np.triu(src)
    src[mask & (src == 0)] = 1e-300  # keep the full tridiagonal pattern
    c = layout.from_sparray(sparse.coo_matrix(src), bs, (E,), symmetry, op, device=device)
    rows, cols, pos = c._coordinates()
    vals = np.zeros((E, c.nnz), np.complex128)
    vals[:, pos] = dense[:, rows, cols]
    c.data.copy_(torch.as_tensor(vals, device=device))
    return c


@pytest.mark.parametrize("layout", LAYOUTS)
def test_dropin_selected_solve_on_block_sparse_stacks(device, layout):
    bs = [8, 12, 8, 10, 6]
    geo = Geometry(tuple(bs))
    E = 3
    dA, dL, dG = _system(geo, E, 1e-2, 5151)
    skew = lambda x: -np.conj(x)  # noqa: E731
    a = _container(layout, dA, bs, False, np.conj, device)
    sl = _container(layout, dL, bs, True, skew, device)
    sg = _container(layout, dG, bs, True, skew, device)
    xl, xg, xr = RGF(max_batch_size=2).selected_solve(a, sl, sg, return_retarded=True)
    ref = dense_solve(dA, dL, dG)
    mask = tridiagonal_mask(bs)
    for X, R in zip((xr, xl, xg), (ref[0], ref[1], ref[2])):
        got = X.to_dense().cpu().numpy()
        err = block_rel_errors(got, R * mask, bs)
        assert err <= 1e-10, err


def test_dropin_selected_inv_with_host_block_sparse_stacks(device):
    """numpy-backed containers (the reference's objects on the host): the
    batch's raw entries are staged to the device once and densified there."""
    bs = [6, 10, 6, 8]
    geo = Geometry(tuple(bs))
    E = 4
    dA, _, _ = _system(geo, E, 1e-2, 5252)
    mask = tridiagonal_mask(bs)
    src = np.where(mask, dA[0], 0)
    o = OracleDSDB.from_sparray(sparse.coo_matrix(src), bs, (E,))
    o.data[...] = dA[:, o.rows, o.cols]
    out = OracleDSDB(np.zeros_like(o.data), o.rows, o.cols, bs)
    RGF(max_batch_size=3).selected_inv(o, out=out)
    ref = np.linalg.inv(dA)
    got = np.zeros_like(dA)
    for i in range(len(bs)):
        for j in range(max(0, i - 1), min(len(bs), i + 2)):
            off = np.concatenate(([0], np.cumsum(bs)))
            got[:, off[i]:off[i + 1], off[j]:off[j + 1]] = out.get_block(i, j)
    err = block_rel_errors(got, ref * mask, bs)
    assert err <= 1e-10, err


@pytest.mark.parametrize("layout", LAYOUTS)
def test_dropin_on_stack_views_as_the_electron_solver_calls_it(device, layout, monkeypatch):
    """ElectronSolver.solve's call shape (electron/solver.py:821-874): a system
    matrix holding one energy batch, Sigma^<> and every output passed as
    ``.stack[batch]`` views of whole-grid stacks, OBC blocks of the batch,
    return_retarded and return_current.  The views must take the device
    densify / sparsify path (the per-block host loops are disabled here) and
    give results bit for bit equal to one call on the whole stacks."""
    import torch

    from quatrex_amd.greens_function_solver import OBCBlocks
    from quatrex_amd.greens_function_solver import rgf as rgf_mod
    from quatrex_amd.utils.solvers_utils import get_batches

    bs = [8, 12, 8, 10, 6]
    geo = Geometry(tuple(bs))
    E = 6
    dA, dL, dG = _system(geo, E, 1e-2, 5353)
    skew = lambda x: -np.conj(x)  # noqa: E731
    sl = _container(layout, dL, bs, True, skew, device)
    sg = _container(layout, dG, bs, True, skew, device)
    rng = np.random.default_rng(9)

    def contact(b, f):
        br = -0.5j * np.eye(b) + 0.05 * (rng.standard_normal((b, b)) + 1j * rng.standard_normal((b, b)))
        gam = 1j * (br - br.conj().T)
        rep = lambda x: np.broadcast_to(x, (E, b, b)).copy()  # noqa: E731
        return rep(br), rep(1j * gam * f), rep(1j * gam * (f - 1))

    L, R = contact(bs[0], 0.7), contact(bs[-1], 0.2)

    def obc_for(s):
        o = OBCBlocks(len(bs))
        for k, lst in enumerate((o.retarded, o.lesser, o.greater)):
            lst[0], lst[-1] = L[k][s], R[k][s]
        return o

    def outputs():
        return tuple(_container(layout, np.zeros_like(dA), bs, sym, op, device)
                     for sym, op in ((True, skew), (True, skew), (False, np.conj)))

    # one call on the whole stacks
    whole = outputs()
    a_all = _container(layout, dA, bs, False, np.conj, device)
    cur_all = RGF(max_batch_size=100).selected_solve(a_all, sl, sg, obc_for(slice(None)), out=whole,
                                                     return_retarded=True, return_current=True)

    def no_host_loop(*_a, **_k):
        raise AssertionError("stack views went through the per-block host path")

    monkeypatch.setattr(rgf_mod, "_gather", no_host_loop)
    monkeypatch.setattr(rgf_mod, "_scatter", no_host_loop)
    # the electron solver's loop: batches of the energy grid, views everywhere
    batched = outputs()
    cur = []
    sizes, offs = get_batches(E, 3)  # [3, 3]; inside, batches of one: views of views
    for k in range(len(sizes)):
        s = slice(int(offs[k]), int(offs[k + 1]))
        a_batch = _container(layout, dA[s], bs, False, np.conj, device)
        out_views = tuple(x.stack[s] for x in batched)
        cur.append(RGF(max_batch_size=1).selected_solve(a_batch, sl.stack[s], sg.stack[s], obc_for(s),
                                                        out=out_views, return_retarded=True,
                                                        return_current=True))
    for w, b in zip(whole, batched):
        assert torch.equal(w.data, b.data)
    assert torch.equal(torch.cat(cur), cur_all)
    # and the values are right
    obc_np = [[L[k]] + [None] * (len(bs) - 2) + [R[k]] for k in range(3)]
    ref = dense_solve(dA, dL, dG, geo.block_offsets, obc_np)
    mask = tridiagonal_mask(bs)
    for X, Rf in zip((batched[2], batched[0], batched[1]), ref):
        err = block_rel_errors(X.to_dense().cpu().numpy(), Rf * mask, bs)
        assert err <= 1e-10, err
