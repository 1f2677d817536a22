"""Host-resident stacks through the drop-in ``RGF.selected_solve``.

A batch of CPU-resident block-tridiagonal stacks is split into energy chunks
whose host copies, transfers in both directions and solves overlap (pinned
staging, two copy streams, the context stream).  Energies are independent
and every kernel's arithmetic per energy is the same at any batch size, so
the results must equal, bit for bit, those of the same solve on
device-resident stacks.
"""

import numpy as np
import pytest

from quatrex_amd import workloads as W
from quatrex_amd.datastructures import BlockTridiagonal
from quatrex_amd.greens_function_solver import RGF, OBCBlocks
from quatrex_amd.greens_function_solver import rgf as rgf_mod
from quatrex_amd.layout import Geometry

pytestmark = pytest.mark.gpu


def _case(E, seed):
    import torch

    bs = (24, 40, 16, 33)
    g = Geometry(bs)
    model = W.make_device(g, seed)
    ad, au, al = W.system_matrix(model, W.draw_energies(E), 1e-2, fold_boundary=False)
    sld, slu, sgd, sgu = W.sigmas(model, E)
    rng = np.random.default_rng(seed)

    def one(b, f):
        br = -0.5j * np.eye(b) + 0.05 * (rng.standard_normal((E, b, b)) + 1j * rng.standard_normal((E, b, b)))
        gam = 1j * (br - np.conj(np.swapaxes(br, -1, -2)))
        return br, 1j * gam * f, 1j * gam * (f - 1)

    L, R = one(bs[0], 0.8), one(bs[-1], 0.3)
    obc = OBCBlocks(num_blocks=len(bs))
    for k, lst in enumerate((obc.retarded, obc.lesser, obc.greater)):
        lst[0], lst[-1] = L[k], R[k]
    t = torch.from_numpy
    host = (BlockTridiagonal.from_packed(bs, t(ad), t(au), t(al)),
            BlockTridiagonal.from_packed(bs, t(sld), t(slu), torch.zeros_like(t(slu))),
            BlockTridiagonal.from_packed(bs, t(sgd), t(sgu), torch.zeros_like(t(sgu))))
    return bs, host, obc


def _to(m, device):
    return BlockTridiagonal.from_packed(list(m.block_sizes), *(x.to(device) for x in m.packed()))


@pytest.mark.parametrize("E,max_batch,chunk_bytes", [(5, 100, None), (9, 4, None), (7, 100, 1)])
def test_host_pipeline_equals_device_solve(device, monkeypatch, E, max_batch, chunk_bytes):
    if chunk_bytes is not None:
        # one energy per chunk: every staging slot is reused several times
        monkeypatch.setattr(rgf_mod, "_HOST_CHUNK_BYTES", chunk_bytes)
    bs, host, obc = _case(E, 41 + E)
    outs_h = tuple(BlockTridiagonal(bs, (E,)) for _ in range(3))
    solver = RGF(max_batch_size=max_batch)
    cur_h = solver.selected_solve(*host, obc_blocks=obc, out=[outs_h[1], outs_h[2], outs_h[0]],
                                  return_retarded=True, return_current=True)
    assert solver._pipe is not None, "the host-resident batch did not take the chunked path"
    dev = tuple(_to(m, device) for m in host)
    outs_d = tuple(BlockTridiagonal(bs, (E,), device=device) for _ in range(3))
    cur_d = RGF(max_batch_size=max_batch).selected_solve(*dev, obc_blocks=obc,
                                                         out=[outs_d[1], outs_d[2], outs_d[0]],
                                                         return_retarded=True, return_current=True)
    for xh, xd in zip(outs_h, outs_d):
        for ph, pd in zip(xh.packed(), xd.packed()):
            assert ph.device.type == "cpu"
            assert np.array_equal(ph.numpy(), pd.cpu().numpy())
    assert np.array_equal(np.asarray(cur_h), cur_d.cpu().numpy())
    assert np.abs(cur_d.cpu().numpy()).max() > 0


@pytest.mark.parametrize("E,max_batch,chunk_bytes", [(6, 100, None), (7, 100, 1)])
def test_host_pipeline_selected_inv_equals_device(device, monkeypatch, E, max_batch, chunk_bytes):
    if chunk_bytes is not None:
        monkeypatch.setattr(rgf_mod, "_HOST_CHUNK_BYTES", chunk_bytes)
    bs, host, obc = _case(E, 71 + E)
    A = host[0]
    xh = BlockTridiagonal(bs, (E,))
    solver = RGF(max_batch_size=max_batch)
    solver.selected_inv(A, obc_blocks=obc, out=xh)
    assert solver._pipe is not None, "the host-resident batch did not take the chunked path"
    xd = BlockTridiagonal(bs, (E,), device=device)
    RGF(max_batch_size=max_batch).selected_inv(_to(A, device), obc_blocks=obc, out=xd)
    for ph, pd in zip(xh.packed(), xd.packed()):
        assert np.array_equal(ph.numpy(), pd.cpu().numpy())
    assert np.abs(xh.packed()[0].numpy()).max() > 0
