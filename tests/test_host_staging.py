"""Host logic of the chunked staging of host-resident stacks (no GPU): which
operands take it and when a block-sparse output must also be staged in.

The device maps are built on the CPU here; ``writes_every_entry`` only reads
their index and flag arrays.
"""

import numpy as np
import pytest
from scipy import sparse

from oracle.dsdb import OracleDSDB
from quatrex_amd.dsdb import DSDBCOO, DSDBCSR, writes_every_entry
from quatrex_amd.greens_function_solver import rgf as rgf_mod
from quatrex_amd.layout import Geometry
from tests.helpers import tridiagonal_mask

BS = (4, 6, 5)


def _pattern(extra: bool):
    mask = tridiagonal_mask(list(BS))
    if extra:
        mask[0, -2:] = mask[-1, 0] = True  # blocks (0, 2) and (2, 0)
    return mask


@pytest.mark.parametrize("layout", [DSDBCOO, DSDBCSR])
@pytest.mark.parametrize("extra", [False, True])
def test_writes_every_entry_follows_the_pattern(layout, extra):
    g = Geometry(BS)
    mask = _pattern(extra)
    m = layout.from_sparray(sparse.coo_matrix(mask.astype(np.complex128)), BS, (2,), device="cpu")
    assert writes_every_entry(m, g, "dul", m.nnz, "cpu") is (not extra)
    # a part left out leaves its stored entries as they were
    assert writes_every_entry(m, g, "du", m.nnz, "cpu") is False


def test_writes_every_entry_symmetric_storage():
    """Only the upper triangle is stored: diag + upper write all of it, the
    mirrored lower blocks write nothing."""
    g = Geometry(BS)
    up = np.triu(_pattern(False)).astype(np.complex128)
    m = DSDBCOO.from_sparray(sparse.coo_matrix(up), BS, (2,), symmetry=True, device="cpu")
    assert writes_every_entry(m, g, "du", m.nnz, "cpu") is True
    assert writes_every_entry(m, g, "d", m.nnz, "cpu") is False
    # the lower blocks' map entries are images of the stored upper blocks:
    # sparsifying them writes nothing, so the upper entries stay unwritten
    assert writes_every_entry(m, g, "dl", m.nnz, "cpu") is False


def test_host_operand_predicates():
    import torch

    mask = _pattern(False)
    r, c = np.nonzero(mask)
    o = OracleDSDB(np.zeros((3, r.size), np.complex128), r, c, BS)
    assert rgf_mod._host_dsdb(o) and rgf_mod._host_operand(o)
    m = DSDBCOO.from_sparray(sparse.coo_matrix(mask.astype(np.complex128)), BS, (3,), device="cpu")
    assert rgf_mod._host_dsdb(m)
    # strided energies, another dtype or byte order: not staged in chunks
    o.data = np.zeros((6, r.size), np.complex128)[::2]
    assert not rgf_mod._host_dsdb(o)
    o.data = np.zeros((3, r.size), np.complex64)
    assert not rgf_mod._host_dsdb(o)
    o.data = np.zeros((3, r.size), ">c16")
    assert not rgf_mod._host_dsdb(o)
    m.data = m.data.to(torch.complex64)
    assert not rgf_mod._host_dsdb(m)
    # a dense-block container is not block-sparse
    from quatrex_amd.datastructures import BlockTridiagonal
    bt = BlockTridiagonal(list(BS), (3,))
    assert not rgf_mod._host_dsdb(bt) and rgf_mod._host_packed(bt)
