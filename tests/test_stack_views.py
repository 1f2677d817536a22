"""Substack views of the block-sparse stacks (CPU): the shape of the
reference's ``_DStackView`` (parent + stack index, no structure of its own),
index composition, writes reaching the parent, and the drop-in's recognition
of such views as block-sparse stacks."""

import numpy as np
import pytest
from scipy import sparse

from quatrex_amd.dsdb import DSDBCOO, DSDBCSR, StackView, structure_of
from quatrex_amd.greens_function_solver.rgf import _is_dsdb


def _stack(layout, E=7):
    bs = [3, 4, 3]
    rng = np.random.default_rng(4)
    m = sparse.random(10, 10, density=0.4, random_state=rng, format="coo").astype(np.complex128)
    m.setdiag(np.arange(1, 11))
    c = layout.from_sparray(m, bs, (E,), device="cpu")
    import torch

    vals = torch.arange(E * c.nnz, dtype=torch.float64).reshape(E, c.nnz).to(torch.complex128)
    c.data.copy_(vals)
    return c


@pytest.mark.parametrize("layout", [DSDBCOO, DSDBCSR])
def test_view_is_reference_shaped_and_composes(layout):
    c = _stack(layout)
    v = c.stack[1:6]
    assert isinstance(v, StackView)
    assert v._dsdbsparse is c and structure_of(v) is c
    assert not hasattr(v, "rows") and not hasattr(v, "rowptr_map") and not hasattr(v, "cols")
    assert v.shape == (5, 10, 10) and v.num_blocks == 3 and v.symmetry == c.symmetry
    assert np.array_equal(v.data.numpy(), c.data[1:6].numpy())
    w = v.stack[1:4]
    assert w._dsdbsparse is c and w._stack_index == (slice(2, 5, 1),)
    assert np.array_equal(w.data.numpy(), c.data[2:5].numpy())
    s = c.stack[0:7:2].stack[1:3]
    assert np.array_equal(s.data.numpy(), c.data[2:6:2].numpy())
    assert _is_dsdb(c) and _is_dsdb(v) and _is_dsdb(w)


def test_view_writes_reach_parent():
    c = _stack(DSDBCOO)
    v = c.stack[2:4]
    v.data[...] = 0
    assert float(c.data[2:4].abs().sum()) == 0.0
    assert float(c.data[1].abs().sum()) > 0 and float(c.data[4].abs().sum()) > 0


def test_non_stack_objects_are_not_block_sparse():
    class Dense:
        data = np.zeros((2, 3))
        block_sizes = [1, 2]

    class View:
        def __init__(self):
            self._dsdbsparse = Dense()
            self.data = np.zeros((1, 3))
            self.block_sizes = [1, 2]

    assert not _is_dsdb(Dense()) and not _is_dsdb(View())
