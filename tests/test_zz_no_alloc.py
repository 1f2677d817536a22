"""The allocation-free selected solve (rgf_selected_solve_noalloc_z) against
rgf_selected_solve_z on the same inputs: two calls in a row, the second with
a smaller input, on the batched sweep (blocks of 256, several chunks of the
fixed array) and on the persistent sweep (blocks <= 32), plus the refusal of
a geometry that does not fit the array."""

import numpy as np
import pytest

from quatrex_amd import _lib
from quatrex_amd import solve as S
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry
from tests.test_gpu_persistent import _mixed_case
from tests.test_gpu_rgf import _t

pytestmark = pytest.mark.gpu

TOL = 1e-12  # same kernels and per-energy arithmetic: agreement to FP64 rounding


def _case(bs, E, seed):
    rng = np.random.default_rng(seed)
    g, A, SL, SG, obc = _mixed_case(list(bs), E, seed, rng)
    return g, A, SL, SG, obc


def _run(device, ctx, case, no_alloc):
    import torch

    g, A, SL, SG, obc = case
    E = A[0].shape[0]
    t = lambda x: _t(x, device)  # noqa: E731
    GR = S.packed_empty(g, E, device)
    GL = S.packed_empty(g, E, device)
    GG = S.packed_empty(g, E, device)
    cur = torch.empty((E, g.n_blocks + 1), dtype=torch.complex128, device=device)
    ob = [[t(x) if x is not None else None for x in lst] for lst in obc]
    S.selected_solve(g, tuple(map(t, A)), tuple(map(t, SL)), tuple(map(t, SG)), g_lesser=GL,
                     g_greater=GG, g_retarded=GR, obc=ob, current=cur, ctx=ctx, no_alloc=no_alloc)
    ctx.status()
    return [x for X in (GR, GL, GG) for x in X if x is not None] + [cur]


def _compare(got, ref):
    import torch

    for x, r in zip(got, ref):
        assert torch.isfinite(x).all()
        err = float(torch.linalg.norm(x - r) / torch.linalg.norm(r))
        assert err <= TOL, err


@pytest.mark.parametrize("first,second", [
    # batched sweep: 24 energies of 256-blocks need several chunks of the array
    (((256, 256, 256, 256), 24), ((256, 192, 256), 7)),
    # persistent sweep: more energy points than workgroups, then a smaller one
    (((32, 24, 32, 16, 32, 32), 600), ((16, 8, 16), 9)),
])
def test_no_alloc_equals_allocating_solve(device, first, second):
    ctx = S.Context(device)
    for k, (bs, E) in enumerate((first, second)):
        print(f"no-alloc case {k}: blocks {bs}, {E} energies", flush=True)
        case = _case(bs, E, 400 + 7 * k + len(bs))
        ref = _run(device, ctx, case, no_alloc=False)
        got = _run(device, ctx, case, no_alloc=True)
        _compare(got, ref)
    ctx.close()


def test_no_alloc_refuses_what_does_not_fit(device):
    """One energy point of two 1024-blocks needs more scratch than the fixed
    array holds: a clear error before anything is launched."""
    import torch

    g = Geometry((1024, 1024))
    E = 1
    z = lambda *sel: S.packed_empty(g, E, device, *sel)  # noqa: E731
    A, SL, SG = z(), z(("d", "u")), z(("d", "u"))
    for X in (A, SL, SG):
        for x in X:
            if x is not None:
                x.zero_()
    GL, GG = z(("d", "u")), z(("d", "u"))
    ctx = S.Context(device)
    with pytest.raises(_lib.RgfError, match="fixed array holds"):
        S.selected_solve(g, A, SL[:2], SG[:2], g_lesser=GL, g_greater=GG, ctx=ctx, no_alloc=True)
    ctx.close()
    torch.cuda.synchronize()
