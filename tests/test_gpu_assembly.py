"""GPU parity of the full system-matrix assembly (rgf_assemble_system_z)
against the oracle restatement of the electron solver's assembly, on uneven
block sizes, with and without overlap, potential and each self-energy, with
one Hamiltonian for all energies or one per energy.  Bar: 1e-14 relative per
element scale (elementwise arithmetic; FMA contraction may change the last
bit), and the assembled matrix drives the selected solve like the generator's."""

import numpy as np
import pytest

from oracle import assembly as AS
from quatrex_amd import solve as S
from quatrex_amd.layout import Geometry

pytestmark = pytest.mark.gpu


def _herm(rng, n, scale=1.0):
    x = scale * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return 0.5 * (x + x.conj().T)


def _packed(g, dense, device):
    import torch

    d = dense if dense.ndim == 3 else dense[None]
    parts = g.from_dense(d)
    out = tuple(torch.as_tensor(np.ascontiguousarray(p), device=device) for p in parts)
    return out if dense.ndim == 3 else tuple(t[0].contiguous() for t in out)


@pytest.mark.parametrize("bs", [(5, 3, 7, 4), (16,), (32, 32, 24)])
@pytest.mark.parametrize("overlap", [False, True])
@pytest.mark.parametrize("per_energy_h", [False, True])
def test_assembly_vs_oracle(device, bs, overlap, per_energy_h):
    import torch

    rng = np.random.default_rng(sum(bs) + 2 * overlap + per_energy_h)
    g = Geometry(bs)
    n = g.n
    E = 5
    mask = AS.tridiagonal_mask(bs)
    en = rng.uniform(-1, 1, E)
    eta = 1e-3
    h = np.stack([_herm(rng, n) for _ in range(E)]) if per_energy_h else _herm(rng, n)
    s = np.where(mask, _herm(rng, n, 0.1) + np.eye(n), 0) if overlap else None
    v = rng.standard_normal(n)
    sr = np.stack([_herm(rng, n, 0.3) for _ in range(E)])
    sl = np.stack([1j * _herm(rng, n, 0.2) for _ in range(E)])
    sg = np.stack([1j * _herm(rng, n, 0.2) for _ in range(E)])
    z = torch.as_tensor(en + 1j * eta, device=device)
    for with_sigma in (False, True):
        for with_v in (False, True):
            ref = AS.assemble(bs, en, eta, h, s, v if with_v else None,
                              *((sr, sl, sg) if with_sigma else (None, None, None)))
            a = S.packed_empty(g, E, device)
            S.assemble_system(g, z, _packed(g, h, device), a=a,
                              overlap=_packed(g, s, device) if overlap else None,
                              potential=torch.as_tensor(v, device=device) if with_v else None,
                              sigma_retarded_herm=_packed(g, sr, device) if with_sigma else None,
                              sigma_lesser=_packed(g, sl, device) if with_sigma else None,
                              sigma_greater=_packed(g, sg, device) if with_sigma else None)
            S.default_context(device).status()
            got = g.to_dense(*(t.cpu().numpy() for t in a))
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * np.abs(ref).max())


def test_assembly_refuses_partial_operands(device):
    import torch

    g = Geometry((4, 4))
    E = 2
    z = torch.zeros(E, dtype=torch.complex128, device=device)
    h = tuple(torch.zeros(k, dtype=torch.complex128, device=device)
              for k in (g.diag_size, g.offdiag_size, g.offdiag_size))
    a = S.packed_empty(g, E, device)
    sl = S.packed_empty(g, E, device, ("d", "u"))
    with pytest.raises(ValueError):
        S.assemble_system(g, z, h, a=a, sigma_lesser=sl)
    with pytest.raises(ValueError):
        S.assemble_system(g, z, h[:1], a=a)
