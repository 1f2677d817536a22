"""Seeded synthetic inputs for the RGF solve: tight-binding device + self-energies.

For a geometry ``(N, b)`` (or a list of block sizes) the generator draws, with
``numpy.random.default_rng(seed)``:

* ``H_ii = (X + X^H)/2`` with ``X ~ CN(0,1)/sqrt(b_i)`` (Hermitian diagonal blocks);
* ``H_{i,i+1} = 0.3 * CN(0,1)/sqrt(b_i)``, ``H_{i+1,i} = H_{i,i+1}^H``;
* ``Sigma^<_ii = i * (Y Y^H) / b_i`` with ``Y ~ CN(0,1)``, off-diagonal blocks zero;
* boundary self-energy ``Sigma_B = -0.5i * I`` on blocks ``0`` and ``N-1``;
* ``Sigma^> = Sigma^< - (Sigma^R - Sigma^R^H)`` with ``Sigma^R = Sigma_B``, i.e.
  ``Sigma^>_ii = Sigma^<_ii + i*I`` on the two boundary blocks.

``CN(0, s^2)`` is ``s * (randn + i randn) / sqrt(2)``.  The system matrix for
energy ``E`` and broadening ``eta`` is ``A = (E + i eta) I - H - Sigma_B``,
as assembled by the electron solver of the reference
(``src/quatrex/electron/solver.py:668-709``).  Energies are ``U(-1, 1)``
drawn from ``default_rng(777)``; device seeds are ``20260 + geometry_id``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from quatrex_amd.layout import Geometry

DEVICE_SEED_BASE = 20260
ENERGY_SEED = 777
ARRIVAL_SEED = 4242
GEOMETRY_SEED = 20265
BOUNDARY_SIGMA = -0.5j


def _cn(rng: np.random.Generator, shape, sigma: float = 1.0) -> np.ndarray:
    return sigma * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)


@dataclass
class DeviceModel:
    """Energy-independent blocks of one device geometry (packed, complex128)."""

    geometry: Geometry
    h_diag: np.ndarray  # (diag_size,)
    h_upper: np.ndarray  # (offdiag_size,)
    h_lower: np.ndarray  # (offdiag_size,)
    sl_diag: np.ndarray  # Sigma^< diagonal blocks
    sg_diag: np.ndarray  # Sigma^> diagonal blocks
    seed: int

    @property
    def sl_upper(self) -> np.ndarray:
        return np.zeros(self.geometry.offdiag_size, dtype=np.complex128)

    @property
    def sg_upper(self) -> np.ndarray:
        return np.zeros(self.geometry.offdiag_size, dtype=np.complex128)


def make_device(geometry: Geometry | tuple, seed: int) -> DeviceModel:
    """Draw H, Sigma^< and Sigma^> for ``geometry`` with ``default_rng(seed)``."""
    if not isinstance(geometry, Geometry):
        geometry = Geometry.uniform(*geometry)
    g = geometry
    rng = np.random.default_rng(seed)
    h_diag = np.empty(g.diag_size, dtype=np.complex128)
    h_upper = np.empty(g.offdiag_size, dtype=np.complex128)
    h_lower = np.empty(g.offdiag_size, dtype=np.complex128)
    sl_diag = np.empty(g.diag_size, dtype=np.complex128)
    for i, b in enumerate(g.block_sizes):
        x = _cn(rng, (b, b)) / np.sqrt(b)
        g.diag(h_diag, i)[...] = 0.5 * (x + x.conj().T)
    for i in range(g.n_blocks - 1):
        bi = g.block_sizes[i]
        u = 0.3 * _cn(rng, (bi, g.block_sizes[i + 1])) / np.sqrt(bi)
        g.upper(h_upper, i)[...] = u
        g.lower(h_lower, i)[...] = u.conj().T
    for i, b in enumerate(g.block_sizes):
        y = _cn(rng, (b, b))
        g.diag(sl_diag, i)[...] = 1j * (y @ y.conj().T) / b
    sg_diag = sl_diag.copy()
    # Sigma^R - Sigma^R^H = -1j * I on the boundary blocks.
    for i in sorted({0, g.n_blocks - 1}):
        blk = g.diag(sg_diag, i)
        blk[...] += 1j * np.eye(g.block_sizes[i])
    return DeviceModel(g, h_diag, h_upper, h_lower, sl_diag, sg_diag, seed)


def draw_energies(n: int, seed: int = ENERGY_SEED) -> np.ndarray:
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=n)


def boundary_blocks(geometry: Geometry) -> list[int]:
    return sorted({0, geometry.n_blocks - 1})


def system_matrix(
    model: DeviceModel, energies, eta, fold_boundary: bool = True
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Packed ``A = (E + i eta) I - H - Sigma_B`` for each energy (numpy).

    ``eta`` is a scalar or one value per energy.  With ``fold_boundary=False``
    the boundary term is left out, for callers that pass it as an OBC block.
    """
    g = model.geometry
    energies = np.atleast_1d(np.asarray(energies, dtype=np.float64))
    eta = np.broadcast_to(np.asarray(eta, dtype=np.float64), energies.shape)
    z = energies + 1j * eta
    ne = energies.size
    a_diag = np.broadcast_to(-model.h_diag, (ne, g.diag_size)).copy()
    for i, b in enumerate(g.block_sizes):
        blk = g.diag(a_diag, i)
        idx = np.arange(b)
        blk[:, idx, idx] += z[:, None]
        if fold_boundary and i in boundary_blocks(g):
            blk[:, idx, idx] -= BOUNDARY_SIGMA
    a_upper = np.broadcast_to(-model.h_upper, (ne, g.offdiag_size)).copy()
    a_lower = np.broadcast_to(-model.h_lower, (ne, g.offdiag_size)).copy()
    return a_diag, a_upper, a_lower


def sigmas(model: DeviceModel, n_energies: int):
    """Packed ``(sl_diag, sl_upper, sg_diag, sg_upper)`` repeated over energies."""
    g = model.geometry
    rep = lambda v: np.broadcast_to(v, (n_energies, v.shape[-1])).copy()  # noqa: E731
    return (
        rep(model.sl_diag),
        np.zeros((n_energies, g.offdiag_size), np.complex128),
        rep(model.sg_diag),
        np.zeros((n_energies, g.offdiag_size), np.complex128),
    )


def boundary_obc(model: DeviceModel, n_energies: int) -> dict[int, np.ndarray]:
    """Retarded OBC blocks ``{block: (E, b, b)}`` equal to ``Sigma_B``."""
    g = model.geometry
    return {
        i: np.broadcast_to(
            BOUNDARY_SIGMA * np.eye(g.block_sizes[i], dtype=np.complex128),
            (n_energies, g.block_sizes[i], g.block_sizes[i]),
        ).copy()
        for i in boundary_blocks(g)
    }


# --- torch (device) assembly for large benchmarks ---------------------------


def system_matrix_torch(model: DeviceModel, energies, eta, device, out=None):
    """Assemble packed ``A`` for many energies directly on ``device`` (torch).

    Used where the numpy route would not fit in host memory (thousands of
    256x256 blocks); the arithmetic is identical to :func:`system_matrix`.
    """
    import torch

    g = model.geometry
    energies = np.atleast_1d(np.asarray(energies, dtype=np.float64))
    eta = np.broadcast_to(np.asarray(eta, dtype=np.float64), energies.shape)
    ne = energies.size
    z = torch.as_tensor(energies + 1j * eta, dtype=torch.complex128, device=device)
    if out is None:
        a_diag = torch.empty((ne, g.diag_size), dtype=torch.complex128, device=device)
        a_upper = torch.empty((ne, g.offdiag_size), dtype=torch.complex128, device=device)
        a_lower = torch.empty((ne, g.offdiag_size), dtype=torch.complex128, device=device)
    else:
        a_diag, a_upper, a_lower = out
    h_diag = torch.as_tensor(model.h_diag, device=device)
    a_diag.copy_((-h_diag).expand(ne, -1))
    a_upper.copy_((-torch.as_tensor(model.h_upper, device=device)).expand(ne, -1))
    a_lower.copy_((-torch.as_tensor(model.h_lower, device=device)).expand(ne, -1))
    for i, b in enumerate(g.block_sizes):
        o = int(g.diag_offsets[i])
        idx = o + torch.arange(b, device=device) * (b + 1)
        shift = z[:, None].expand(ne, b)
        if i in boundary_blocks(g):
            shift = shift - BOUNDARY_SIGMA
        a_diag[:, idx] += shift
    return a_diag, a_upper, a_lower
