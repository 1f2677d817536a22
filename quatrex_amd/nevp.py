"""Polynomial (non-linear) eigenvalue solvers on the device.

Mirrors ``qttools.nevp`` (``src/qttools/nevp/nevp.py:8-40``,
``beyn.py:14-379``) with the reference's names and signatures:
``Beyn(r_o, r_i, m_0, num_quad_points, ...)(a_xx)`` returns ``(ws, vs)`` for
the problem ``(sum_n a_n w^n) v = 0``, ``n = -p .. p``, with the coefficient
blocks ``a_xx`` given lowest order first.  The whole batch runs in one call of
``rgf_beyn_z``: operator inverses on the quadrature nodes, the moments, the
SVD projection and the reduced eigenproblem are HIP kernels.

Differences from the reference, all outside the returned eigenpairs:

* the probe block ``Y`` comes from a counter-based generator on the device
  (seed 42, advanced by one per call, as the reference's
  ``default_rng(42)`` is advanced by every call), not from numpy's stream;
* columns beyond the rank of ``P_0`` hold zeros with eigenvalue 0, where the
  reference leaves the probe columns (eigenvalue 0 as well);
* ``eig_compute_location``, ``project_compute_location``,
  ``num_threads_contour``, ``contour_batch_size`` and ``use_pinned_memory``
  are accepted and ignored (everything runs on the device);
  ``use_qr=True`` (the QR projection, ``beyn.py:147-181``) is not built and
  raises ``NotImplementedError``.
"""

from __future__ import annotations

import ctypes
from abc import ABC, abstractmethod

import numpy as np

from quatrex_amd import _lib
from quatrex_amd import solve as _solve

__all__ = ["NEVP", "Beyn"]


class NEVP(ABC):
    """Abstract polynomial eigenvalue solver (``nevp.py:8-40``)."""

    @abstractmethod
    def __call__(self, a_xx):
        """Returns ``(ws, vs)`` for the coefficient blocks ``a_xx``."""


def _to_device(x, device=None):
    import torch

    if isinstance(x, torch.Tensor):
        t = x.to(dtype=torch.complex128)
        if not t.is_cuda:
            t = t.to(device or torch.device("cuda", torch.cuda.current_device()))
        return t, False
    dev = device or torch.device("cuda", torch.cuda.current_device())
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.complex128)).to(dev), True


class Beyn(NEVP):
    """Beyn's contour-integral method (``beyn.py:14-379``) on the device."""

    def __init__(self, r_o: float, r_i: float, m_0: int, num_quad_points: int,
                 num_threads_contour: int = 1024, eig_compute_location: str = "numpy",
                 project_compute_location: str = "numpy", use_qr: bool = False,
                 contour_batch_size: int | None = None, use_pinned_memory: bool = True):
        if use_qr:
            raise NotImplementedError("the QR projection of Beyn's method is not built; use the SVD one")
        self.r_o = r_o
        self.r_i = r_i
        self.m_0 = m_0
        self.num_quad_points = num_quad_points
        self.num_threads_contour = num_threads_contour
        self.eig_compute_location = eig_compute_location
        self.project_compute_location = project_compute_location
        self.use_qr = use_qr
        self.contour_batch_size = contour_batch_size or num_quad_points
        self.use_pinned_memory = use_pinned_memory
        self.seed = 42
        self._calls = 0

    def next_seed(self) -> int:
        """Seed of the probe block of the next call (advanced per call)."""
        s = (self.seed * 0x9E3779B97F4A7C15 + self._calls) & 0xFFFFFFFFFFFFFFFF
        self._calls += 1
        return s

    def options(self) -> "_lib.BeynOpts_t":
        return _lib.BeynOpts_t(float(self.r_o), float(self.r_i), int(self.m_0),
                               int(self.num_quad_points), self.next_seed())

    def __call__(self, a_xx):
        import torch

        if len(a_xx) % 2 == 0 or not 3 <= len(a_xx) <= 7:
            raise ValueError("a_xx must hold 3, 5 or 7 coefficient blocks")
        first, from_numpy = _to_device(a_xx[0])
        blocks = [first] + [_to_device(a, first.device)[0] for a in a_xx[1:]]
        # 2-D blocks are one batch entry (beyn.py:355-357)
        batch_shape = tuple(first.shape[:-2]) if first.dim() > 2 else (1,)
        d = int(first.shape[-1])
        blocks = [b.reshape(-1, d, d).contiguous() for b in blocks]
        E = blocks[0].shape[0]
        for b in blocks[1:]:
            if tuple(b.shape) != tuple(blocks[0].shape):
                raise ValueError("coefficient blocks differ in shape")
        m = int(self.m_0)
        ws = torch.empty((E, m), dtype=torch.complex128, device=first.device)
        vs = torch.empty((E, d, m), dtype=torch.complex128, device=first.device)
        ptrs = (ctypes.c_void_p * len(blocks))(*[b.data_ptr() for b in blocks])
        ctx = _solve.default_context(first.device)
        opts = self.options()
        rc = _lib.load().rgf_beyn_z(ctx.handle, d, E, len(blocks), ptrs, ctypes.byref(opts),
                                    ctypes.c_void_p(ws.data_ptr()), ctypes.c_void_p(vs.data_ptr()),
                                    None)
        _lib.check(rc, ctx.handle)
        ctx.status()
        ws = ws.reshape(*batch_shape, m)
        vs = vs.reshape(*batch_shape, d, m)
        if from_numpy:
            return ws.cpu().numpy(), vs.cpu().numpy()
        return ws, vs
