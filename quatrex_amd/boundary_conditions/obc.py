"""Open boundary conditions backed by the gfx950 HIP library.

Mirrors, with the reference's names, signatures and warnings:

* ``OBCSolver`` / ``SanchoRubio`` (``src/qttools/boundary_conditions/obc/
  obc.py:11-48``, ``sancho_rubio.py:10-85``): ``SanchoRubio(max_iterations=100,
  convergence_tol=1e-6)(a_ii, a_ij, a_ji, contact)`` returns the surface
  Green's function of a stack of boundary blocks.  The whole stack runs in one
  call of ``rgf_sancho_rubio_z`` with the reference's batch-wide stopping rule;
  a ``RuntimeWarning`` is raised when no iteration met it.
* ``OBCSystem`` (``obc.py:51-202`` on ``boundary_system.py:42-373``): the
  memoising wrapper with its fixed-point refinement and residuals
  (``rgf_obc_fixed_point_z``), for one process (stack communicator of size 1).
  The ``"auto"`` mode keeps the reference's test: the number of converged
  stack entries (divided by the communicator size) is compared with
  ``agreement_threshold``, and a count above it recomputes from scratch
  (``boundary_system.py:300-313``).
* ``compute_contact_obc`` / ``compute_obc`` (``src/quatrex/electron/
  solver.py:420-529``, ``block_sections = 1``): boundary system
  (``rgf_boundary_system_z``), surface Green's function, and the retarded,
  lesser and greater boundary self-energies (``rgf_contact_obc_z``), written
  into an ``OBCBlocks`` for the RGF solver.  The right contact is evaluated on
  the unreversed blocks (``a_ii = A[n,n]``, ``a_ij = A[n,n-1]``,
  ``a_ji = A[n-1,n]``, ``sigma = A[n-1,n] g A[n,n-1]``), which equals the
  reference's reversed evaluation followed by the inverse reversal because
  every step commutes with the block-reversal permutation.

Inputs may be torch device tensors (used in place) or numpy arrays (staged to
the device; numpy results are returned for numpy inputs).  Every arithmetic
operation runs in the HIP library.
"""

from __future__ import annotations

import ctypes
import warnings
from abc import ABC, abstractmethod

import numpy as np

from quatrex_amd import _lib
from quatrex_amd import solve as _solve
from quatrex_amd.greens_function_solver.solver import OBCBlocks

__all__ = ["OBCSolver", "SanchoRubio", "OBCSystem", "compute_contact_obc", "compute_obc"]


def _torch():
    import torch

    return torch


def _ctx_for(t):
    return _solve.default_context(t.device)


def _stage(x, device=None):
    """``(tensor (E, b, b) on the device, lead shape, from_numpy)``."""
    torch = _torch()
    from_numpy = not isinstance(x, torch.Tensor)
    if from_numpy:
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.complex128)).to(device)
    else:
        t = x.to(device=device if device is not None else x.device, dtype=torch.complex128)
        if not t.is_cuda:
            t = t.to(torch.device("cuda", torch.cuda.current_device()))
    if t.dim() < 2 or t.shape[-1] != t.shape[-2]:
        raise ValueError("boundary blocks must be (..., b, b)")
    lead = tuple(t.shape[:-2])
    b = int(t.shape[-1])
    return t.reshape(-1, b, b).contiguous(), lead, from_numpy


def _unstage(t, lead, from_numpy):
    t = t.reshape(*lead, *t.shape[-2:])
    return t.cpu().numpy() if from_numpy else t


def _same(a, *others):
    for o in others:
        if o is not None and tuple(o.shape) != tuple(a.shape):
            raise ValueError(f"boundary blocks differ in shape: {tuple(a.shape)} vs {tuple(o.shape)}")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class OBCSolver(ABC):
    r"""Solver of the surface Green's function recursion
    ``x_ii = (a_ii - a_ji x_ii a_ij)^{-1}`` (``obc.py:11-48``)."""

    @abstractmethod
    def __call__(self, a_ii, a_ij, a_ji, contact: str):
        """Returns the surface Green's function of the boundary blocks."""


class SanchoRubio(OBCSolver):
    """Sancho-Rubio decimation on the device (``sancho_rubio.py:10-85``).

    After a call, ``iterations`` and ``delta`` hold the number of decimation
    steps run and the last value of the stopping criterion.
    """

    def __init__(self, max_iterations: int = 100, convergence_tol: float = 1e-6):
        self.max_iterations = max_iterations
        self.convergence_tol = convergence_tol
        self.iterations = 0
        self.delta = float("nan")

    def __call__(self, a_ii, a_ij, a_ji, contact: str):
        A, lead, from_numpy = _stage(a_ii)
        B, _, _ = _stage(a_ij, A.device)
        C, _, _ = _stage(a_ji, A.device)
        _same(A, B, C)
        torch = _torch()
        ctx = _ctx_for(A)
        E, b = A.shape[0], A.shape[-1]
        x = torch.empty_like(A)
        it = ctypes.c_int32()
        delta = ctypes.c_double()
        rc = _lib.load().rgf_sancho_rubio_z(ctx.handle, b, E, _p(A), _p(B), _p(C),
                                            int(self.max_iterations), float(self.convergence_tol),
                                            _p(x), ctypes.byref(it), ctypes.byref(delta))
        _lib.check(rc, ctx.handle)
        ctx.status()
        self.iterations = it.value
        self.delta = delta.value
        if not (self.delta < self.convergence_tol):
            warnings.warn("Surface Green's function did not converge.", RuntimeWarning)
        return _unstage(x, lead, from_numpy)


class _IdentityCompressor:
    """Keeps a private copy of the cached solution (no compression)."""

    def compress(self, x):
        return x.clone() if hasattr(x, "clone") else (x.copy() if hasattr(x, "copy") else x)

    def decompress(self, x):
        return x


_MEMO_MODES = ("off", "auto", "force", "force-after-first")


class OBCSystem:
    """Memoising OBC boundary system for one process, with the interface of
    the reference's ``OBCSystem`` (``obc.py:51-202`` on
    ``boundary_system.py:42-373``).

    A call returns ``(solution, rel_residuals, abs_residuals)``.  Where the
    solution comes from depends on ``memoization_mode``:

    * ``"off"``: the boundary solver, nothing cached;
    * ``"force-after-first"`` / ``"auto"``: the boundary solver on the first
      call for a contact, afterwards the cached solution of the last call
      refined by fixed-point steps;
    * ``"force"``: like ``"force-after-first"``, but the first call starts the
      fixed-point steps from ``inv(a_ii)`` instead of calling the solver;
    * ``"auto"`` re-solves from scratch when the refined cached solution
      passes the residual test (absolute or relative) on more than
      ``agreement_threshold`` times the stack-communicator size entries.  That
      is the reference's test as written (``boundary_system.py:300-313``):
      a count of converged entries, not a fraction, so with one process a
      cached solution survives only while no stack entry passes the test.

    Every returned solution has had one more fixed-point step (the one that
    yields its residuals) and, unless the mode is ``"off"``, is what the next
    call for the same contact starts from.  ``system_reducer`` must be
    ``None``: the reference's reducers act on its own array types.
    """

    def __init__(self, boundary_solver: OBCSolver, cache_compressor=None, system_reducer=None,
                 num_ref_iterations: int = 2, relative_tol: float = 2e-1,
                 absolute_tol: float = 1e-6, warning_threshold: float = 1e-1,
                 memoization_mode: str = "auto", agreement_threshold: float = 0.999) -> None:
        compressor = cache_compressor if cache_compressor is not None else _IdentityCompressor()
        for method in ("compress", "decompress"):
            if not callable(getattr(compressor, method, None)):
                raise ValueError(f"cache_compressor needs a {method}() method")
        if not 0.0 <= agreement_threshold <= 1.0:
            raise ValueError(f"agreement_threshold must lie in [0, 1], got {agreement_threshold}")
        if system_reducer is not None:
            raise NotImplementedError("only the identity system reducer is supported")
        if num_ref_iterations < 2:
            warnings.warn(f"num_ref_iterations={num_ref_iterations} raised to 2: the residual "
                          "test needs at least 2 fixed-point steps", RuntimeWarning)
            num_ref_iterations = 2
        self.boundary_solver = boundary_solver
        self.cache_compressor = compressor
        self.num_ref_iterations = num_ref_iterations
        self.relative_tol = relative_tol
        self.absolute_tol = absolute_tol
        self.warning_threshold = warning_threshold
        self.memoization_mode = memoization_mode
        self.agreement_threshold = agreement_threshold
        self._cache = {}

    # -- device steps -----------------------------------------------------
    def _refine(self, system, solution, residuals: bool):
        torch = _torch()
        a_ii, a_ij, a_ji = system
        _same(a_ii, a_ij, a_ji, solution)
        b = int(a_ii.shape[-1])
        E = a_ii.numel() // (b * b)
        ctx = _ctx_for(a_ii)
        x_ref = torch.empty_like(a_ii)
        abs_res = rel_res = None
        if residuals:
            abs_res = torch.empty(a_ii.shape[:-2], dtype=torch.float64, device=a_ii.device)
            rel_res = torch.empty_like(abs_res)
        rc = _lib.load().rgf_obc_fixed_point_z(ctx.handle, b, E, _p(a_ii), _p(a_ij), _p(a_ji),
                                               _p(solution), _p(x_ref), _p(abs_res), _p(rel_res))
        _lib.check(rc, ctx.handle)
        ctx.status()
        return x_ref, abs_res, rel_res

    def _fixed_point_step(self, boundary_system, solution):
        """``inv(a_ii - a_ji x a_ij)`` (``obc.py:102-127``)."""
        return self._refine(boundary_system, solution, False)[0]

    def _get_starting_guess(self, boundary_system):
        """``inv(a_ii)`` (``obc.py:129-152``)."""
        return _solve.inv(boundary_system[0], ctx=_ctx_for(boundary_system[0]))

    def _get_residuals(self, boundary_system, test_solution):
        """``(rel, abs, x_ref)`` after one fixed-point step (``obc.py:154-202``)."""
        if not isinstance(test_solution, _torch().Tensor):
            raise TypeError(f"test_solution must be a torch.Tensor, not {type(test_solution).__name__}")
        x_ref, abs_res, rel_res = self._refine(boundary_system, test_solution, True)
        return rel_res, abs_res, x_ref

    # -- where a solution comes from ----------------------------------------
    def _from_solver(self, system, contact: str, kwargs):
        x = self.boundary_solver(*system, contact, **kwargs)
        self._remember(contact, x)
        return x

    def _remember(self, contact: str, x) -> None:
        if self.memoization_mode != "off":
            self._cache[contact] = self.cache_compressor.compress(x)

    def _starting_point(self, system, contact: str):
        """A solution to refine instead of solving, or None to solve afresh."""
        mode = self.memoization_mode
        if mode == "off":
            return None
        cached = self.cache_compressor.decompress(self._cache.get(contact))
        if cached is not None:
            return cached
        if mode == "force":
            return self._get_starting_guess(system)
        if mode in ("auto", "force-after-first"):
            return None
        raise RuntimeError(f"memoization_mode {mode!r} is none of {', '.join(_MEMO_MODES)}")

    def _stack_converged_count(self, rel, ab) -> int:
        ok = (ab < self.absolute_tol) | (rel < self.relative_tol)
        return int(ok.sum().item())

    def _memoized_solve(self, boundary_system, contact: str, **kwargs):
        start = self._starting_point(boundary_system, contact)
        if start is None:
            return self._from_solver(boundary_system, contact, kwargs)
        rel, ab, x = self._get_residuals(boundary_system, start)
        if self.memoization_mode == "auto":
            ranks = 1  # this process holds the whole stack
            if self._stack_converged_count(rel, ab) / ranks > self.agreement_threshold:
                return self._from_solver(boundary_system, contact, kwargs)
        for _step in range(self.num_ref_iterations - 2):
            x = self._fixed_point_step(boundary_system, x)
        return x

    def __call__(self, boundary_system, contact: str, **kwargs):
        """Returns ``(solution, rel_residuals, abs_residuals)``
        (``boundary_system.py:318-373``)."""
        staged = [_stage(x) for x in boundary_system]
        lead, from_numpy = staged[0][1], staged[0][2]
        system = tuple(s[0] for s in staged)
        x = self._memoized_solve(system, contact, **kwargs)
        if not isinstance(x, _torch().Tensor):
            x = _stage(x, system[0].device)[0]
        rel, ab, x = self._get_residuals(system, x.reshape(system[0].shape))
        if bool(((rel > self.warning_threshold) & (ab > self.absolute_tol)).any().item()):
            warnings.warn(
                f"{type(self.boundary_solver).__name__} on contact {contact}: fixed-point "
                f"residual up to {rel.max().item():.3e} relative, {ab.max().item():.3e} absolute",
                RuntimeWarning)
        self._remember(contact, x)
        rel, ab = rel.reshape(lead), ab.reshape(lead)
        if from_numpy:
            rel, ab = rel.cpu().numpy(), ab.cpu().numpy()
        return _unstage(x, lead, from_numpy), rel, ab


def compute_contact_obc(m_10, m_00, m_01, occupancies, eta_obc: float, obc, contact: str = "left",
                        overlap_blocks=None):
    """Retarded, lesser and greater boundary self-energies of one contact
    (``_compute_contact_obc``, ``solver.py:420-495``).

    ``m_10, m_00, m_01``: the contact's boundary blocks, in the unreversed
    orientation for both contacts (left: ``A[1,0], A[0,0], A[0,1]``; right:
    ``A[n-1,n], A[n,n], A[n,n-1]``); ``occupancies``: ``(E,)`` real;
    ``obc``: an :class:`OBCSystem` (or any callable returning the solution
    first); ``overlap_blocks``: ``None`` or ``(s_10, s_00, s_01)`` overlap
    blocks in the same orientation.  Returns device tensors (numpy for numpy
    inputs) shaped like ``m_00``.
    """
    torch = _torch()
    M00, lead, from_numpy = _stage(m_00)
    dev = M00.device
    M10, _, _ = _stage(m_10, dev)
    M01, _, _ = _stage(m_01, dev)
    _same(M00, M10, M01)
    E, b = M00.shape[0], M00.shape[-1]
    S = [None, None, None]
    if overlap_blocks is not None:
        S = [_stage(s, dev)[0] for s in overlap_blocks]
        _same(M00, *S)
    ctx = _ctx_for(M00)
    lib = _lib.load()
    a_ii = torch.empty_like(M00)
    with_overlap = overlap_blocks is not None
    a_ij = torch.empty_like(M00) if with_overlap else M01
    a_ji = torch.empty_like(M00) if with_overlap else M10
    rc = lib.rgf_boundary_system_z(ctx.handle, b, E, float(eta_obc), _p(M00), _p(M01), _p(M10),
                                   _p(S[1]), _p(S[2]), _p(S[0]), _p(a_ii),
                                   _p(a_ij) if with_overlap else None,
                                   _p(a_ji) if with_overlap else None)
    _lib.check(rc, ctx.handle)
    result = obc((a_ii, a_ij, a_ji), contact="G: " + contact)
    g_00 = result[0] if isinstance(result, tuple) else result
    g_00 = _stage(g_00, dev)[0]
    _same(M00, g_00)
    occ = torch.as_tensor(np.asarray(occupancies, dtype=np.float64) if not isinstance(
        occupancies, torch.Tensor) else occupancies, dtype=torch.float64).to(dev).reshape(-1)
    if occ.numel() != E:
        raise ValueError("The shape of the stack and the factor do not match.")
    occ = occ.contiguous()
    sigma = torch.empty_like(M00)
    lesser = torch.empty_like(M00)
    greater = torch.empty_like(M00)
    rc = lib.rgf_contact_obc_z(ctx.handle, b, E, _p(M10), _p(g_00), _p(M01), _p(occ), _p(sigma),
                               _p(lesser), _p(greater))
    _lib.check(rc, ctx.handle)
    return tuple(_unstage(t, lead, from_numpy) for t in (sigma, lesser, greater))


def _reversed(x):
    """Block reversal ``x[..., ::-1, ::-1]`` (``order_block(x, "reverse")``)."""
    if isinstance(x, _torch().Tensor):
        return x.flip((-2, -1))
    return np.ascontiguousarray(np.asarray(x)[..., ::-1, ::-1])


def compute_obc(system_matrix, left_occupancies, right_occupancies, eta_obc: float, obc,
                obc_blocks: OBCBlocks | None = None, batch_slice=None, overlap=None) -> OBCBlocks:
    """Boundary self-energies of both contacts into ``obc_blocks``
    (``_compute_obc``, ``solver.py:497-529``, one process holding every block).

    ``system_matrix`` (and ``overlap``) expose ``.blocks[i, j]`` stacks
    (``BlockTridiagonal``, the block-sparse containers, or the reference's
    ``DSDBSparse``); ``batch_slice`` names the cache entries like the reference.
    """
    n = int(system_matrix.num_blocks) - 1
    if n < 1:
        raise ValueError("open boundaries need at least two blocks")
    blocks = system_matrix.blocks
    tag = str(batch_slice)
    obc_blocks = obc_blocks if obc_blocks is not None else OBCBlocks(n + 1)
    ov = overlap.blocks if overlap is not None else None
    left = compute_contact_obc(blocks[1, 0], blocks[0, 0], blocks[0, 1], left_occupancies, eta_obc,
                               obc, contact="left-" + tag,
                               overlap_blocks=None if ov is None else (ov[1, 0], ov[0, 0], ov[0, 1]))
    obc_blocks.retarded[0], obc_blocks.lesser[0], obc_blocks.greater[0] = left
    m = n - 1
    # The reference reverses the system-matrix blocks of the right contact but
    # adds the overlap blocks unreversed (solver.py:461-476); in the unreversed
    # frame used here that is the reversed overlap.
    right = compute_contact_obc(blocks[m, n], blocks[n, n], blocks[n, m], right_occupancies,
                                eta_obc, obc, contact="right-" + tag,
                                overlap_blocks=None if ov is None else
                                tuple(_reversed(ov[i, j]) for i, j in ((m, n), (n, n), (n, m))))
    obc_blocks.retarded[-1], obc_blocks.lesser[-1], obc_blocks.greater[-1] = right
    return obc_blocks
