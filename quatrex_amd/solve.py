"""Device-level API over the C ABI: packed block-tridiagonal torch tensors.

All operands are contiguous complex128 torch tensors on a ROCm device in the
packed layout of :mod:`quatrex_amd.layout` (``(E, diag_size)`` and
``(E, offdiag_size)``).  PyTorch only provides the device buffers and the
stream; every arithmetic operation runs in the library's HIP kernels.
"""

from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from quatrex_amd import _lib
from quatrex_amd.layout import Geometry

__all__ = [
    "Context",
    "default_context",
    "selected_solve",
    "selected_inv",
    "inv",
    "workspace_size",
    "LinAlgError",
]

LinAlgError = np.linalg.LinAlgError


def _torch():
    import torch

    return torch


class Context:
    """One ``rgf_ctx``: a device, a HIP stream, workspace and tuning state."""

    def __init__(self, device=None, stream=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("no ROCm device visible: the RGF kernels need a GPU")
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else
                           torch.device(device).index or 0)
        self.device = dev
        if stream is None:
            stream = torch.cuda.current_stream(dev)
        self.stream = stream
        lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(lib.rgf_create(dev.index, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self._workspace = None

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.rgf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, option: int, value: int) -> None:
        _lib.check(self._lib.rgf_set_option(self._h, option, int(value)), self._h)

    def set_tile_for_block(self, block_size: int, cfg: int) -> None:
        self.set_option(_lib.RGF_OPT_TILE_FOR_BLOCK + int(block_size), cfg)

    def set_workspace(self, tensor) -> None:
        """Use a caller-owned device tensor as workspace (None: context-owned)."""
        self._workspace = tensor
        if tensor is None:
            _lib.check(self._lib.rgf_set_workspace(self._h, None, 0), self._h)
        else:
            nbytes = tensor.numel() * tensor.element_size()
            _lib.check(self._lib.rgf_set_workspace(self._h, ctypes.c_void_p(tensor.data_ptr()),
                                                   nbytes), self._h)

    def profile(self, on: bool = True) -> None:
        """Time every sweep GEMM launch and block inversion with HIP events."""
        self.set_option(_lib.RGF_OPT_PROFILE, 1 if on else 0)

    def profile_read(self, reset: bool = True) -> dict:
        p = _lib.Profile_t()
        _lib.check(self._lib.rgf_profile_read(self._h, ctypes.byref(p), int(reset)), self._h)
        return {k: getattr(p, k) for k, _ in _lib.Profile_t._fields_}

    def status(self) -> None:
        """Synchronise; raise ``LinAlgError`` if a pivot was exactly zero."""
        info = (ctypes.c_int64 * 3)()
        rc = self._lib.rgf_status(self._h, info)
        if rc == _lib.RGF_ESINGULAR:
            raise LinAlgError(f"Singular matrix (energy {info[0]}, block {info[1]}, "
                              f"column {info[2]})")
        _lib.check(rc, self._h)


_ctx_cache: dict = {}
_ctx_lock = threading.Lock()


def default_context(device=None) -> Context:
    """Context bound to ``device`` and torch's current stream there (cached)."""
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else
                       torch.device(device).index or 0)
    stream = torch.cuda.current_stream(dev)
    key = (dev.index, stream.cuda_stream)
    with _ctx_lock:
        ctx = _ctx_cache.get(key)
        if ctx is None:
            ctx = Context(dev, stream)
            _ctx_cache[key] = ctx
    return ctx


def _check(t, shape, name, dtype=None):
    torch = _torch()
    dtype = dtype or torch.complex128
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name}: must be a device tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    return t.data_ptr()


def _bt(triple, geo: Geometry, E: int, name: str, need=(True, True, True), const=True):
    triple = tuple(triple) + (None,) * (3 - len(tuple(triple)))
    d, u, l = triple
    shapes = ((E, geo.diag_size), (E, geo.offdiag_size), (E, geo.offdiag_size))
    ptrs = []
    for t, shp, req, part in zip((d, u, l), shapes, need, ("diag", "upper", "lower")):
        if t is None and req and not (part != "diag" and geo.n_blocks == 1):
            raise ValueError(f"{name}.{part} is required")
        ptrs.append(_check(t, shp, f"{name}.{part}") if t is not None else None)
    cls = _lib.BtConst_t if const else _lib.Bt_t
    return cls(*ptrs)


def _obc_lists(obc, geo: Geometry, E: int):
    """Normalise OBC input to three per-block lists of device tensors or None."""
    if obc is None:
        return None
    if hasattr(obc, "retarded"):
        lists = (obc.retarded, obc.lesser, obc.greater)
    else:
        lists = tuple(obc)
    out = []
    for kind, lst in zip(("retarded", "lesser", "greater"), lists):
        if lst is None:
            out.append([None] * geo.n_blocks)
            continue
        if isinstance(lst, dict):
            lst = [lst.get(i) for i in range(geo.n_blocks)]
        if len(lst) != geo.n_blocks:
            raise ValueError(f"obc.{kind} must have one entry per block")
        row = []
        for i, t in enumerate(lst):
            if t is None:
                row.append(None)
                continue
            b = geo.block_sizes[i]
            _check(t, (E, b, b), f"obc.{kind}[{i}]")
            row.append(t)
        out.append(row)
    return out


def _obc_struct(lists):
    if lists is None:
        return None, []
    keep = []
    ptrs = []
    for row in lists:
        if all(t is None for t in row):
            ptrs.append(None)
            continue
        arr, holder = _lib.ptr_array([t.data_ptr() if t is not None else None for t in row])
        keep.append(holder)
        ptrs.append(arr)
    return _lib.Obc_t(*ptrs), keep


@dataclass
class MixedOptions:
    refine: int = 1  # FP64 refinement steps; -1: FP64 Schur complements instead
    fallback_tol: float = 1e-8
    fallback_flags: object = None  # device int32 (E,)
    residuals: object = None  # device float64 (E,)


def selected_solve(geometry: Geometry, a, sigma_lesser, sigma_greater, *, g_lesser, g_greater,
                   g_retarded=None, obc=None, current=None, ctx: Context | None = None,
                   precision: str = "fp64", mixed: MixedOptions | None = None,
                   no_alloc: bool = False) -> None:
    """Selected solve on device buffers (asynchronous on the context stream).

    ``a``: ``(diag, upper, lower)``; ``sigma_*``: ``(diag, upper)``;
    ``g_lesser``/``g_greater``: ``(diag, upper[, lower])`` outputs;
    ``g_retarded``: ``None`` or ``(diag[, upper, lower])``; ``current``:
    ``None`` or a ``(E, N+1)`` complex128 tensor.  ``no_alloc``: the FP64
    solve through ``rgf_selected_solve_noalloc_z``, which allocates no device
    memory (its workspace is a fixed per-GPU array, shared by every such call
    on that GPU: overlapping calls race).
    """
    if no_alloc and precision != "fp64":
        raise ValueError("no_alloc is available for precision='fp64' only")
    ctx = ctx or default_context()
    geo = geometry
    E = int(a[0].shape[0])
    A = _bt(a, geo, E, "a")
    SL = _bt(sigma_lesser, geo, E, "sigma_lesser", (True, True, False))
    SG = _bt(sigma_greater, geo, E, "sigma_greater", (True, True, False))
    GL = _bt(g_lesser, geo, E, "g_lesser", (True, True, False), const=False)
    GG = _bt(g_greater, geo, E, "g_greater", (True, True, False), const=False)
    GR = (_bt(g_retarded, geo, E, "g_retarded", (True, False, False), const=False)
          if g_retarded is not None else None)
    cur = _check(current, (E, geo.n_blocks + 1), "current") if current is not None else None
    obc_lists = _obc_lists(obc, geo, E)
    obc_s, keep = _obc_struct(obc_lists)
    gstruct, gkeep = _lib.geometry_struct(geo.block_sizes)
    lib = _lib.load()
    byref = ctypes.byref
    if precision == "fp64":
        fn = lib.rgf_selected_solve_noalloc_z if no_alloc else lib.rgf_selected_solve_z
        rc = fn(ctx.handle, byref(gstruct), E, byref(A), byref(SL),
                                      byref(SG), byref(obc_s) if obc_s else None,
                                      byref(GR) if GR else None, byref(GL), byref(GG), cur)
    elif precision == "mixed":
        m = mixed or MixedOptions()
        torch = _torch()
        ff = _check(m.fallback_flags, (E,), "fallback_flags", torch.int32) if m.fallback_flags is not None else None
        rs = _check(m.residuals, (E,), "residuals", torch.float64) if m.residuals is not None else None
        opts = _lib.MixedOpts_t(int(m.refine), float(m.fallback_tol), ff, rs)
        rc = lib.rgf_selected_solve_mixed(ctx.handle, byref(gstruct), E, byref(A), byref(SL),
                                          byref(SG), byref(obc_s) if obc_s else None, byref(opts),
                                          byref(GR) if GR else None, byref(GL), byref(GG), cur)
    else:
        raise ValueError(f"unknown precision {precision!r}")
    del keep, gkeep
    if rc == _lib.RGF_EINVAL and current is not None and obc_lists is None:
        raise TypeError(_lib.last_error(ctx.handle))
    if rc == _lib.RGF_EINVAL:
        raise ValueError(_lib.last_error(ctx.handle))
    _lib.check(rc, ctx.handle)


@dataclass
class SolveProblem:
    """One selected solve of a batch (the arguments of :func:`selected_solve`)."""

    geometry: Geometry
    a: tuple
    sigma_lesser: tuple
    sigma_greater: tuple
    g_lesser: tuple
    g_greater: tuple
    g_retarded: tuple | None = None
    obc: object = None
    current: object = None


def selected_solve_batch(problems, ctx: Context | None = None) -> None:
    """Several FP64 selected solves in one library call (``rgf_selected_solve_batch_z``):
    problems whose blocks are all at most 32 share one persistent kernel
    launch (energy points of every such geometry in one device work queue),
    the others run one after the other.  Asynchronous on the context stream."""
    ctx = ctx or default_context()
    structs = (_lib.SolveProblem_t * max(len(problems), 1))()
    keep = []
    for k, p in enumerate(problems):
        geo = p.geometry
        E = int(p.a[0].shape[0])
        gstruct, gkeep = _lib.geometry_struct(geo.block_sizes)
        obc_lists = _obc_lists(p.obc, geo, E)
        obc_s, okeep = _obc_struct(obc_lists)
        cur = _check(p.current, (E, geo.n_blocks + 1), "current") if p.current is not None else None
        if cur is not None and obc_lists is None:
            raise TypeError("the current needs lesser and greater OBC blocks at both ends")
        st = structs[k]
        st.geom = gstruct
        st.n_energies = E
        st.a = _bt(p.a, geo, E, "a")
        st.sigma_lesser = _bt(p.sigma_lesser, geo, E, "sigma_lesser", (True, True, False))
        st.sigma_greater = _bt(p.sigma_greater, geo, E, "sigma_greater", (True, True, False))
        st.g_lesser = _bt(p.g_lesser, geo, E, "g_lesser", (True, True, False), const=False)
        st.g_greater = _bt(p.g_greater, geo, E, "g_greater", (True, True, False), const=False)
        if p.g_retarded is not None:
            st.g_retarded = _bt(p.g_retarded, geo, E, "g_retarded", (True, False, False), const=False)
        if obc_s:
            keep.append(obc_s)
            st.obc = ctypes.pointer(obc_s)
        st.current = cur
        keep.extend([gkeep, okeep])
    rc = _lib.load().rgf_selected_solve_batch_z(ctx.handle, len(problems), structs)
    del keep
    if rc == _lib.RGF_EINVAL:
        raise ValueError(_lib.last_error(ctx.handle))
    _lib.check(rc, ctx.handle)


def selected_inv(geometry: Geometry, a, *, x, obc_retarded=None, ctx: Context | None = None) -> None:
    """Selected inversion on device buffers (asynchronous)."""
    ctx = ctx or default_context()
    geo = geometry
    E = int(a[0].shape[0])
    A = _bt(a, geo, E, "a")
    X = _bt(x, geo, E, "x", (True, False, False), const=False)
    keep = None
    optr = None
    if obc_retarded is not None:
        lists = _obc_lists((obc_retarded, None, None), geo, E)
        optr, keep = _lib.ptr_array([t.data_ptr() if t is not None else None for t in lists[0]])
    gstruct, gkeep = _lib.geometry_struct(geo.block_sizes)
    rc = _lib.load().rgf_selected_inv_z(ctx.handle, ctypes.byref(gstruct), E, ctypes.byref(A), optr,
                                        ctypes.byref(X))
    del keep, gkeep
    if rc == _lib.RGF_EINVAL:
        raise ValueError(_lib.last_error(ctx.handle))
    _lib.check(rc, ctx.handle)


def inv(a, out=None, ctx: Context | None = None):
    """Batched inverse of ``(..., n, n)`` complex128 device matrices (``linalg.inv``)."""
    torch = _torch()
    ctx = ctx or default_context()
    if a.dim() < 2 or a.shape[-1] != a.shape[-2]:
        raise ValueError("inv expects (..., n, n)")
    a = a.contiguous()
    out = torch.empty_like(a) if out is None else out
    n = a.shape[-1]
    nm = a.numel() // (n * n)
    _check(a, a.shape, "a")
    _check(out, a.shape, "out")
    rc = _lib.load().rgf_inv_z(ctx.handle, n, nm, ctypes.c_void_p(a.data_ptr()),
                               ctypes.c_void_p(out.data_ptr()))
    _lib.check(rc, ctx.handle)
    return out


def assemble(geometry: Geometry, z, h, *, a, sigma_b=None, ctx: Context | None = None) -> None:
    """``A_e = z_e I - H - S_B`` on the device (asynchronous).

    ``z``: ``(E,)`` complex128 device tensor; ``h``: ``(diag, upper, lower)`` of
    ONE energy-independent H (1-D packed tensors); ``sigma_b``: ``None`` or a
    per-block list of ``(b_i, b_i)`` device tensors; ``a``: packed outputs.
    """
    ctx = ctx or default_context()
    geo = geometry
    E = int(z.shape[0])
    _check(z, (E,), "z")
    H = _lib.BtConst_t(*(_check(t, (n,), f"h.{p}") if t is not None else None
                         for t, n, p in zip(h, (geo.diag_size, geo.offdiag_size, geo.offdiag_size),
                                            ("diag", "upper", "lower"))))
    Aout = _bt(a, geo, E, "a", const=False)
    keep = None
    sb = None
    if sigma_b is not None:
        for i, t in enumerate(sigma_b):
            if t is not None:
                b = geo.block_sizes[i]
                _check(t, (b, b), f"sigma_b[{i}]")
        sb, keep = _lib.ptr_array([t.data_ptr() if t is not None else None for t in sigma_b])
    gstruct, gkeep = _lib.geometry_struct(geo.block_sizes)
    rc = _lib.load().rgf_assemble_z(ctx.handle, ctypes.byref(gstruct), E, ctypes.c_void_p(z.data_ptr()),
                                    ctypes.byref(H), sb, ctypes.byref(Aout))
    del keep, gkeep
    _lib.check(rc, ctx.handle)


def assemble_system(geometry: Geometry, z, hamiltonian, *, a, overlap=None, potential=None,
                    sigma_retarded_herm=None, sigma_lesser=None, sigma_greater=None,
                    ctx: Context | None = None) -> None:
    """The electron solver's system matrix on the tridiagonal blocks
    (``src/quatrex/electron/solver.py:668-709``), asynchronous:
    ``A_e = z_e S - P - (Sigma^R_herm + (Sigma^> - Sigma^<) / 2 + H)``.

    ``z``: ``(E,)`` complex128 (``E + i eta``); ``hamiltonian``: packed
    ``(diag, upper, lower)``, 1-D (one H for every energy) or ``(E, size)``;
    ``overlap``: ``None`` (identity) or packed 1-D ``(diag, upper, lower)``;
    ``potential``: ``None`` or a float64 ``(n,)`` tensor (``S_ab (v_a + v_b) / 2``
    with an overlap, ``diag(v)`` without); self-energies: ``None`` or packed
    ``(E, size)`` ``(diag, upper, lower)``; ``a``: packed outputs.
    """
    torch = _torch()
    ctx = ctx or default_context()
    geo = geometry
    E = int(z.shape[0])
    _check(z, (E,), "z")
    per_energy = hamiltonian[0].dim() == 2
    sizes = (geo.diag_size, geo.offdiag_size, geo.offdiag_size)
    names = ("diag", "upper", "lower")

    def shared(tri, name):
        tri = tuple(tri) + (None,) * (3 - len(tuple(tri)))
        out = []
        for t, n, part in zip(tri, sizes, names):
            if t is None:
                if part == "diag" or geo.n_blocks > 1:
                    raise ValueError(f"{name}.{part} is required")
                out.append(None)
            else:
                out.append(_check(t, (n,), f"{name}.{part}"))
        return _lib.BtConst_t(*out)

    H = _bt(hamiltonian, geo, E, "hamiltonian") if per_energy else shared(hamiltonian, "hamiltonian")
    S = shared(overlap, "overlap") if overlap is not None else None
    v = (_check(potential, (geo.n,), "potential", torch.float64) if potential is not None else None)
    sig = [_bt(t, geo, E, name) if t is not None else None
           for t, name in ((sigma_retarded_herm, "sigma_retarded_herm"), (sigma_lesser, "sigma_lesser"),
                           (sigma_greater, "sigma_greater"))]
    ref = lambda t: ctypes.pointer(t) if t is not None else None  # noqa: E731
    spec = _lib.Assembly_t(ctypes.c_void_p(z.data_ptr()), H, 1 if per_energy else 0, ref(S), v,
                           ref(sig[0]), ref(sig[1]), ref(sig[2]))
    Aout = _bt(a, geo, E, "a", const=False)
    gstruct, gkeep = _lib.geometry_struct(geo.block_sizes)
    rc = _lib.load().rgf_assemble_system_z(ctx.handle, ctypes.byref(gstruct), E, ctypes.byref(spec),
                                           ctypes.byref(Aout))
    del gkeep
    _lib.check(rc, ctx.handle)


class SolveGraph:
    """A captured selected-solve sweep for fixed device buffers (hipGraph).

    ``launch()`` replays the sweep on the context stream with the data the
    buffers hold at that time (e.g. after re-assembling ``A`` in place).
    """

    def __init__(self, geometry: Geometry, a, sigma_lesser, sigma_greater, *, g_lesser, g_greater,
                 g_retarded=None, obc=None, current=None, ctx: Context | None = None,
                 precision: str = "fp64", refine: int = 1):
        self.ctx = ctx or default_context()
        geo = geometry
        E = int(a[0].shape[0])
        A = _bt(a, geo, E, "a")
        SL = _bt(sigma_lesser, geo, E, "sigma_lesser", (True, True, False))
        SG = _bt(sigma_greater, geo, E, "sigma_greater", (True, True, False))
        GL = _bt(g_lesser, geo, E, "g_lesser", (True, True, False), const=False)
        GG = _bt(g_greater, geo, E, "g_greater", (True, True, False), const=False)
        GR = (_bt(g_retarded, geo, E, "g_retarded", (True, False, False), const=False)
              if g_retarded is not None else None)
        obc_lists = _obc_lists(obc, geo, E)
        obc_s, self._keep = _obc_struct(obc_lists)
        cur = _check(current, (E, geo.n_blocks + 1), "current") if current is not None else None
        gstruct, gkeep = _lib.geometry_struct(geo.block_sizes)
        h = ctypes.c_void_p()
        byref = ctypes.byref
        rc = _lib.load().rgf_graph_capture_solve(
            self.ctx.handle, byref(gstruct), E, byref(A), byref(SL), byref(SG),
            byref(obc_s) if obc_s else None, 1 if precision == "mixed" else 0, int(refine),
            byref(GR) if GR else None, byref(GL), byref(GG), cur, byref(h))
        if rc == _lib.RGF_EINVAL and current is not None and obc_lists is None:
            raise TypeError(_lib.last_error(self.ctx.handle))
        _lib.check(rc, self.ctx.handle)
        self._h = h
        # keep the buffers alive as long as the graph references them
        self._buffers = (a, sigma_lesser, sigma_greater, g_lesser, g_greater, g_retarded, obc,
                         current, obc_lists)

    def launch(self) -> None:
        _lib.check(_lib.load().rgf_graph_launch(self._h), self.ctx.handle)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().rgf_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def workspace_size(geometry: Geometry, n_energies: int, solve: bool = True, current: bool = False,
                   retarded: bool = True) -> int:
    """Device workspace bytes for one sweep over ``n_energies`` energies."""
    gstruct, keep = _lib.geometry_struct(geometry.block_sizes)
    out = ctypes.c_size_t()
    _lib.check(_lib.load().rgf_workspace_size(ctypes.byref(gstruct), int(n_energies), int(solve),
                                              int(current), int(retarded), ctypes.byref(out)))
    return int(out.value)


def packed_empty(geometry: Geometry, n_energies: int, device, parts: Sequence[str] = ("d", "u", "l"),
                 dtype=None):
    """Allocate packed ``(diag, upper, lower)`` device tensors (None for absent parts)."""
    torch = _torch()
    dtype = dtype or torch.complex128
    sizes = {"d": geometry.diag_size, "u": geometry.offdiag_size, "l": geometry.offdiag_size}
    return tuple(
        torch.empty((n_energies, sizes[p]), dtype=dtype, device=device) if p in parts else None
        for p in ("d", "u", "l")
    )
