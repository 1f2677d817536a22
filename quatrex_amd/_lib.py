"""ctypes binding of ``librgf_hip.so`` (the C ABI declared in ``include/rgf.h``).

There is no fallback: if the library is missing or fails to load, every
solver call raises.  Build it with ``python -m quatrex_amd.build``.
"""

from __future__ import annotations

import ctypes
import threading
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "librgf_hip.so"

RGF_OK = 0
RGF_EINVAL = 1
RGF_ESINGULAR = 2
RGF_EHIP = 3
RGF_ENOMEM = 4
RGF_EUNSUPPORTED = 5

RGF_OPT_MAX_CHUNK = 1
RGF_OPT_WORKSPACE_BUDGET = 2
RGF_OPT_TILE_CFG = 3
RGF_OPT_PROFILE = 4
RGF_OPT_UNFUSED = 5
RGF_OPT_OVERLAP = 6
RGF_OPT_PERSISTENT = 7
RGF_OPT_TILE_FOR_BLOCK = 1000  # + block size (FP64 products)
RGF_OPT_TILE_FOR_BLOCK_F32 = 3000  # + block size (FP32-compute products)
RGF_NUM_TILE_CFGS = 7
TILE_CFG_NAMES = ("64x64", "32x32", "64x32", "128x64", "64x64-3m", "32x32-3m", "128x64-3m")

# Every function declared in include/rgf.h (checked by the tests).
EXPORTED_SYMBOLS = (
    "rgf_create",
    "rgf_destroy",
    "rgf_set_stream",
    "rgf_set_option",
    "rgf_last_error",
    "rgf_status",
    "rgf_workspace_size",
    "rgf_set_workspace",
    "rgf_inv_z",
    "rgf_assemble_z",
    "rgf_assemble_system_z",
    "rgf_selected_inv_z",
    "rgf_selected_solve_z",
    "rgf_selected_solve_noalloc_z",
    "rgf_selected_solve_batch_z",
    "rgf_selected_solve_mixed",
    "rgf_profile_read",
    "rgf_autotune",
    "rgf_graph_capture_solve",
    "rgf_graph_launch",
    "rgf_graph_destroy",
    "rgf_densify_z",
    "rgf_sparsify_z",
    "rgf_sancho_rubio_z",
    "rgf_beyn_z",
    "rgf_spectral_obc_z",
    "rgf_obc_fixed_point_z",
    "rgf_contact_obc_z",
    "rgf_boundary_system_z",
    "rgf_version",
)


class BeynOpts_t(ctypes.Structure):
    _fields_ = [("r_o", ctypes.c_double), ("r_i", ctypes.c_double), ("m_0", c_int32),
                ("num_quad_points", c_int32), ("seed", ctypes.c_uint64)]


class SpectralOpts_t(ctypes.Structure):
    _fields_ = [("min_decay", ctypes.c_double), ("max_decay", ctypes.c_double),
                ("min_propagation", ctypes.c_double), ("residual_tolerance", ctypes.c_double),
                ("eta_decay", ctypes.c_double), ("num_ref_iterations", c_int32),
                ("residual_normalization", c_int32)]


class Geometry_t(ctypes.Structure):
    _fields_ = [("n_blocks", c_int32), ("block_sizes", POINTER(c_int64))]


class BtConst_t(ctypes.Structure):
    _fields_ = [("diag", c_void_p), ("upper", c_void_p), ("lower", c_void_p)]


class Bt_t(ctypes.Structure):
    _fields_ = [("diag", c_void_p), ("upper", c_void_p), ("lower", c_void_p)]


class Assembly_t(ctypes.Structure):
    _fields_ = [("z", c_void_p), ("hamiltonian", BtConst_t), ("hamiltonian_per_energy", c_int32),
                ("overlap", POINTER(BtConst_t)), ("potential", c_void_p),
                ("sigma_retarded_herm", POINTER(BtConst_t)), ("sigma_lesser", POINTER(BtConst_t)),
                ("sigma_greater", POINTER(BtConst_t))]


class Obc_t(ctypes.Structure):
    _fields_ = [
        ("retarded", POINTER(c_void_p)),
        ("lesser", POINTER(c_void_p)),
        ("greater", POINTER(c_void_p)),
    ]


class MixedOpts_t(ctypes.Structure):
    _fields_ = [
        ("refine", c_int32),
        ("fallback_tol", c_double),
        ("fallback_flags", c_void_p),
        ("residuals", c_void_p),
    ]


class SolveProblem_t(ctypes.Structure):
    _fields_ = [
        ("geom", Geometry_t),
        ("n_energies", c_int64),
        ("a", BtConst_t),
        ("sigma_lesser", BtConst_t),
        ("sigma_greater", BtConst_t),
        ("obc", POINTER(Obc_t)),
        ("g_retarded", Bt_t),
        ("g_lesser", Bt_t),
        ("g_greater", Bt_t),
        ("current", c_void_p),
    ]


class BlockMapPart_t(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("src", ctypes.c_void_p), ("dst", ctypes.c_void_p),
                ("op", ctypes.c_void_p), ("covers_part", ctypes.c_int32)]


class BlockMap_t(ctypes.Structure):
    _fields_ = [("part", BlockMapPart_t * 3)]


class Profile_t(ctypes.Structure):
    _fields_ = [
        ("gemm_launches", c_int64),
        ("gemm_ms", c_double),
        ("gemm_flops", c_double),
        ("inv_calls", c_int64),
        ("inv_ms", c_double),
        ("inv_flops", c_double),
        ("gemm_bytes", c_double),
    ]


class RgfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rgf error {code}: {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def _declare(lib: ctypes.CDLL) -> None:
    P = POINTER
    lib.rgf_create.argtypes = [c_int, c_void_p, P(c_void_p)]
    lib.rgf_destroy.argtypes = [c_void_p]
    lib.rgf_set_stream.argtypes = [c_void_p, c_void_p]
    lib.rgf_set_option.argtypes = [c_void_p, c_int32, c_int64]
    lib.rgf_last_error.argtypes = [c_void_p, c_char_p, c_size_t]
    lib.rgf_status.argtypes = [c_void_p, P(c_int64)]
    lib.rgf_workspace_size.argtypes = [P(Geometry_t), c_int64, c_int32, c_int32, c_int32, P(c_size_t)]
    lib.rgf_set_workspace.argtypes = [c_void_p, c_void_p, c_size_t]
    lib.rgf_inv_z.argtypes = [c_void_p, c_int64, c_int64, c_void_p, c_void_p]
    lib.rgf_assemble_z.argtypes = [c_void_p, P(Geometry_t), c_int64, c_void_p, P(BtConst_t),
                                   P(c_void_p), P(Bt_t)]
    lib.rgf_assemble_system_z.argtypes = [c_void_p, P(Geometry_t), c_int64, P(Assembly_t), P(Bt_t)]
    lib.rgf_selected_inv_z.argtypes = [c_void_p, P(Geometry_t), c_int64, P(BtConst_t), P(c_void_p),
                                       P(Bt_t)]
    lib.rgf_selected_solve_z.argtypes = [c_void_p, P(Geometry_t), c_int64, P(BtConst_t),
                                         P(BtConst_t), P(BtConst_t), P(Obc_t), P(Bt_t), P(Bt_t),
                                         P(Bt_t), c_void_p]
    lib.rgf_selected_solve_noalloc_z.argtypes = lib.rgf_selected_solve_z.argtypes
    lib.rgf_selected_solve_batch_z.argtypes = [c_void_p, c_int32, P(SolveProblem_t)]
    lib.rgf_selected_solve_mixed.argtypes = [c_void_p, P(Geometry_t), c_int64, P(BtConst_t),
                                             P(BtConst_t), P(BtConst_t), P(Obc_t),
                                             P(MixedOpts_t), P(Bt_t), P(Bt_t), P(Bt_t), c_void_p]
    lib.rgf_profile_read.argtypes = [c_void_p, P(Profile_t), c_int32]
    lib.rgf_autotune.argtypes = [c_void_p, c_int64, c_int64, c_int32, P(c_int32), P(ctypes.c_float)]
    lib.rgf_graph_capture_solve.argtypes = [c_void_p, P(Geometry_t), c_int64, P(BtConst_t),
                                            P(BtConst_t), P(BtConst_t), P(Obc_t), c_int32, c_int32,
                                            P(Bt_t), P(Bt_t), P(Bt_t), c_void_p, P(c_void_p)]
    lib.rgf_graph_launch.argtypes = [c_void_p]
    lib.rgf_graph_destroy.argtypes = [c_void_p]
    lib.rgf_densify_z.argtypes = [c_void_p, P(Geometry_t), c_int64, c_void_p, c_int64, P(BlockMap_t),
                                  P(Bt_t)]
    lib.rgf_sparsify_z.argtypes = [c_void_p, P(Geometry_t), c_int64, P(BtConst_t), P(BlockMap_t),
                                   c_void_p, c_int64]
    lib.rgf_sancho_rubio_z.argtypes = [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                       c_int32, ctypes.c_double, c_void_p, P(c_int32),
                                       P(ctypes.c_double)]
    lib.rgf_beyn_z.argtypes = [c_void_p, c_int64, c_int64, c_int32, P(c_void_p), P(BeynOpts_t),
                               c_void_p, c_void_p, c_void_p]
    lib.rgf_spectral_obc_z.argtypes = [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                       P(BeynOpts_t), P(SpectralOpts_t), c_void_p, c_void_p, c_void_p]
    lib.rgf_obc_fixed_point_z.argtypes = [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, c_void_p]
    lib.rgf_contact_obc_z.argtypes = [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p]
    lib.rgf_boundary_system_z.argtypes = [c_void_p, c_int64, c_int64, ctypes.c_double] + [c_void_p] * 9
    lib.rgf_version.argtypes = []
    lib.rgf_version.restype = c_char_p
    for name in EXPORTED_SYMBOLS:
        if name != "rgf_version":
            getattr(lib, name).restype = c_int


def load() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build the HIP extension with "
                    "`python -m quatrex_amd.build` (there is no CPU fallback)"
                )
            lib = ctypes.CDLL(str(LIB_PATH))
            _declare(lib)
            _lib = lib
    return _lib


def last_error(ctx) -> str:
    buf = ctypes.create_string_buffer(1024)
    load().rgf_last_error(ctx, buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, ctx=None) -> None:
    if rc == RGF_OK:
        return
    msg = last_error(ctx) if ctx else ""
    raise RgfError(rc, msg)


def geometry_struct(block_sizes) -> tuple[Geometry_t, object]:
    arr = (c_int64 * len(block_sizes))(*[int(b) for b in block_sizes])
    return Geometry_t(len(block_sizes), ctypes.cast(arr, POINTER(c_int64))), arr


def ptr_array(ptrs) -> tuple[object, object]:
    """Host array of device pointers (None -> NULL)."""
    arr = (c_void_p * len(ptrs))(*[p if p else None for p in ptrs])
    return ctypes.cast(arr, POINTER(c_void_p)), arr
