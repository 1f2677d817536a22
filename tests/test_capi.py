"""C-ABI checks that need no GPU: the library loads and exports exactly the
entry points declared in include/rgf.h."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from quatrex_amd import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "rgf.h"


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rgf_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert header_functions() == sorted(_lib.EXPORTED_SYMBOLS)


def test_library_built_and_exports_header_symbols():
    assert _lib.LIB_PATH.exists(), "run python -m quatrex_amd.build"
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(rgf_[a-z0-9_]+)$", out, flags=re.M))
    assert set(header_functions()) <= exported


def test_binding_declares_signatures():
    lib = _lib.load()
    assert lib.rgf_version().decode().startswith("quatrex_amd-rgf")
    assert lib.rgf_selected_solve_z.argtypes is not None


def test_workspace_query_without_gpu():
    from quatrex_amd.layout import Geometry
    from quatrex_amd.solve import workspace_size

    g = Geometry.uniform(32, 256)
    one = workspace_size(g, 1)
    many = workspace_size(g, 64)
    assert one > 11 * 256 * 256 * 16
    assert many > 60 * one
    assert workspace_size(g, 1, current=True) > one
    with pytest.raises(_lib.RgfError):
        workspace_size(Geometry.uniform(2, 2000), 1)


def test_null_context_is_rejected():
    lib = _lib.load()
    assert lib.rgf_selected_solve_z(None, None, 0, None, None, None, None, None, None, None,
                                    None) == _lib.RGF_EINVAL


def _inverse_aux_need(n: int) -> int:
    """Bytes the blocked inverse of one n x n block uses per energy: two
    pivot-row buffers of n x NB complex128 values plus two int tables of n
    (NB = 32 up to 512 rows, 16 above), rounded up to 256 bytes."""
    nb = 32 if n <= 512 else 16
    raw = 2 * n * nb * 16 + 2 * n * 4
    return (raw + 255) // 256 * 256


@pytest.mark.parametrize("sizes", [(520, 512), (600, 400, 600), (1024, 300), (700, 513, 512, 700)])
def test_workspace_holds_every_block_inverse(sizes):
    """Mixed geometries whose largest block uses 16-column panels and a smaller
    block 32-column panels: the per-energy inverse scratch must cover the
    larger need (a 512-row block needs more than a 600-row one)."""
    from quatrex_amd.layout import Geometry
    from quatrex_amd.solve import workspace_size

    g = Geometry(sizes)
    bmax = max(sizes)
    for E in (1, 3):
        need = E * (11 * bmax * bmax * 16 + max(_inverse_aux_need(b) for b in sizes))
        assert workspace_size(g, E) >= need, (sizes, E)


def test_autotune_energy_count_exceeds_the_cache():
    """The autotuner times products over an operand stack larger than the
    256 MB Infinity Cache, within [64, 1024] energies."""
    from quatrex_amd.autotune import default_energies

    for b in (16, 32, 64, 128, 256, 384, 512, 1024):
        E = default_energies(b)
        assert 64 <= E <= 1024
        assert E == 1024 or E == 64 or b * b * E * 16 >= 256 << 20
    assert default_energies(256) == 256 and default_energies(512) == 64
    assert default_energies(128) == 1024
