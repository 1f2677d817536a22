"""GEMM tile autotuner keyed on block size, with a JSON cache.

For each block size ``b`` (and precision) the library times every tile
configuration of the grouped complex GEMM on a representative set of sweep
products (``rgf_autotune``) and keeps the fastest; results are cached per
device name in ``~/.cache/quatrex_amd/tiles.json`` (override with
``QX_TILE_CACHE``) and applied to a context with :func:`apply`.
"""

from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path

from quatrex_amd import _lib


def cache_path() -> Path:
    env = os.environ.get("QX_TILE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "quatrex_amd" / "tiles.json"


def _device_key(ctx) -> str:
    import torch

    return torch.cuda.get_device_name(ctx.device)


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except Exception:
        return {}


def default_energies(block_size: int) -> int:
    """Energies per timed product: enough that one operand stack (b^2 E
    complex128) exceeds the 256 MB Infinity Cache, as in a production sweep,
    within [64, 1024].  The ranking of the tiles changes with it: at b = 256
    the 128x64 3M tile wins with 64 energies (operands cache-resident) and the
    64x64 3M tile with 256-512."""
    b = max(1, int(block_size))
    return int(min(1024, max(64, -(-(256 << 20) // (16 * b * b)))))


def tune(ctx, block_size: int, n_energies: int | None = None, f32: bool = False,
         path: Path | None = None, force: bool = False) -> dict:
    """Best tile configuration for ``block_size`` (timed or from the cache;
    a cached entry timed at another energy count is timed again)."""
    path = path or cache_path()
    db = _load(path)
    dev = _device_key(ctx)
    key = f"{'f32' if f32 else 'f64'}:{int(block_size)}"
    if n_energies is None:
        n_energies = default_energies(block_size)
    entry = db.get(dev, {}).get(key)
    if entry is not None and entry.get("n_energies") != int(n_energies):
        entry = None
    if entry is None or force:
        best = ctypes.c_int32(-1)
        ms = (ctypes.c_float * _lib.RGF_NUM_TILE_CFGS)()
        rc = _lib.load().rgf_autotune(ctx.handle, int(block_size), int(n_energies), int(f32),
                                      ctypes.byref(best), ms)
        _lib.check(rc, ctx.handle)
        entry = {
            "best": int(best.value),
            "best_name": _lib.TILE_CFG_NAMES[best.value],
            "ms": {name: float(ms[k]) for k, name in enumerate(_lib.TILE_CFG_NAMES) if ms[k] >= 0},
            "n_energies": int(n_energies),
        }
        db.setdefault(dev, {})[key] = entry
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(db, indent=1, sort_keys=True))
        except OSError:
            pass
    base = _lib.RGF_OPT_TILE_FOR_BLOCK_F32 if f32 else _lib.RGF_OPT_TILE_FOR_BLOCK
    ctx.set_option(base + int(block_size), entry["best"])
    return entry


def apply(ctx, block_sizes, n_energies: int = 64, precisions=("f64",), path: Path | None = None) -> dict:
    """Tune (or load) and install the tile choice for every distinct block size."""
    out = {}
    for b in sorted({int(x) for x in block_sizes}):
        for p in precisions:
            out[f"{p}:{b}"] = tune(ctx, b, n_energies, f32=(p == "f32"), path=path)
    return out
