"""bench.py's contract pieces that need no GPU: the one-GPU refusal, the
per-energy work it divides by, the arguments it forwards to its child legs,
and the committed HBM-traffic file it reports."""

import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 2
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


@pytest.mark.parametrize("n_blocks,b,flops", [(8, 32, 6.37e7), (32, 256, 1.421e11),
                                              (48, 256, 2.15e11), (16, 128, 8.64e9),
                                              (24, 384, 3.57e11), (12, 512, 4.07e11),
                                              (64, 512, 2.31e12)])
def test_flops_per_energy_matches_the_op_count(n_blocks, b, flops):
    """8 b^3 (34 N - 29): per energy one inverse and 4 products for block 0,
    12 products and one inverse per forward step, 21 products per backward
    step, each 8 b^3 real flops; the published values are rounded to three
    digits."""
    f = bench.flops_per_energy(n_blocks, b)
    gemms = 4 + 12 * (n_blocks - 1) + 21 * (n_blocks - 1)
    invs = n_blocks
    assert f == 8.0 * b**3 * (gemms + invs)
    assert abs(f - flops) / flops < 6e-3


def test_leg_arguments_exist():
    saved = sys.argv
    sys.argv = ["bench.py"]
    try:
        args = bench.parse()
    finally:
        sys.argv = saved
    for name in bench.LEG_ARGS:
        assert hasattr(args, name), name
    assert args.gpus == 1 and args.energies == 512 and args.n_blocks == 32 and args.block_size == 256


def test_traffic_file_has_the_fields_bench_reads():
    t = json.loads((ROOT / "profiles" / "zgemm_traffic_32x256.json").read_text())
    assert t["kernel"].startswith("zgemm_grouped_kernel<64, 64, 0")
    assert t["hbm_bytes_per_launch"] > t["algorithmic_operand_bytes_per_launch"] * 0.5
    cal = t["copy_calibration"]
    assert cal["fetch_scale"] > 0 and cal["write_scale"] > 0
