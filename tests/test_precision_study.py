"""The precision the lesser recursion needs (CPU): complex64 products cannot
reach 1e-8 per block at any of the configured eta, while products of
operands truncated to 44 bits can (the input size an integer-emulated FP64
product would need).  Against exact FP64 products of the same recurrences.

The closed form pins the harness: with exact products it is the FP64 solve,
so its error against itself is 0, and 24-bit truncation (an FP32 mantissa)
lands at the FP32 product's error level."""

import numpy as np
import pytest

from oracle import precision as PR
from quatrex_amd import workloads as W
from quatrex_amd.layout import Geometry


def _case(eta):
    g = Geometry.uniform(16, 128)
    model = W.make_device(g, W.DEVICE_SEED_BASE + 4)
    en = W.draw_energies(1)
    ad, au, al = (x[0] for x in W.system_matrix(model, en, eta))
    sld, slu, _, _ = (x[0] for x in W.sigmas(model, 1))
    blk = lambda f, x: [f(x, i) for i in range(g.n_blocks if f is g.diag else g.n_blocks - 1)]  # noqa: E731
    return (blk(g.diag, ad), blk(g.upper, au), blk(g.lower, al), blk(g.diag, sld), blk(g.upper, slu))


@pytest.mark.parametrize("eta", [1e-2, 1e-3, 1e-4])
def test_fp32_products_miss_1e8_and_44_bits_meet_it(eta):
    a_d, a_u, a_l, s_d, s_u = _case(eta)
    run = lambda mm: PR.solve_lesser(a_d, a_u, a_l, s_d, s_u, mm)  # noqa: E731
    ref = run(PR.exact_product)
    assert PR.max_block_error(run(PR.exact_product), ref) == 0.0
    e32 = PR.max_block_error(run(PR.fp32_product), ref)
    e24 = PR.max_block_error(run(PR.truncated_product(24)), ref)
    e44 = PR.max_block_error(run(PR.truncated_product(44)), ref)
    print(f"eta={eta:g}: fp32 {e32:.1e}, 24-bit {e24:.1e}, 44-bit {e44:.1e}")
    assert e32 > 1e-6
    assert 0.1 < e24 / e32 < 30
    assert e44 < 1e-9
