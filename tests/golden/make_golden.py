"""Writes the golden fixtures under tests/golden/ from the DENSE oracle.

Inputs come from the seeded workload generator (quatrex_amd.workloads) or the
reference-test-style generator (tests/helpers.py); expected outputs are
numpy.linalg.inv of the full matrix and G^R Sigma G^R^H, selected on the
tridiagonal pattern.  Run:  python tests/golden/make_golden.py
"""

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.dense import dense_solve  # noqa: E402
from quatrex_amd import workloads as W  # noqa: E402
from quatrex_amd.layout import Geometry  # noqa: E402
from tests.helpers import bt_dense, skew_tridiagonal  # noqa: E402

OUT = Path(__file__).resolve().parent


def write(name, g, a, sl, sg, extra=None):
    gr, gl, gg = dense_solve(a, sl, sg)
    data = {"block_sizes": np.asarray(g.block_sizes, np.int64)}
    for key, m in (("a", a), ("sl", sl), ("sg", sg), ("gr", gr), ("gl", gl), ("gg", gg)):
        d, u, l = g.from_dense(m)
        data[f"{key}_diag"], data[f"{key}_upper"], data[f"{key}_lower"] = d, u, l
    # stored diagonal G^<> are (X - X^H)/2 of the exact value (identical up to rounding)
    data.update(extra or {})
    np.savez_compressed(OUT / name, **data)
    print(name, sum(v.nbytes for v in data.values()), "bytes")


def main():
    # 1) generator workload, N=4, b=8, 3 energies, block-diagonal Sigma
    g = Geometry.uniform(4, 8)
    m = W.make_device(g, W.DEVICE_SEED_BASE + 100)
    en = W.draw_energies(3)
    ad, au, al = W.system_matrix(m, en, 1e-3)
    a = g.to_dense(ad, au, al)
    sld, slu, sgd, sgu = W.sigmas(m, 3)
    z = np.zeros_like(slu)
    write("rgf_4x8_generator.npz", g, a, g.to_dense(sld, slu, z), g.to_dense(sgd, sgu, z),
          {"energies": en, "eta": np.float64(1e-3), "seed": np.int64(W.DEVICE_SEED_BASE + 100)})

    # 2) reference-test-style matrix with mixed block sizes and off-diagonal Sigma
    bs = [3, 5, 2, 4]
    g = Geometry(tuple(bs))
    a = bt_dense(bs, seed=31)
    rng = np.random.default_rng(32)
    sl = skew_tridiagonal(bs, rng)
    sg = skew_tridiagonal(bs, rng)
    write("rgf_mixed_3-5-2-4.npz", g, a[None], sl[None], sg[None])

    # 3) the one-energy 8x32 CPU workload: only input checksums and diagonal outputs
    g = Geometry.uniform(8, 32)
    seed = W.DEVICE_SEED_BASE + 1
    m = W.make_device(g, seed)
    en = W.draw_energies(1)
    ad, au, al = W.system_matrix(m, en, 1e-3)
    sld, slu, sgd, sgu = W.sigmas(m, 1)
    z = np.zeros_like(slu)
    gr, gl, _ = dense_solve(g.to_dense(ad, au, al), g.to_dense(sld, slu, z), g.to_dense(sgd, sgu, z))
    np.savez_compressed(
        OUT / "cfg1_8x32.npz",
        seed=np.int64(seed), energies=en, eta=np.float64(1e-3),
        a_diag_sum=ad.sum(), sl_diag_sum=m.sl_diag.sum(),
        gr_diag=g.from_dense(gr)[0], gl_diag=g.from_dense(gl)[0],
    )
    print("cfg1_8x32.npz")


if __name__ == "__main__":
    main()
