"""Write the golden fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

The reference ships no vectors for this path (its regression results live
in an external psresults/ tree that is not part of it), so the fixtures are
the oracle's outputs on fixed, seeded inputs.  They pin the oracle against
silent change (tests/test_golden.py, CPU) and give the GPU tests a stored
answer that does not need the oracle at run time.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from polyswift_plusplus_amd import fields  # noqa: E402

CASES = {
    # The reference's own CPU case: 64x64 2D melt, box 4.0 Rg, fA 0.30,
    # chiN 20, Ns 100, steepest descent [0.1, 0] (Picard), 200 iterations,
    # seeded N(0, 0.1) fields.
    "c1_2d_picard": dict(shape=(64, 64, 1), box=(4.0, 4.0, 4.0 / 64), fA=0.30, chiN=20.0, ds=0.01,
                         sigma=0.1, seed=101, seed_kind=None, mix=0, lambda0=0.1, lambda1=0.0,
                         depth=0, beta=0.0, start=1e300, iters=200),
    # small 3D lamellar-seeded run with Anderson mixing (depth 4)
    "small_3d_anderson": dict(shape=(12, 12, 12), box=(6.0, 6.0, 6.0), fA=0.35, chiN=18.0, ds=0.05,
                              sigma=0.1, seed=202, seed_kind="lamellar", mix=1, lambda0=0.1,
                              lambda1=0.0, depth=4, beta=0.5, start=0.5, iters=15),
    # steepest descent with both relaxation parameters and an even point count
    "small_3d_sd2": dict(shape=(8, 6, 10), box=(3.2, 2.4, 4.0), fA=0.45, chiN=14.0, ds=0.05,
                         sigma=0.3, seed=303, seed_kind="lamellar", mix=0, lambda0=0.15,
                         lambda1=0.05, depth=0, beta=0.0, start=1e300, iters=8),
}


def run_case(name, c):
    shape = c["shape"]
    cell = tuple(b / n for b, n in zip(c["box"], shape))
    wA, wB = fields.seeded_fields(shape, cell, sigma=c["sigma"], seed=c["seed"],
                                  seed_kind=c["seed_kind"])
    P = O.Params(*shape, *cell, ds=c["ds"], fA=c["fA"], chiN=c["chiN"], mix=c["mix"],
                 lambda0=c["lambda0"], lambda1=c["lambda1"], depth=c["depth"], beta=c["beta"],
                 anderson_start=c["start"])
    r = O.scft_run(P, wA, wB, n_iters=c["iters"])
    return dict(shape=np.array(shape), cell=np.array(cell), fA=c["fA"], chiN=c["chiN"], ds=c["ds"],
                mix=c["mix"], lambda0=c["lambda0"], lambda1=c["lambda1"], depth=c["depth"],
                beta=c["beta"], start=c["start"], iters=c["iters"],
                wA0=wA, wB0=wB, F=r["F"], Q=r["Q"], err=r["err"], wA=r["wA"], wB=r["wB"],
                p=r["p"], phiA=r["phiA"], phiB=r["phiB"])


def fft_pair_case():
    rng = np.random.default_rng(404)
    d = rng.standard_normal((6, 10, 8))
    k = rng.random((6, 10, 8))
    return dict(data=d, kdata=k, res=O.scaled_fft_pair(d, k))


def main():
    O.set_threads(min(8, os.cpu_count() or 1))
    for name, c in CASES.items():
        out = run_case(name, c)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, "F[-1] =", out["F"][-1])
    np.savez_compressed(os.path.join(HERE, "fft_pair_6x10x8.npz"), **fft_pair_case())


if __name__ == "__main__":
    main()
