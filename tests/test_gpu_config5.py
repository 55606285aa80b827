"""The differentiation workload at its full size: 96^3, box (7.2, 7.6, 8.0)
Rg, f_A = 0.32, chiN = 22, Ns = 200 (ds = 0.005), FP64, seeded fields.
Size-independent checks: the contour-checkpointed sweep (every 20 steps)
gives the stored-slice sweep's Q, densities and box stress, and the adjoint
stress equals the central finite difference (h = 1e-5 Rg per box length) of
-ln Q at fixed fields."""
import numpy as np
import pytest

from polyswift_plusplus_amd import DiblockSCFT, fields

pytestmark = pytest.mark.gpu

SHAPE = (96, 96, 96)
BOX = (7.2, 7.6, 8.0)
CELL = tuple(b / n for b, n in zip(BOX, SHAPE))


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def test_config5_checkpointed_sweep_and_stress():
    wA, wB = fields.seeded_fields(SHAPE, CELL, sigma=0.1, seed=5, seed_kind="lamellar", amplitude=1.0)
    e = DiblockSCFT(SHAPE, CELL, 0.32, 0.005, 22.0)
    c = DiblockSCFT(SHAPE, CELL, 0.32, 0.005, 22.0)
    try:
        assert e.Ns == 200
        e.setFields(wA, wB)
        e.solveQ()
        Q = e.calcBigQ()
        phiA, phiB = e.densities()
        g = e.boxStress()
        print(f"config 5: Q = {Q:.12e}, dF/dL = {g}", flush=True)
        c.setFields(wA, wB)
        c.setCheckpointing(20)
        c.solveQ()
        assert abs(c.calcBigQ() / Q - 1) < 1e-12
        cA, cB = c.densities()
        assert rel(cA, phiA) < 1e-11 and rel(cB, phiB) < 1e-11
        assert rel(c.boxStress(), g) < 1e-9
        h = 1e-5
        for d in range(3):
            lnQ = []
            for sgn in (1, -1):
                cell = list(CELL)
                cell[d] = (BOX[d] + sgn * h) / SHAPE[d]
                e.setCell(cell)
                e.solveQ()
                lnQ.append(np.log(e.calcBigQ()))
            fd = -(lnQ[0] - lnQ[1]) / (2 * h)
            print(f"  axis {d}: adjoint {g[d]:.10e}  finite difference {fd:.10e}", flush=True)
            assert abs(fd - g[d]) < 1e-6 * np.max(np.abs(g)) + 1e-10
        e.setCell(CELL)
    finally:
        e.close()
        c.close()
