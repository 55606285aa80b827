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


def test_config5_box_relaxation_to_zero_stress():
    """The box relaxation at full size (checkpoints every 20 contour steps,
    trapezoid contour rule) to |dF/dL| < 1e-7 from seeded N(0, 0.1) fields
    plus the hexagonal-cylinder seed.  Size-independent check: the cylinders
    run along z, so dF/dL_z vanishes and L_z stays, and the relaxed x-y cell
    is the regular hexagonal one for the seed's wavevectors (2, 0), (-1, +-2):
    |k| equal for all three gives L_y / L_x = 2 / sqrt(3)."""
    import bench
    print("config 5 box relaxation", flush=True)
    r = bench.config5_relax_leg()
    h = r["history"]
    print(f"  {r['outer_steps']} outer steps, L = {r['L']}, max|dF/dL| = {r['max_abs_dFdL']:.2e}, "
          f"F {h[0]['F']:.10f} -> {h[-1]['F']:.10f}", flush=True)
    assert r["converged"] and r["max_abs_dFdL"] < 1e-7
    assert all(step["scft_err"] < 1e-10 for step in h)
    assert abs(r["L"][2] - 8.0) < 1e-12
    assert abs(r["L"][1] / r["L"][0] - 2.0 / np.sqrt(3.0)) < 1e-4
    assert h[-1]["F"] < h[0]["F"] - 1e-3
