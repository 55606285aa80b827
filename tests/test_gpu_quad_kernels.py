"""The two contour-quadrature kernels: on small FP32 grids the default
splits each cell's contour sum over a block's four waves (k_quad_split); the
serial kernel (k_quad_update) sums every cell in slice order.  They sum the
same FP64 terms in different orders, so densities and the checkpointed chunk
sums (raw sums carried between chunks) agree to FP64 rounding; steepest-
descent and Anderson trajectories agree until an FP32 w-factor rounds
differently (far below the FP32 sweep's own error)."""
import numpy as np
import pytest

from polyswift_plusplus_amd import DiblockSCFT, Mixing, fields

pytestmark = pytest.mark.gpu

SHAPE, CELL = (32, 32, 32), (0.125,) * 3


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(ckpt=0):
    wA, wB = fields.seeded_fields(SHAPE, CELL, sigma=0.2, seed=31, seed_kind="lamellar")
    out = []
    for kind in ("auto", "serial"):
        e = DiblockSCFT(SHAPE, CELL, 0.3, 0.01, 16.0, prec="fp32")
        e.setQuadKernel(kind)
        if ckpt:
            e.setCheckpointing(ckpt)
        e.setFields(wA, wB)
        out.append(e)
    return out


def test_split_and_serial_densities():
    print("quad kernels: densities", flush=True)
    a, b = pair()
    try:
        for e in (a, b):
            e.solveQ()
        pa, pb = a.densities(), b.densities()
        assert rel(pa[0], pb[0]) < 1e-14 and rel(pa[1], pb[1]) < 1e-14
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("mix", [Mixing("steepest", 0.1, 0.05),
                                 Mixing("anderson", lambda0=0.1, depth=4, beta=0.5)],
                         ids=["steepest", "anderson"])
def test_split_and_serial_trajectories(mix):
    print(f"quad kernels: {mix.kind} trajectory", flush=True)
    a, b = pair()
    try:
        # the first iteration starts from identical fields; later ones from
        # fields 1e-15 apart, whose FP32 w-factors can round differently
        ha, hb = a.iterate(6, mix), b.iterate(6, mix)
        assert abs(ha["F"][0] - hb["F"][0]) < 1e-13
        assert rel(ha["F"], hb["F"]) < 1e-8
        assert rel(ha["err"], hb["err"]) < 1e-8
        for x, y in zip(a.getFields(), b.getFields()):
            assert rel(x, y) < 1e-8
    finally:
        a.close()
        b.close()


def test_split_and_serial_checkpointed_chunks():
    print("quad kernels: checkpointed chunk sums", flush=True)
    a, b = pair(ckpt=7)
    try:
        for e in (a, b):
            e.solveQ()
        pa, pb = a.densities(), b.densities()
        assert rel(pa[0], pb[0]) < 1e-14 and rel(pa[1], pb[1]) < 1e-14
        ha, hb = a.iterate(3, Mixing("anderson", lambda0=0.1, depth=3, beta=0.5)), \
            b.iterate(3, Mixing("anderson", lambda0=0.1, depth=3, beta=0.5))
        assert rel(ha["F"], hb["F"]) < 1e-8
    finally:
        a.close()
        b.close()
