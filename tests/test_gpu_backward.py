"""The reverse mode of (w_A, w_B, box) -> (phi_A, phi_B, Q)
(psx_scft_backward, adjoint.DensityOp) against central finite differences of
the oracle's forward evaluation (an independent propagation) and of the
engine's own, and against two closed forms: with g_phi = 0 and g_Q = -1/Q it
is the derivative of -ln Q, whose w part is phi / N under the trapezoid
contour rule and whose box part is the box stress."""
import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import Block, BlockCopolymerSCFT, DiblockSCFT, PsxError, fields
from polyswift_plusplus_amd.adjoint import DensityOp

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = {
    "diblock3d": dict(shape=(12, 10, 8), cell=(0.3, 0.35, 0.25), blocks=None, fA=0.3),
    "diblock2d": dict(shape=(16, 12, 1), cell=(0.25, 0.3, 1.0), blocks=None, fA=0.35),
    "triblock": dict(shape=(10, 8, 6), cell=(0.35, 0.3, 0.4),
                     blocks=[Block("A", 0.25, 1.0, (), (1,)), Block("B", 0.5, 1.2, (0,), (2,)),
                             Block("A", 0.25, 1.0, (1,), ())], fA=0.5),
}


def setup(name, ds=0.05, chiN=15.0):
    c = CASES[name]
    if c["blocks"] is None:
        e = DiblockSCFT(c["shape"], c["cell"], c["fA"], ds, chiN)
        P = O.Params(*c["shape"], *c["cell"], ds=ds, fA=c["fA"], chiN=chiN)
    else:
        e = BlockCopolymerSCFT(c["shape"], c["cell"], c["blocks"], ds, chiN)
        ob = tuple(O.Block(0 if b.species == "A" else 1, b.lengthfrac, b.bSegRatio, b.head, b.tail)
                   for b in c["blocks"])
        P = O.Params(*c["shape"], *c["cell"], ds=ds, fA=0.5, chiN=chiN, blocks=ob)
    wA, wB = fields.seeded_fields(c["shape"], c["cell"], sigma=0.4, seed=61, seed_kind="lamellar")
    return e, P, wA, wB


def oracle_L(P, wA, wB, gA, gB, gQ):
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    return float(np.sum(gA * phiA) + np.sum(gB * phiB) + gQ * Q)


@pytest.mark.parametrize("name", list(CASES))
def test_backward_against_oracle_finite_differences(name):
    print(f"backward {name}", flush=True)
    e, P, wA, wB = setup(name)
    try:
        rng = np.random.default_rng(5)
        gA, gB = rng.standard_normal(wA.shape), rng.standard_normal(wA.shape)
        gQ = 0.7
        op = DensityOp(e)
        op.forward(wA, wB)
        dA, dB, dL = op.backward(gA, gB, gQ)
        # directional derivatives along random directions of w_A and w_B
        h = 1e-5
        for which in (0, 1):
            v = rng.standard_normal(wA.shape)
            pa, pb = (wA + h * v, wB) if which == 0 else (wA, wB + h * v)
            ma, mb = (wA - h * v, wB) if which == 0 else (wA, wB - h * v)
            fd = (oracle_L(P, pa, pb, gA, gB, gQ) - oracle_L(P, ma, mb, gA, gB, gQ)) / (2 * h)
            an = float(np.sum((dA if which == 0 else dB) * v))
            print(f"  dw_{'AB'[which]}: adjoint {an:.12e} finite difference {fd:.12e}", flush=True)
            assert abs(an - fd) < 2e-8 * max(1.0, abs(fd))
        # single cells
        for r in ((0, 0, 0), (3, 2, 0), (5, 7, 0)):
            for which in (0, 1):
                p = [wA.copy(), wB.copy()]
                m = [wA.copy(), wB.copy()]
                p[which][r] += h
                m[which][r] -= h
                fd = (oracle_L(P, *p, gA, gB, gQ) - oracle_L(P, *m, gA, gB, gQ)) / (2 * h)
                an = (dA if which == 0 else dB)[r]
                assert abs(an - fd) < 1e-8 * max(1.0, np.max(np.abs(dA))), (r, which, an, fd)
        # box lengths
        box = op.box()
        for d in range(3):
            if P.shape[d] == 1:
                assert dL[d] == 0.0
                continue
            vals = []
            for sgn in (1, -1):
                cell = [P.dx, P.dy, P.dz]
                cell[d] = (box[d] + sgn * h) / P.shape[d]
                Pd = O.Params(*P.shape, *cell, ds=P.ds, fA=P.fA, chiN=P.chiN, blocks=P.blocks)
                vals.append(oracle_L(Pd, wA, wB, gA, gB, gQ))
            fd = (vals[0] - vals[1]) / (2 * h)
            print(f"  dL_{'xyz'[d]}: adjoint {dL[d]:.12e} finite difference {fd:.12e}", flush=True)
            assert abs(dL[d] - fd) < 2e-8 * max(1.0, abs(fd))
    finally:
        e.close()


def test_backward_of_log_Q_closed_forms():
    """g_phi = 0, g_Q = -1/Q: dL/dw is the derivative of -ln Q, equal to
    phi / N under the trapezoid contour rule, and dL/dL_box is the box stress."""
    print("backward: -ln Q", flush=True)
    e, P, wA, wB = setup("diblock3d")
    try:
        e.setQuadrature("trapezoid")
        e.setFields(wA, wB)
        e.solveQ()
        Q = e.calcBigQ()
        phiA, phiB = e.densities()
        dA, dB, dL = e.backward(None, None, -1.0 / Q)
        N = float(np.prod(P.shape))
        assert rel(dA, phiA / N) < 1e-10 and rel(dB, phiB / N) < 1e-10
        assert rel(dL, e.boxStress()) < 1e-10
        # the reference quadrature changes only the densities' weights, not Q
        e.setQuadrature("reference")
        e.solveQ()
        dA2, dB2, dL2 = e.backward(None, None, -1.0 / Q)
        assert rel(dA2, dA) < 1e-12 and rel(dL2, dL) < 1e-12
    finally:
        e.close()


def test_backward_refusals_and_device_entry():
    e, P, wA, wB = setup("diblock3d")
    try:
        with pytest.raises(PsxError):
            e.backward(np.zeros(P.shape))            # before propagate
        e.setFields(wA, wB)
        e.solveQ()
        e.setCheckpointing(5)
        e.solveQ()
        with pytest.raises(PsxError):
            e.backward(np.zeros(P.shape))
        e.setCheckpointing(0)
        e.setPrecision("fp32")
        e.solveQ()
        with pytest.raises(PsxError):
            e.backward(np.zeros(P.shape))
    finally:
        e.close()
    star = BlockCopolymerSCFT((8, 8, 8), (0.3,) * 3,
                              [Block("A", 0.3, 1.0, (), (1, 2)), Block("B", 0.4, 1.0, (), (0, 2)),
                               Block("A", 0.3, 1.0, (0, 1), ())], 0.05, 12.0)
    try:
        star.setFields(*fields.seeded_fields((8, 8, 8), (0.3,) * 3, sigma=0.3, seed=62))
        star.solveQ()
        with pytest.raises(PsxError):
            star.backward(np.zeros((8, 8, 8)))
    finally:
        star.close()
