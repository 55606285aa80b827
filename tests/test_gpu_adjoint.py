"""Differentiation: box stress dF/dL (adjoint of the split-step sweep) and
dF/d(chiN, fA), against the oracle at fixed fields and against central
finite differences of the converged free energy.  The reference has no
analytic stress to compare with."""
import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import DiblockSCFT, fields
from polyswift_plusplus_amd.relax import converge, relax_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,prec,tol", [((8, 6, 4), "fp64", 1e-10), ((16, 12, 10), "fp64", 1e-10),
                                            ((32, 16, 8), "fp64", 1e-10), ((16, 12, 10), "fp32", 2e-4)])
def test_stress_matches_oracle(shape, prec, tol):
    print(f"stress {shape} {prec}", flush=True)
    cell = (0.45, 0.4, 0.5)
    P = O.Params(*shape, *cell, ds=0.05, fA=0.3, chiN=12.0)
    rng = np.random.default_rng(21)
    wA, wB = rng.normal(0, 0.8, shape), rng.normal(0, 0.8, shape)
    ref = O.box_stress(P, wA, wB)
    e = DiblockSCFT(shape, cell, 0.3, 0.05, 12.0, prec=prec)
    e.setFields(wA, wB)
    e.solveQ()
    g = e.boxStress()
    assert np.max(np.abs(g - ref)) < tol * np.max(np.abs(ref))


def test_param_grad_fixed_fields():
    shape, cell = (16, 12, 10), (0.45, 0.4, 0.5)
    P = O.Params(*shape, *cell, ds=0.05, fA=0.3, chiN=12.0)
    rng = np.random.default_rng(22)
    wA, wB = rng.normal(0, 0.8, shape), rng.normal(0, 0.8, shape)
    phiA, phiB, Q, (qA, qB, qtB, qtA) = O.diblock_densities(P, wA, wB, want_slices=True)
    e = DiblockSCFT(shape, cell, 0.3, 0.05, 12.0)
    e.setFields(wA, wB)
    e.update()
    gx, gf = e.paramGrad()
    assert abs(gx - (np.mean(phiA * phiB) - 0.3 * 0.7)) < 1e-12
    want = -np.mean(qA[-1] * qtB[-1] * (wB - wA)) / Q
    assert abs(gf - want) < 1e-10 * abs(want)


def _lamellar(shape=(32, 1, 1), L=4.0, fA=0.3, chiN=16.0, ds=0.01, quad="trapezoid"):
    cell = (L / shape[0], 1.0, 1.0)
    wA, wB = fields.seeded_fields(shape, cell, sigma=0.05, seed=5, seed_kind="lamellar", amplitude=2.0)
    e = DiblockSCFT(shape, cell, fA, ds, chiN, quadrature=quad)
    e.setFields(wA, wB)
    return e, wA, wB


@pytest.mark.parametrize("quad,tol", [("trapezoid", 1e-6), ("reference", 2e-3)])
def test_saddle_stress_finite_difference(quad, tol):
    # With the split-step-consistent quadrature F is stationary at the fixed
    # point and the adjoint stress is the exact total derivative.  With the
    # reference's Simpson(0.33333) densities the fixed point is not exactly
    # stationary, so the two differ at O(1e-5 chiN).
    e, wA, wB = _lamellar(quad=quad)
    converge(e, tol=1e-11)
    g = e.boxStress()[0]
    F0 = e.getFreeE()
    h = 1e-4
    Fs = []
    for sgn in (+1, -1):
        e2, _, _ = _lamellar(L=4.0 + sgn * h, quad=quad)
        converge(e2, tol=1e-11)
        Fs.append(e2.getFreeE())
    fd = (Fs[0] - Fs[1]) / (2 * h)
    print(f"{quad}: dF/dL adjoint {g:.10e} fd {fd:.10e} F {F0}", flush=True)
    assert abs(F0) > 1e-3            # an ordered (lamellar) saddle point
    assert abs(g - fd) < tol * max(1.0, abs(fd)) + 1e-9


def test_trapezoid_density_is_dlnQ_dw():
    # phi(r) = -N d(ln Q)/dw(r) / ... : with trapezoid weights the density is
    # the exact derivative of the discrete Q (central difference in one cell).
    shape, cell = (8, 6, 4), (0.45, 0.4, 0.5)
    rng = np.random.default_rng(23)
    wA, wB = rng.normal(0, 0.5, shape), rng.normal(0, 0.5, shape)
    e = DiblockSCFT(shape, cell, 0.3, 0.05, 12.0, quadrature="trapezoid")
    e.setFields(wA, wB)
    phiA, _ = e.update()
    N = np.prod(shape)
    h = 1e-5
    for idx in [(0, 0, 0), (3, 2, 1), (7, 5, 3)]:
        Qs = []
        for sgn in (+1, -1):
            w = wA.copy()
            w[idx] += sgn * h
            e.setFields(w, wB)
            e.solveQ()
            Qs.append(e.calcBigQ())
        dlnQ = (np.log(Qs[0]) - np.log(Qs[1])) / (2 * h)
        assert abs(-N * dlnQ - phiA[idx]) < 1e-8 * abs(phiA[idx])


def test_saddle_chiN_finite_difference():
    e, wA, wB = _lamellar()
    converge(e, tol=1e-11)
    gx, _ = e.paramGrad()
    h = 1e-4
    Fs = []
    for sgn in (+1, -1):
        e2, _, _ = _lamellar(chiN=16.0 + sgn * h)
        converge(e2, tol=1e-11)
        Fs.append(e2.getFreeE())
    fd = (Fs[0] - Fs[1]) / (2 * h)
    print(f"dF/dchiN {gx:.10e} fd {fd:.10e}", flush=True)
    assert abs(gx - fd) < 1e-6 * max(1.0, abs(fd))


def test_saddle_fA_finite_difference():
    # fA moves in whole contour steps: central difference over +-ds at a fine
    # ds; the difference converges to the derivative as O(ds^2).
    ds = 0.0025
    e, _, _ = _lamellar(fA=0.3, ds=ds)
    converge(e, tol=1e-11)
    _, gf = e.paramGrad()
    Fs = []
    for sgn in (+1, -1):
        e2, _, _ = _lamellar(fA=0.3 + sgn * ds, ds=ds)
        converge(e2, tol=1e-11)
        Fs.append(e2.getFreeE())
    fd = (Fs[0] - Fs[1]) / (2 * ds)
    print(f"dF/dfA {gf:.8e} fd {fd:.8e} F+ {Fs[0]:.12f} F- {Fs[1]:.12f}", flush=True)
    assert abs(gf - fd) < 5e-3 * abs(fd)


def test_box_relaxation_lamellar():
    e, _, _ = _lamellar(L=3.6, quad="trapezoid")
    r = relax_box(e, tol_stress=1e-7)
    print("relax:", [(round(h["L"][0], 8), h["dFdL"][0]) for h in r["history"]], flush=True)
    assert r["converged"]
    Fs = [h["F"] for h in r["history"]]
    assert Fs[-1] <= Fs[0] + 1e-12
