"""The oracle's blends (several block copolymers sharing the A and B fields,
pspolymer/PsPolymerHldr.cpp:35-83), pinned by:

* the homogeneous-field closed forms: Q_p = exp(-ds (a sA_p + b sB_p)) per
  polymer, densities from the Simpson 0.33333 quirk weighted by
  volfrac / lengthRatio (pspolymer/PsBlock.cpp:170-173), hand-evaluated;
* an independent numpy propagation (np.fft split steps, hand-built
  quadrature weights, per-polymer Q, densities and F) on random fields;
* linearity: a diblock split into two identical polymers of volfrac 0.3 and
  0.7 is the diblock itself, along a steepest-descent trajectory;
* the box stress of a blend against central finite differences of
  -sum_p volfrac ln Q_p / lengthRatio at fixed fields;
* the species averages sum volfrac * lengthfrac (psbase/PsBlockBase.cpp:185-189)
  and the rejected descriptions."""
import numpy as np
import pytest

import oracle as O
from oracle import Block, Params, Polymer

SHAPE, CELL, DS = (12, 10, 8), (0.3, 0.35, 0.25), 0.05


def blend3():
    """Diblock (vf 0.7, length 100: A 0.4, B 0.6), A homopolymer (vf 0.2,
    length 50) and B homopolymer (vf 0.1, length 150)."""
    return (Polymer(0.7, 1.0, (Block(0, 0.4, 1.0, (), (1,)), Block(1, 0.6, 1.0, (0,), ()))),
            Polymer(0.2, 0.5, (Block(0, 1.0, 1.0, (), ()),)),
            Polymer(0.1, 1.5, (Block(1, 1.0, 1.0, (), ()),)))


def params(polymers, chiN=14.0, **kw):
    return Params(*SHAPE, *CELL, ds=DS, fA=0.5, chiN=chiN, polymers=tuple(polymers), **kw)


def simpson_quirk(M, ds):
    """The reference's contour weights for M points (PsBlock.cpp:256-380):
    0.33333 ds [1, 4, 2, ..., 4, 1] over an odd count, a trapezoid last
    interval when M is even."""
    odd = M if M % 2 == 1 else M - 1
    w = np.zeros(M)
    if odd == 1:
        w[0] = 0.0
    else:
        w[:odd] = 2.0
        w[1:odd - 1:2] = 4.0
        w[0] = w[odd - 1] = 1.0
        w[:odd] *= 0.33333 * ds
    if M % 2 == 0:
        w[M - 2] += 0.5 * ds
        w[M - 1] += 0.5 * ds
    return w


def test_species_fracs_and_steps():
    P = params(blend3())
    fA, fB = O.species_fracs(P)
    assert abs(fA - (0.7 * 0.4 + 0.2)) < 1e-15 and abs(fB - (0.7 * 0.6 + 0.1)) < 1e-15
    assert O.lib().orc_polymer_count(P.c()) == 3
    # slices per block: steps = lengthRatio * lengthfrac / ds -> 8, 12, 10, 30
    assert O.lib().orc_graph_slice_count(P.c()) == 2 * (9 + 13 + 11 + 31)


def test_homogeneous_closed_forms():
    a, b, chiN = 0.7, -0.4, 14.0
    P = params(blend3(), chiN=chiN)
    wA, wB = np.full(SHAPE, a), np.full(SHAPE, b)
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    steps = [(8, 12), (10, 0), (0, 30)]
    want_Q = [np.exp(-DS * (a * sa + b * sb)) for sa, sb in steps]
    assert np.allclose(Q, want_Q, rtol=1e-13, atol=0)
    # Simpson quirk sums: M = 9, 13, 11, 31 points, all odd: 0.33333 * ds * 3 * steps
    c = lambda s: 0.33333 * DS * 3 * s
    want_A = 0.7 / 1.0 * c(8) + 0.2 / 0.5 * c(10)
    want_B = 0.7 / 1.0 * c(12) + 0.1 / 1.5 * c(30)
    assert abs(want_A - 0.479995200) < 1e-12 and abs(want_B - 0.5199948) < 1e-12
    assert np.allclose(phiA, want_A, rtol=1e-13, atol=0) and np.allclose(phiB, want_B, rtol=1e-13, atol=0)
    fA, fB = 0.48, 0.52
    F = O.free_energy(P, wA, wB, phiA, phiB, Q)
    want_F = (-(0.7 * np.log(want_Q[0])) / 1.0 - (0.2 * np.log(want_Q[1])) / 0.5
              - (0.1 * np.log(want_Q[2])) / 1.5
              + (-want_A * a - want_B * b) + chiN * (want_A * want_B - fA * fB))
    assert abs(F - want_F) < 1e-13


def numpy_blend(polymers, wA, wB, chiN, shape=SHAPE, cell=CELL, ds=DS):
    """Independent restatement with numpy FFTs: per-polymer split-step
    propagation, junction hand-off, Q from the first block's head, densities
    weighted by volfrac / lengthRatio / Q_p, and F."""
    k2 = sum(np.meshgrid(*[(2 * np.pi * np.fft.fftfreq(n, d=h)) ** 2 for n, h in zip(shape, cell)],
                         indexing="ij"))
    kf = np.exp(-ds * k2)
    wf = {0: np.exp(-0.5 * ds * wA), 1: np.exp(-0.5 * ds * wB)}

    def sweep(q0, species, steps):
        qs = [q0]
        for _ in range(steps):
            u = wf[species] * qs[-1]
            qs.append(wf[species] * np.real(np.fft.ifftn(kf * np.fft.fftn(u))))
        return qs

    phi = {0: np.zeros(shape), 1: np.zeros(shape)}
    Qs = []
    for p in polymers:
        bl = p.blocks
        if len(bl) == 2:      # A (head free) - B (tail free)
            sA, sB = (round(p.lratio * b.frac / ds) for b in bl)
            qA = sweep(np.ones(shape), bl[0].species, sA)
            qB = sweep(qA[-1], bl[1].species, sB)
            tB = sweep(np.ones(shape), bl[1].species, sB)
            tA = sweep(tB[-1], bl[0].species, sA)
            pairs = [(bl[0], qA, tA), (bl[1], qB, tB)]
            Q = np.mean(tA[-1])
        else:                 # homopolymer: both ends free
            s = round(p.lratio * bl[0].frac / ds)
            q = sweep(np.ones(shape), bl[0].species, s)
            pairs = [(bl[0], q, q)]
            Q = np.mean(q[-1])
        Qs.append(Q)
        for b, H, T in pairs:
            M = len(H)
            w = simpson_quirk(M, ds)
            qqt = sum(w[t] * H[t] * T[M - 1 - t] for t in range(M)) / Q
            phi[b.species] += qqt * (p.volfrac / p.lratio)
    fA = sum(p.volfrac * b.frac for p in polymers for b in p.blocks if b.species == 0)
    fB = sum(p.volfrac * b.frac for p in polymers for b in p.blocks if b.species == 1)
    F = (sum(-(p.volfrac * np.log(Q)) / p.lratio for p, Q in zip(polymers, Qs))
         + np.mean(-phi[0] * wA - phi[1] * wB) + chiN * np.mean(phi[0] * phi[1]) - chiN * fA * fB)
    return phi[0], phi[1], np.array(Qs), F


def test_blend_against_numpy_propagation():
    rng = np.random.default_rng(8)
    wA, wB = 0.6 * rng.standard_normal(SHAPE), 0.6 * rng.standard_normal(SHAPE)
    polys = blend3()
    P = params(polys)
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    nA, nB, nQ, nF = numpy_blend(polys, wA, wB, P.chiN)
    assert np.max(np.abs(Q / nQ - 1)) < 1e-12
    assert np.max(np.abs(phiA - nA)) < 1e-12 and np.max(np.abs(phiB - nB)) < 1e-12
    assert abs(O.free_energy(P, wA, wB, phiA, phiB, Q) - nF) < 1e-12
    # even point counts too (trapezoid tails): steps 7 and 13 at ds = 0.05
    polys2 = (Polymer(0.6, 1.0, (Block(0, 0.35, 1.0, (), (1,)), Block(1, 0.65, 1.0, (0,), ()))),
              Polymer(0.4, 0.7, (Block(1, 1.0, 1.0, (), ()),)))
    P2 = params(polys2)
    phiA, phiB, Q = O.graph_densities(P2, wA, wB)
    nA, nB, nQ, nF = numpy_blend(polys2, wA, wB, P2.chiN)
    assert np.max(np.abs(Q / nQ - 1)) < 1e-12
    assert np.max(np.abs(phiA - nA)) < 1e-12 and np.max(np.abs(phiB - nB)) < 1e-12
    assert abs(O.free_energy(P2, wA, wB, phiA, phiB, Q) - nF) < 1e-12


def test_split_diblock_is_the_diblock():
    rng = np.random.default_rng(9)
    wA, wB = 0.5 * rng.standard_normal(SHAPE), 0.5 * rng.standard_normal(SHAPE)
    di = (Block(0, 0.3, 1.0, (), (1,)), Block(1, 0.7, 1.0, (0,), ()))
    Pd = Params(*SHAPE, *CELL, ds=DS, fA=0.3, chiN=15.0)
    Pb = params((Polymer(0.3, 1.0, di), Polymer(0.7, 1.0, di)), chiN=15.0)
    a, b, Q = O.diblock_densities(Pd, wA, wB)
    c, d, Qb = O.graph_densities(Pb, wA, wB)
    assert np.allclose(Qb, [Q, Q], rtol=1e-15, atol=0)
    assert np.max(np.abs(c - a)) < 1e-14 and np.max(np.abs(d - b)) < 1e-14
    rd = O.scft_run(Pd, wA, wB, n_iters=4)
    rb = O.scft_run(Pb, wA, wB, n_iters=4)
    assert np.max(np.abs(rb["F"] - rd["F"])) < 1e-13
    assert np.max(np.abs(rb["wA"] - rd["wA"])) < 1e-12


def test_blend_stress_against_finite_differences():
    rng = np.random.default_rng(10)
    wA, wB = 0.5 * rng.standard_normal(SHAPE), 0.5 * rng.standard_normal(SHAPE)
    polys = blend3()
    P = params(polys)
    g = O.box_stress(P, wA, wB)
    h = 1e-5
    for d in range(3):
        vals = []
        for sgn in (1, -1):
            cell = list(CELL)
            cell[d] = (SHAPE[d] * CELL[d] + sgn * h) / SHAPE[d]
            Pd = Params(*SHAPE, *cell, ds=DS, fA=0.5, chiN=P.chiN, polymers=polys)
            _, _, Q = O.graph_densities(Pd, wA, wB)
            vals.append(sum(-(p.volfrac * np.log(q)) / p.lratio for p, q in zip(polys, Q)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[d]) < 1e-7 * max(1.0, abs(g[d])), (d, fd, g[d])


def test_bad_blend_descriptions():
    di = (Block(0, 0.4, 1.0, (), (1,)), Block(1, 0.6, 1.0, (0,), ()))
    bad = [(Polymer(0.5, 1.0, di), Polymer(0.5, 1.0, (Block(0, 1.0, 1.0, (1,), ()),))),   # join outside
           (Polymer(0.5, 1.0, di), Polymer(0.5, 0.0, (Block(0, 1.0),))),                   # lengthRatio 0
           (Polymer(0.5, 1.0, di), Polymer(0.5, 0.33, (Block(0, 1.0),)))]                  # steps 6.6
    for polys in bad:
        assert O.lib().orc_graph_slice_count(params(polys).c()) == -1
