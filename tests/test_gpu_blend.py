"""Blends on the GPU (several block copolymers sharing the A and B fields):
per-polymer Q, densities weighted by volfrac / lengthRatio, the species
averages, F and the box stress against the oracle (FP64 to 1e-10), in line
and plane sweeps, steepest-descent and Anderson trajectories, FP32 within
2e-5, and the reference-style blend deck (a diblock with an A homopolymer of
half its length)."""
import os

import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import BlendSCFT, Block, Mixing, Polymer, PsxError, deck, fields

pytestmark = pytest.mark.gpu

RTOL64 = 1e-10
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def blend3():
    """Engine and oracle descriptions of a diblock (vf 0.7, length 100),
    an A homopolymer (vf 0.2, length 50) and a B homopolymer (vf 0.1,
    length 150)."""
    eng = [Polymer(0.7, 100.0, (Block("A", 0.4, 1.0, (), (1,)), Block("B", 0.6, 1.0, (0,), ()))),
           Polymer(0.2, 50.0, (Block("A", 1.0),)),
           Polymer(0.1, 150.0, (Block("B", 1.0, 1.3),))]
    orc = (O.Polymer(0.7, 1.0, (O.Block(0, 0.4, 1.0, (), (1,)), O.Block(1, 0.6, 1.0, (0,), ()))),
           O.Polymer(0.2, 0.5, (O.Block(0, 1.0, 1.0, (), ()),)),
           O.Polymer(0.1, 1.5, (O.Block(1, 1.0, 1.3, (), ()),)))
    return eng, orc


def case(shape=(12, 10, 8), cell=(0.3, 0.35, 0.25), ds=0.05, chiN=14.0, seed=21, **kw):
    eng, orc = blend3()
    P = O.Params(*shape, *cell, ds=ds, fA=0.5, chiN=chiN, polymers=orc, **kw)
    wA, wB = fields.seeded_fields(shape, cell, sigma=0.5, seed=seed, seed_kind="lamellar")
    return eng, P, wA, wB


@pytest.mark.parametrize("shape,cell", [((12, 10, 8), (0.3, 0.35, 0.25)),
                                        ((4, 32, 32), (0.3, 0.2, 0.2))], ids=["lines", "planes"])
def test_blend_densities_Q_F_stress(shape, cell):
    print(f"blend densities {shape}", flush=True)
    eng, P, wA, wB = case(shape, cell)
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    e = BlendSCFT(shape, cell, eng, P.ds, P.chiN)
    try:
        assert e.graphInfo()[1] == 6   # two paths per polymer
        assert e.sweepMode()[0] == ("planes" if shape[1] == 32 else "lines")
        e.setFields(wA, wB)
        e.solveQ()
        gQ = e.polymerQ()
        assert gQ.shape == (3,) and np.max(np.abs(gQ / Q - 1)) < RTOL64
        assert abs(e.calcBigQ() / Q[0] - 1) < RTOL64
        gA, gB = e.densities()
        assert rel(gA, phiA) < RTOL64 and rel(gB, phiB) < RTOL64
        F = O.free_energy(P, wA, wB, phiA, phiB, Q)
        assert abs(e.getFreeE() - F) < RTOL64 * abs(F)
        g = e.boxStress()
        want = O.box_stress(P, wA, wB)
        print(f"  stress {g} oracle {want}", flush=True)
        assert np.max(np.abs(g - want)) < RTOL64 * np.max(np.abs(want))
    finally:
        e.close()


@pytest.mark.parametrize("mix,omix", [(Mixing("steepest", 0.1, 0.05), dict(mix=0, lambda0=0.1, lambda1=0.05)),
                                      (Mixing("anderson", lambda0=0.1, depth=4, beta=0.3),
                                       dict(mix=1, lambda0=0.1, depth=4, beta=0.3))],
                         ids=["steepest", "anderson"])
def test_blend_trajectories(mix, omix):
    print(f"blend trajectory {mix.kind}", flush=True)
    eng, P, wA, wB = case(**omix)
    ref = O.scft_run(P, wA, wB, n_iters=8)
    e = BlendSCFT(P.shape, (P.dx, P.dy, P.dz), eng, P.ds, P.chiN)
    try:
        e.setFields(wA, wB)
        h = e.iterate(8, mix, use_graph=True)
        assert rel(h["F"], ref["F"]) < 1e-9 and rel(h["Q"], ref["Q"]) < RTOL64
        assert rel(h["err"], ref["err"]) < 1e-9
        gA, gB, gp = e.getFields()
        assert rel(gA, ref["wA"]) < 1e-9 and rel(gB, ref["wB"]) < 1e-9
    finally:
        e.close()


def test_blend_fp32_and_homogeneous():
    eng, P, wA, wB = case()
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    e = BlendSCFT(P.shape, (P.dx, P.dy, P.dz), eng, P.ds, P.chiN, prec="fp32")
    try:
        e.setFields(wA, wB)
        e.solveQ()
        assert np.max(np.abs(e.polymerQ() / Q - 1)) < 2e-5
        gA, gB = e.densities()
        assert rel(gA, phiA) < 2e-5 and rel(gB, phiB) < 2e-5
        # homogeneous fields: the closed forms of the oracle's pins
        e.setPrecision("fp64")
        a, b = 0.7, -0.4
        e.setFields(np.full(P.shape, a), np.full(P.shape, b))
        e.solveQ()
        want = [np.exp(-P.ds * (a * 8 + b * 12)), np.exp(-P.ds * a * 10), np.exp(-P.ds * b * 30)]
        assert np.max(np.abs(e.polymerQ() / want - 1)) < 1e-13
        gA, gB = e.densities()
        assert np.max(np.abs(gA - 0.4799952)) < 1e-12 and np.max(np.abs(gB - 0.5199948)) < 1e-12
    finally:
        e.close()


def test_blend_rejections():
    eng, P, wA, wB = case()
    e = BlendSCFT(P.shape, (P.dx, P.dy, P.dz), eng, P.ds, P.chiN)
    try:
        with pytest.raises(PsxError):
            e.setCheckpointing(5)
    finally:
        e.close()
    with pytest.raises(PsxError) as ex:     # lengthRatio 0.33 * 1.0 / 0.05 = 6.6 steps
        BlendSCFT(P.shape, (P.dx, P.dy, P.dz), [eng[0], Polymer(0.3, 33.0, (Block("A", 1.0),))],
                  P.ds, P.chiN)
    assert ex.value.status == 3
    with pytest.raises(PsxError):           # a join outside the polymer
        BlendSCFT(P.shape, (P.dx, P.dy, P.dz), [eng[0], Polymer(0.3, 50.0, (Block("A", 1.0, 1.0, (1,)),))],
                  P.ds, P.chiN)


def test_blend_deck_matches_oracle():
    """The reference-style blend deck: a diblock (volfrac 0.9) and an A
    homopolymer of half its length (volfrac 0.1), steepest descent from the
    reference's random initial fields, against the oracle step by step."""
    path = os.path.join(HERE, "golden", "decks", "ab_homop_mix.pre")
    cfg = deck.configure(deck.load(path))
    print(f"blend deck {cfg.shape}", flush=True)
    h = deck.run(path)
    g = O.Rng(cfg.seed)
    n = int(np.prod(cfg.shape))
    drawn = {name: (1.0 * (g.uniform(n) + (-0.5))).reshape(cfg.shape) for name in cfg.init_fields}
    wA, wB = drawn[cfg.species_fields[0]], drawn[cfg.species_fields[1]]
    polys = tuple(O.Polymer(p.volfrac, p.length / cfg.polymers[0].length,
                            tuple(O.Block(0 if b.species == "A" else 1, b.lengthfrac, b.bSegRatio,
                                          b.head, b.tail) for b in p.blocks))
                  for p in cfg.polymers)
    P = O.Params(*cfg.shape, *cfg.cell, cfg.ds, 0.5, cfg.chiN, lambda0=cfg.lambdas[0],
                 lambda1=cfg.lambdas[1], polymers=polys)
    ref = O.scft_run(P, wA, wB, n_iters=cfg.nsteps)
    print(f"  F {np.asarray(h['F'])[-3:]} oracle {ref['F'][-3:]}", flush=True)
    assert h["iters"] == cfg.nsteps == 10
    assert rel(h["F"], ref["F"]) < RTOL64 and rel(h["Q"], ref["Q"]) < RTOL64
