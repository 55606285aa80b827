"""GPU parity of general block graphs (linear multiblocks, stars) against the
oracle's restatement of PsBlockCopolymer::update (FP64 to 1e-10 rel).

Architectures follow the reference's own regression decks:
pstests/tests/diblock/triblock.pre (A-B-A), star3ab.pre (arms A and B joined
at their tails to the head of a second A arm), asymmetric.pre (bSegRatio).
Every call goes through the C ABI of libpsx_scft.so.
"""
import os

import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import Block, BlockCopolymerSCFT, DiblockSCFT, Mixing, PsxError, fields

pytestmark = pytest.mark.gpu

RTOL64 = 1e-10
SHAPE = (16, 12, 10)
CELL = (0.3, 0.25, 0.35)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def triblock(b2=1.0):
    return [Block("A", 0.25, 1.0, (), (1,)), Block("B", 0.5, b2, (0,), (2,)),
            Block("A", 0.25, 1.0, (1,), ())]


def star3():
    return [Block("A", 0.25, 1.0, (), (1, 2)), Block("B", 0.5, 1.0, (), (0, 2)),
            Block("A", 0.25, 1.0, (0, 1), ())]


def star3_centre_first():
    # the same star with the centre block first: block 0's head is a
    # junction, so Q = mean(qHeadInitial qHeadFinal) with both factors != 1
    return [Block("A", 0.25, 1.0, (1, 2), ()), Block("A", 0.25, 1.0, (), (0, 2)),
            Block("B", 0.5, 1.0, (), (0, 1))]


def star4():
    # four arms joined at one junction: two A arms (tails), a B arm (tail),
    # a B arm (head)
    return [Block("A", 0.2, 1.0, (), (1, 2, 3)), Block("A", 0.2, 1.0, (), (0, 2, 3)),
            Block("B", 0.3, 1.0, (), (0, 1, 3)), Block("B", 0.3, 1.2, (0, 1, 2), ())]


ARCHS = {"triblock": triblock, "triblock_b": lambda: triblock(1.4), "star3": star3,
         "star3_centre_first": star3_centre_first, "star4": star4}


def oracle_blocks(blocks):
    return tuple(O.Block(0 if b.species == "A" else 1, b.lengthfrac, b.bSegRatio, b.head, b.tail)
                 for b in blocks)


def setup(arch, ds=0.05, chiN=14.0, seed=5, prec="fp64", **kw):
    blocks = ARCHS[arch]()
    P = O.Params(*SHAPE, *CELL, ds, 0.5, chiN, blocks=oracle_blocks(blocks), **kw)
    wA, wB = fields.seeded_fields(SHAPE, CELL, sigma=0.5, seed=seed, seed_kind="lamellar")
    e = BlockCopolymerSCFT(SHAPE, CELL, blocks, ds, chiN, prec=prec)
    e.setFields(wA, wB)
    return P, e, wA, wB


@pytest.mark.parametrize("arch", list(ARCHS))
def test_block_slices_Q_densities(arch):
    print(f"graph slices {arch}", flush=True)
    P, e, wA, wB = setup(arch)
    phiA, phiB, Q, sl = O.graph_densities(P, wA, wB, want_slices=True)
    e.solveQ()
    for b, (qh, qt) in enumerate(sl):
        for s in range(qh.shape[0]):
            assert rel(e.blockSlice(b, False, s), qh[s]) < 1e-12, (b, s)
            assert rel(e.blockSlice(b, True, s), qt[s]) < 1e-12, (b, s)
    assert abs(e.calcBigQ() / Q - 1) < 1e-12
    gA, gB = e.densities()
    assert rel(gA, phiA) < RTOL64 and rel(gB, phiB) < RTOL64


def test_graph_structure():
    _, e, _, _ = setup("triblock")
    assert e.graphInfo() == (3, 2, 1)            # two chain sweeps, one round
    _, e, _, _ = setup("star3")
    assert e.graphInfo() == (3, 6, 2)            # three free ends, then three junction starts
    _, e, _, _ = setup("star4")
    assert e.graphInfo() == (4, 8, 2)


def test_diblock_as_block_graph_is_bitwise_the_diblock():
    P = O.Params(*SHAPE, *CELL, 0.05, 0.3, 14.0)
    wA, wB = fields.seeded_fields(SHAPE, CELL, sigma=0.5, seed=5, seed_kind="lamellar")
    d = DiblockSCFT(SHAPE, CELL, 0.3, 0.05, 14.0)
    g = BlockCopolymerSCFT(SHAPE, CELL, [Block("A", 0.3, 1.0, (), (1,)), Block("B", 0.7, 1.0, (0,), ())],
                           0.05, 14.0)
    for e in (d, g):
        e.setFields(wA, wB)
    hd = d.iterate(5, Mixing("steepest", 0.1, 0.05), use_graph=True)
    hg = g.iterate(5, Mixing("steepest", 0.1, 0.05), use_graph=True)
    assert np.array_equal(hd["F"], hg["F"]) and np.array_equal(hd["Q"], hg["Q"])
    assert all(np.array_equal(a, b) for a, b in zip(d.getFields(), g.getFields()))


@pytest.mark.parametrize("arch", ["triblock_b", "star3", "star3_centre_first"])
def test_steepest_descent_trajectory(arch):
    print(f"graph SD {arch}", flush=True)
    P, e, wA, wB = setup(arch)
    P.lambda0, P.lambda1 = 0.1, 0.05
    ref = O.scft_run(P, wA, wB, n_iters=10)
    h = e.iterate(10, Mixing("steepest", 0.1, 0.05), use_graph=False)
    assert rel(h["F"], ref["F"]) < RTOL64
    assert rel(h["Q"], ref["Q"]) < RTOL64
    assert rel(h["err"], ref["err"]) < 1e-9
    gA, gB, gp = e.getFields()
    assert rel(gA, ref["wA"]) < RTOL64 and rel(gB, ref["wB"]) < RTOL64


@pytest.mark.parametrize("arch", ["triblock", "star4"])
def test_anderson_trajectory_and_graph_replay(arch):
    """Anderson m = 5 from the Picard start; captured graph (junction
    kernels, three or more stream branches) equals eager execution."""
    print(f"graph anderson {arch}", flush=True)
    P, e, wA, wB = setup(arch)
    P.mix, P.lambda0, P.depth, P.beta, P.anderson_start = 1, 0.1, 5, 0.2, 1e300
    ref = O.scft_run(P, wA, wB, n_iters=12)
    mix = Mixing("anderson", lambda0=0.1, depth=5, beta=0.2, anderson_start=1e300)
    h = e.iterate(12, mix, use_graph=True)
    assert rel(h["F"], ref["F"]) < 1e-9
    assert rel(h["err"], ref["err"]) < 1e-8
    _, e2, _, _ = setup(arch)
    h2 = e2.iterate(12, mix, use_graph=False)
    assert np.array_equal(h["F"], h2["F"])


@pytest.mark.parametrize("arch", ["star3", "star3_centre_first"])
def test_box_stress_of_graph(arch):
    P, e, wA, wB = setup(arch)
    ref = O.box_stress(P, wA, wB)
    e.solveQ()
    s = e.boxStress()
    assert rel(s, ref) < 1e-10


def test_trapezoid_quadrature_graph():
    P, e, wA, wB = setup("star3", quad=1)
    e.setQuadrature("trapezoid")
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    e.solveQ()
    gA, gB = e.densities()
    assert rel(gA, phiA) < RTOL64 and rel(gB, phiB) < RTOL64


def test_fp32_graph_within_bounds():
    P, e, wA, wB = setup("star4", prec="fp32")
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    e.solveQ()
    assert abs(e.calcBigQ() / Q - 1) < 2e-5
    gA, gB = e.densities()
    assert rel(gA, phiA) < 2e-5 and rel(gB, phiB) < 2e-5


@pytest.mark.parametrize("arch", ["triblock", "star3"])
def test_one_stream_mode_equals_streams(arch, monkeypatch):
    """PSX_SPLIT_SWEEPS=0 (all paths of a round in one launch when their
    lengths agree, else one after another on one stream) gives the same bits."""
    P, e1, wA, wB = setup(arch)
    monkeypatch.setenv("PSX_SPLIT_SWEEPS", "0")
    _, e2, _, _ = setup(arch)
    for e in (e1, e2):
        e.solveQ()
    assert e1.calcBigQ() == e2.calcBigQ()
    assert all(np.array_equal(a, b) for a, b in zip(e1.densities(), e2.densities()))


def test_param_grad_diblock_only():
    _, e, _, _ = setup("triblock")
    e.update()
    with pytest.raises(PsxError, match="INVALID"):
        e.paramGrad()


def test_bad_graphs_raise():
    bad_index = [Block("A", 0.5, 1.0, (), (5,)), Block("B", 0.5, 1.0, (0,), ())]
    with pytest.raises(PsxError, match="INVALID"):
        BlockCopolymerSCFT(SHAPE, CELL, bad_index, 0.05, 10.0)
    ring = [Block("A", 0.5, 1.0, (1,), (1,)), Block("B", 0.5, 1.0, (0,), (0,))]
    with pytest.raises(PsxError, match="INVALID"):
        BlockCopolymerSCFT(SHAPE, CELL, ring, 0.05, 10.0)
    with pytest.raises(PsxError, match="STEPS"):
        BlockCopolymerSCFT(SHAPE, CELL, [Block("A", 0.333, 1.0, (), (1,)), Block("B", 0.667, 1.0, (0,), ())],
                           0.05, 10.0)
