"""Plane launches (k_plane: the z and y round trips of whole x-planes, one
launch per contour step) against the oracle and against the line launches,
through the C ABI.  FP64 to 1e-12 in slices and 1e-10 in densities and
trajectories; FP32 to 2e-5.  Covers every plane length (32, 48, 64, 96, 128),
odd and even contour step counts (the launch list's start and end groups),
block graphs with two b classes, checkpointing, the quadrature under the
sweeps, batches in both modes, the autotuner and the error paths."""
import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import (BatchSCFT, Block, BlockCopolymerSCFT, DiblockSCFT, Mixing,
                                    PsxError, fields)

pytestmark = pytest.mark.gpu

CELL = (0.3, 0.25, 0.35)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _case(shape, fA, ds, chiN=15.0, seed=3):
    P = O.Params(*shape, *CELL, ds=ds, fA=fA, chiN=chiN)
    wA, wB = fields.seeded_fields(shape, CELL, sigma=0.5, seed=seed, seed_kind="lamellar")
    return P, wA, wB


def _anderson(P):
    P.mix, P.depth, P.beta, P.lambda0, P.anderson_start = 1, 4, 0.5, 0.1, 1e300


# (shape, fA, ds): Ns = 20, 20, 20, 20, 9 (odd), 10, 11 (odd), 20, 9 (odd)
PLANE_CASES = [((20, 32, 32), 0.3, 0.05), ((6, 64, 64), 0.35, 0.05), ((4, 96, 96), 0.4, 0.05),
               ((3, 128, 128), 0.3, 0.05), ((32, 32, 32), 4 / 9, 1 / 9), ((5, 64, 64), 0.5, 0.1),
               ((7, 96, 96), 5 / 11, 1 / 11), ((6, 48, 48), 0.3, 0.05), ((5, 48, 48), 4 / 9, 1 / 9)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape,fA,ds", PLANE_CASES)
def test_plane_slices_and_densities(shape, fA, ds, prec):
    print(f"plane slices {shape} fA={fA} ds={ds} {prec}", flush=True)
    P, wA, wB = _case(shape, fA, ds)
    phiA, phiB, Q, (qA, qB, qtB, qtA) = O.diblock_densities(P, wA, wB, want_slices=True)
    e = DiblockSCFT(shape, CELL, fA, ds, P.chiN, prec=prec)
    try:
        assert e.sweepMode() == ("planes", True)
        e.setFields(wA, wB)
        e.solveQ()
        tol_s, tol_d = (1e-12, 1e-10) if prec == "fp64" else (2e-5, 2e-5)
        for t in range(e.Ns + 1):
            ref = qA[t] if t <= e.sA else qB[t - e.sA]
            assert rel(e.slice(False, t), ref) < tol_s, t
            ref = qtB[t] if t <= e.sB else qtA[t - e.sB]
            assert rel(e.slice(True, t), ref) < tol_s, t
        assert abs(e.calcBigQ() / Q - 1) < tol_s
        gA, gB = e.densities()
        assert rel(gA, phiA) < tol_d and rel(gB, phiB) < tol_d
    finally:
        e.close()


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("mix", ["steepest", "anderson"])
def test_plane_trajectory(mix, use_graph):
    print(f"plane trajectory {mix} graph={use_graph}", flush=True)
    shape = (8, 32, 32)
    P, wA, wB = _case(shape, 0.3, 0.05)
    if mix == "steepest":
        P.lambda0, P.lambda1 = 0.1, 0.05
        ref = O.scft_run(P, wA, wB, n_iters=8)
        m = Mixing("steepest", 0.1, 0.05)
    else:
        _anderson(P)
        ref = O.scft_run(P, wA, wB, n_iters=8)
        m = Mixing("anderson", lambda0=0.1, depth=4, beta=0.5, anderson_start=1e300)
    e = DiblockSCFT(shape, CELL, 0.3, 0.05, P.chiN)
    try:
        e.setFields(wA, wB)
        h = e.iterate(8, m, use_graph=use_graph)
        assert rel(h["F"], ref["F"]) < 1e-9
        gA, gB, _ = e.getFields()
        assert rel(gA, ref["wA"]) < 1e-9 and rel(gB, ref["wB"]) < 1e-9
    finally:
        e.close()


def test_switching_modes_keeps_fields_and_agrees():
    shape = (6, 64, 64)
    P, wA, wB = _case(shape, 0.35, 0.05)
    e = DiblockSCFT(shape, CELL, 0.35, 0.05, P.chiN)
    try:
        e.setFields(wA, wB)
        m = Mixing("anderson", lambda0=0.1, depth=4, beta=0.5, anderson_start=1e300)
        h1 = e.iterate(3, m)
        e.setSweepMode("lines")
        assert e.sweepMode() == ("lines", True)
        h2 = e.iterate(3, m)
        e.setSweepMode("planes")
        h3 = e.iterate(3, m)
        _anderson(P)
        ref = O.scft_run(P, wA, wB, n_iters=9)
        assert rel(np.r_[h1["F"], h2["F"], h3["F"]], ref["F"]) < 1e-9
        ms, passes = e.timeStep(2)
        assert ms > 0 and passes[0] > 0 and passes[1] == 0 and passes[2] > 0
        e.setSweepMode("lines")
        ms, passes = e.timeStep(2)
        assert all(p > 0 for p in passes)
    finally:
        e.close()


@pytest.mark.parametrize("shape", [(8, 16, 16), (8, 32, 64), (32, 32, 1), (8, 40, 40)])
def test_planes_unavailable(shape):
    e = DiblockSCFT(shape, CELL, 0.3, 0.05, 15.0)
    try:
        assert e.sweepMode() == ("lines", False)
        with pytest.raises(PsxError):
            e.setSweepMode("planes")
        e.setSweepMode("lines")
    finally:
        e.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_set_cell_rewrites_the_plane_factors(prec):
    """The plane launches keep their own copy of the z and y factors: a new
    box (the box relaxation's step) must reach them."""
    shape = (6, 32, 32)
    new = (0.31, 0.22, 0.4)
    P = O.Params(*shape, *new, ds=0.05, fA=0.3, chiN=15.0)
    wA, wB = fields.seeded_fields(shape, CELL, sigma=0.5, seed=8, seed_kind="lamellar")
    _, _, Q, (qA, _, _, _) = O.diblock_densities(P, wA, wB, want_slices=True)
    e = DiblockSCFT(shape, CELL, 0.3, 0.05, 15.0, prec=prec)
    try:
        e.setFields(wA, wB)
        e.solveQ()
        e.setCell(new)
        e.solveQ()
        tol = 1e-12 if prec == "fp64" else 2e-5
        assert abs(e.calcBigQ() / Q - 1) < tol
        assert rel(e.slice(False, e.sA), qA[e.sA]) < tol
    finally:
        e.close()


@pytest.mark.parametrize("K", [1, 6, 7])
def test_plane_checkpointing(K):
    print(f"plane ckpt K={K}", flush=True)
    shape = (8, 32, 32)
    P, wA, wB = _case(shape, 0.3, 0.05)
    e = DiblockSCFT(shape, CELL, 0.3, 0.05, P.chiN)
    try:
        e.setFields(wA, wB)
        e.setCheckpointing(K)
        phiA, phiB, Q = O.diblock_densities(P, wA, wB)
        e.solveQ()
        assert abs(e.calcBigQ() / Q - 1) < 1e-12
        gA, gB = e.densities()
        assert rel(gA, phiA) < 1e-10 and rel(gB, phiB) < 1e-10
        assert rel(e.boxStress(), O.box_stress(P, wA, wB)) < 1e-10
    finally:
        e.close()


def test_plane_block_graph_two_b_classes():
    """Four-arm star with a bSegRatio 1.2 arm: two plane factor tables, and
    fused plane launches across block changes."""
    shape = (6, 32, 32)
    blocks = [Block("A", 0.2, 1.0, (), (1, 2, 3)), Block("A", 0.2, 1.0, (), (0, 2, 3)),
              Block("B", 0.3, 1.0, (), (0, 1, 3)), Block("B", 0.3, 1.2, (0, 1, 2), ())]
    ob = tuple(O.Block(0 if b.species == "A" else 1, b.lengthfrac, b.bSegRatio, b.head, b.tail)
               for b in blocks)
    P = O.Params(*shape, *CELL, 0.05, 0.5, 14.0, blocks=ob, lambda0=0.1, lambda1=0.05)
    wA, wB = fields.seeded_fields(shape, CELL, sigma=0.5, seed=5, seed_kind="lamellar")
    e = BlockCopolymerSCFT(shape, CELL, blocks, 0.05, 14.0)
    try:
        assert e.sweepMode() == ("planes", True)
        e.setFields(wA, wB)
        ref = O.scft_run(P, wA, wB, n_iters=4)
        h = e.iterate(4, Mixing("steepest", 0.1, 0.05))
        assert rel(h["F"], ref["F"]) < 1e-10
        gA, gB, _ = e.getFields()
        assert rel(gA, ref["wA"]) < 1e-10 and rel(gB, ref["wB"]) < 1e-10
    finally:
        e.close()


def test_plane_quadrature_under_the_sweeps(monkeypatch):
    monkeypatch.setenv("PSX_OVERLAP_QUAD", "1")
    shape = (8, 32, 32)
    P, wA, wB = _case(shape, 0.3, 0.05)
    P.lambda0, P.lambda1 = 0.1, 0.05
    ref = O.scft_run(P, wA, wB, n_iters=5)
    e = DiblockSCFT(shape, CELL, 0.3, 0.05, P.chiN)
    try:
        assert e.sweepMode()[0] == "planes"
        e.setFields(wA, wB)
        h = e.iterate(5, Mixing("steepest", 0.1, 0.05))
        assert rel(h["F"], ref["F"]) < 1e-10
        e.setSweepMode("lines")   # the chunks' launch indices are re-planned
        e.setSweepMode("planes")
        h = e.iterate(2, Mixing("steepest", 0.1, 0.05))
        ref2 = O.scft_run(P, ref["wA"], ref["wB"], ref["p"], n_iters=2)
        assert rel(h["F"], ref2["F"]) < 1e-10
    finally:
        e.close()


BSPECS = ([dict(shape=(8, 32, 32), cell=(0.3,) * 3, fA=0.3 + 0.05 * i, ds=0.05, chiN=14.0 + i)
           for i in range(2)] +
          [dict(shape=(2, 96, 96), cell=(0.3, 0.08, 0.08), fA=0.3 + 0.05 * (i % 3), ds=0.05,
                chiN=15.0 + i) for i in range(5)])
BMIX = Mixing("anderson", lambda0=0.2, depth=5, beta=1.0, anderson_start=0.05)


def _bfields(i, sp):
    return fields.seeded_fields(sp["shape"], sp["cell"], sigma=0.3, seed=200 + i, seed_kind="lamellar")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_modes_match_single_engines(prec, monkeypatch):
    """Default class modes (planes below n = 96, lines for the five 96^2
    instances), then the classes switched: every instance equals the single
    engine in the same mode bit for bit."""
    print(f"plane batch {prec}", flush=True)
    monkeypatch.setenv("PSX_OVERLAP_QUAD", "0")
    b = BatchSCFT(BSPECS, prec=prec)
    try:
        assert [b.sweepMode(c) for c in range(b.nclasses)] == ["planes", "lines"]
        for i, sp in enumerate(BSPECS):
            b.instances[i].setFields(*_bfields(i, sp))
        h1 = b.iterate(3, BMIX)
        with pytest.raises(PsxError):
            b.instances[0].setSweepMode("lines")
        b.setSweepMode(0, "lines")
        b.setSweepMode(1, "planes")
        h2 = b.iterate(2, BMIX)
        for i, sp in enumerate(BSPECS):
            e = DiblockSCFT(sp["shape"], sp["cell"], sp["fA"], sp["ds"], sp["chiN"], prec=prec)
            try:
                e.setFields(*_bfields(i, sp))
                e.setSweepMode("planes" if i < 2 else "lines")
                r1 = e.iterate(3, BMIX)
                e.setSweepMode("lines" if i < 2 else "planes")
                r2 = e.iterate(2, BMIX)
                assert np.array_equal(np.r_[h1["F"][i], h2["F"][i]], np.r_[r1["F"], r2["F"]]), i
                for got, want in zip(b.instances[i].getFields(), e.getFields()):
                    assert np.array_equal(got, want), i
            finally:
                e.close()
    finally:
        b.close()


def test_autotune_chooses_a_mode_and_keeps_results():
    shape = (8, 64, 64)
    P, wA, wB = _case(shape, 0.3, 0.05)
    P.lambda0, P.lambda1 = 0.1, 0.05
    ref = O.scft_run(P, wA, wB, n_iters=3)
    e = DiblockSCFT(shape, CELL, 0.3, 0.05, P.chiN)
    try:
        e.setFields(wA, wB)
        e.autotune(2)
        assert e.sweepMode()[0] in ("planes", "lines")
        h = e.iterate(3, Mixing("steepest", 0.1, 0.05))
        assert rel(h["F"], ref["F"]) < 1e-10
    finally:
        e.close()
    specs = BSPECS[:3]
    b = BatchSCFT(specs, prec="fp64")
    try:
        for i, sp in enumerate(specs):
            b.instances[i].setFields(*_bfields(i, sp))
        b.autotune(2)
        modes = [b.sweepMode(c) for c in range(b.nclasses)]
        print("batch autotune modes:", modes, flush=True)
        assert all(m in ("planes", "lines") for m in modes)
        h = b.iterate(4, BMIX)
        for i, sp in enumerate(specs):
            e = DiblockSCFT(sp["shape"], sp["cell"], sp["fA"], sp["ds"], sp["chiN"])
            try:
                e.setFields(*_bfields(i, sp))
                ref = e.iterate(4, BMIX)
                assert np.max(np.abs(h["F"][i] - ref["F"])) <= 1e-11 * np.max(np.abs(ref["F"])), i
            finally:
                e.close()
    finally:
        b.close()


@pytest.mark.parametrize("n,nx,count", [(96, 48, 6), (64, 60, 9)])
def test_fp32_plane_layouts_agree(n, nx, count, monkeypatch):
    """FP32 plane launches at n = 64 and 96 alias the transpose buffers with
    the staged plane only when the launch has more blocks than the CUs hold
    in the barrier-free layout.  A batch class of `count` (nx, n, n)
    instances (2 * count * nx plane blocks per launch) takes the aliased
    layout, each single engine the barrier-free one: every instance must
    equal its single engine bit for bit."""
    print(f"fp32 plane layouts n={n} nx={nx} count={count}", flush=True)
    monkeypatch.setenv("PSX_OVERLAP_QUAD", "0")
    specs = [dict(shape=(nx, n, n), cell=(0.3, 0.08, 0.08), fA=0.3 + 0.05 * (i % 3), ds=0.05,
                  chiN=15.0 + i) for i in range(count)]
    b = BatchSCFT(specs, prec="fp32")
    try:
        b.setSweepMode(0, "planes")
        for i, sp in enumerate(specs):
            b.instances[i].setFields(*_bfields(i, sp))
        h = b.iterate(2, BMIX)
        for i, sp in enumerate(specs):
            e = DiblockSCFT(sp["shape"], sp["cell"], sp["fA"], sp["ds"], sp["chiN"], prec="fp32")
            try:
                e.setFields(*_bfields(i, sp))
                e.setSweepMode("planes")
                r = e.iterate(2, BMIX)
                assert np.array_equal(h["F"][i], r["F"]), i
                for got, want in zip(b.instances[i].getFields(), e.getFields()):
                    assert np.array_equal(got, want), i
            finally:
                e.close()
    finally:
        b.close()
