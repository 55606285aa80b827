"""The FFT plugin's other consumers on the GPU against the oracle: the rest
of PsFFTBase (|F|^2, convolution, the imaginary-input pair, forward and
backward real parts), the simple and multi spectral filters (including the
reference's serial region layout on non-square grids), the Poisson update,
the engine's filter between iterations and a deck with a multiSpecFilter.
FP64 to 1e-12 (1e-10 through SCFT iterations), FP32 to 2e-5."""
import os

import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import DiblockSCFT, PsFFT, deck, fields
from polyswift_plusplus_amd.scft import Mixing

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("shape", [(16, 12, 10), (32, 32, 32), (16, 16, 1), (12, 1, 1), (6, 5, 3),
                                   (128, 4, 6), (4, 96, 8), (64, 48, 2), (256, 1, 1)])
def test_fft_family(shape, prec, tol):
    print(f"fft family {shape} {prec}", flush=True)
    rng = np.random.default_rng(11)
    x, y, k = (rng.standard_normal(shape) for _ in range(3))
    p = PsFFT(shape, prec=prec)
    try:
        assert rel(p.forwardFFTAbs(x), O.forward_fft_abs(x)) < tol
        assert rel(p.calcForwardFFT(x), O.forward_fft_re(x)) < tol
        assert rel(p.calcBackwardFFT(x), O.backward_fft_re(x)) < tol
        assert rel(p.convolveRe(x, y), O.convolve_re(x, y)) < tol
        assert rel(p.scaledFFTPairIm(x, k), O.scaled_fft_pair_im(x, k)) < tol
    finally:
        p.close()


@pytest.mark.parametrize("cells", [None, (4, 2, 2), (2, 4, 1), (16, 8, 4)])
@pytest.mark.parametrize("strength,vwall", [(0.0, 0.0), (0.3, 5.0)])
def test_spec_filter_fp64(cells, strength, vwall):
    """Non-square grid: the region of a cell follows the reference's serial
    (transposed) map, which the oracle restates."""
    print(f"spec filter cells={cells} s={strength}", flush=True)
    shape = (16, 8, 4)
    rng = np.random.default_rng(12)
    w = rng.standard_normal(shape)
    p = PsFFT(shape)
    try:
        out = p.specFilter(w, 0.2, strength, cells, vwall)
    finally:
        p.close()
    assert rel(out, O.spec_filter(w, 0.2, strength, cells, vwall)) < 1e-12


@pytest.mark.parametrize("shape,cells", [((32, 32, 1), (4, 4, 1)), ((32, 16, 1), (4, 2, 1)), ((64, 1, 1), (8, 1, 1))])
def test_spec_filter_2d_1d(shape, cells):
    """2D and 1D grids (the reference's filter decks are 2D, specCellSizes
    [4 4 1]), simple and multi."""
    print(f"spec filter {shape} {cells}", flush=True)
    w = np.random.default_rng(15).standard_normal(shape)
    p = PsFFT(shape)
    try:
        for c in (None, cells):
            assert rel(p.specFilter(w, 0.3, 0.1, c), O.spec_filter(w, 0.3, 0.1, c)) < 1e-12
    finally:
        p.close()


def test_spec_filter_fp32_two_modes():
    shape = (32, 8, 8)
    x = np.arange(32)[:, None, None] * np.ones(shape)
    w = 0.7 + np.cos(2 * np.pi * x / 32) + 0.5 * np.cos(2 * np.pi * 3 * x / 32)
    p = PsFFT(shape, prec="fp32")
    try:
        out = p.specFilter(w, 0.5, 0.0)
        multi = p.specFilter(w, 0.5, 0.0, (4, 4, 4))
    finally:
        p.close()
    assert rel(out, np.cos(2 * np.pi * x / 32)) < 2e-5
    assert rel(multi, O.spec_filter(w, 0.5, 0.0, (4, 4, 4))) < 2e-5


def test_spec_filter_cutoff_zero_and_bad_cells():
    shape = (8, 8, 2)
    w = np.random.default_rng(13).standard_normal(shape)
    p = PsFFT(shape)
    try:
        assert np.array_equal(p.specFilter(w, 0.0, 0.0), w)
        with pytest.raises(Exception, match="specCellSizes"):
            p.specFilter(w, 0.5, 0.0, (3, 4, 1))
    finally:
        p.close()


@pytest.mark.parametrize("shape,cell", [((16, 12, 10), (0.3, 0.25, 0.35)), ((32, 32, 1), (0.1, 0.1, 0.1))])
def test_poisson_update(shape, cell):
    print(f"poisson {shape}", flush=True)
    rng = np.random.default_rng(14)
    charge = 0.3 + rng.standard_normal(shape)
    k2 = O.k2(shape, cell)
    kinv = np.where(k2 < 1e-9, 0.0, 4 * 3.1415 * 0.7 / np.where(k2 < 1e-9, 1.0, k2))
    exact = O.scaled_fft_pair(charge - charge.mean(), kinv) / charge.size
    # old psi near the new one in half the cells, far in the rest, 0 in a few
    psi_old = np.where(rng.random(shape) < 0.5, exact * 1.05, exact * 3.0 + 0.1)
    psi_old.flat[:7] = 0.0
    oc, op = O.poisson_update(charge, psi_old, 0.7, cell)
    p = PsFFT(shape)
    try:
        gc, gp = p.poissonUpdate(charge, psi_old, 0.7, cell)
    finally:
        p.close()
    assert rel(gc, oc) < 1e-13
    assert rel(gp, op) < 1e-12


def test_engine_filter_between_iterations():
    """Steepest descent, a multi filter on both fields, more steepest descent:
    the oracle runs the same sequence."""
    shape, cell = (16, 12, 10), (0.3, 0.25, 0.35)
    P = O.Params(*shape, *cell, 0.05, 0.3, 15.0, lambda0=0.1, lambda1=0.05)
    wA, wB = fields.seeded_fields(shape, cell, sigma=0.3, seed=21)
    r1 = O.scft_run(P, wA, wB, n_iters=3)
    fA, fB = (O.spec_filter(w, 0.05, 0.2, (4, 4, 2)) for w in (r1["wA"], r1["wB"]))
    r2 = O.scft_run(P, fA, fB, r1["p"], n_iters=3)
    e = DiblockSCFT(shape, cell, 0.3, 0.05, 15.0)
    try:
        e.setFields(wA, wB)
        mix = Mixing("steepest", 0.1, 0.05)
        h1 = e.iterate(3, mix)
        e.specFilter("AB", 0.05, 0.2, (4, 4, 2))
        gA, gB, _ = e.getFields()
        assert rel(gA, fA) < 1e-10 and rel(gB, fB) < 1e-10
        h2 = e.iterate(3, mix)
        assert rel(np.r_[h1["F"], h2["F"]], np.r_[r1["F"], r2["F"]]) < 1e-10
        gA, gB, _ = e.getFields()
        assert rel(gA, r2["wA"]) < 1e-10 and rel(gB, r2["wB"]) < 1e-10
    finally:
        e.close()


def test_filter_deck_matches_oracle():
    """melt_filter.pre: multiSpecFilter every 4 steps from step 4."""
    root = deck.load(os.path.join(DECKS, "melt_filter.pre"), search_path=[os.path.join(DECKS, "macros")])
    cfg = deck.configure(root)
    h = deck.run(root, init="gaussian")
    (f,) = cfg.filters
    P = O.Params(*cfg.shape, *cfg.cell, cfg.ds, 0.3, cfg.chiN, lambda0=cfg.lambdas[0],
                 lambda1=cfg.lambdas[1])
    wA, wB = fields.seeded_fields(cfg.shape, cfg.cell, sigma=0.1, seed=cfg.seed)
    p, F = None, []
    for _ in range(cfg.nsteps // 4):
        r = O.scft_run(P, wA, wB, p, n_iters=4)
        F.extend(r["F"])
        wA = O.spec_filter(r["wA"], f.cutoffFactor, f.filterStrength, f.specCellSizes)
        wB = O.spec_filter(r["wB"], f.cutoffFactor, f.filterStrength, f.specCellSizes)
        p = r["p"]
    assert rel(h["F"], F) < 1e-10
    gA, gB, _ = h["engine"].getFields()
    assert rel(gA, wA) < 1e-10 and rel(gB, wB) < 1e-10
    h["engine"].close()


def test_filter_before_update_deck_matches_oracle():
    """melt_filter_before.pre: the multiSpecFilter listed before the
    steepest-descent Updater, due at steps 3, 7, 11.  Such a step takes the
    densities of the unfiltered fields, filters the fields, then updates from
    the filtered fields (pstrol/PsDomain.cpp:354-365,
    pseffhamil/PsSteepDUpdater.cpp:170-184); the oracle runs that order."""
    root = deck.load(os.path.join(DECKS, "melt_filter_before.pre"),
                     search_path=[os.path.join(DECKS, "macros")])
    cfg = deck.configure(root)
    print("filter-before deck", cfg.shape, flush=True)
    h = deck.run(root, init="gaussian")
    (f,) = cfg.filters
    due = f.schedule.resolve(cfg.nsteps)
    P = O.Params(*cfg.shape, *cfg.cell, cfg.ds, 0.3, cfg.chiN, lambda0=cfg.lambdas[0],
                 lambda1=cfg.lambdas[1])
    wA, wB = fields.seeded_fields(cfg.shape, cfg.cell, sigma=0.1, seed=cfg.seed)
    p = np.zeros(cfg.shape)
    F, Q = [], []
    for t in range(1, cfg.nsteps + 1):
        if due.due(t):
            phiA, phiB, q = O.diblock_densities(P, wA, wB)
            wA = O.spec_filter(wA, f.cutoffFactor, f.filterStrength, f.specCellSizes)
            wB = O.spec_filter(wB, f.cutoffFactor, f.filterStrength, f.specCellSizes)
            F.append(O.sd_update(P, wA, wB, p, phiA, phiB, q))
            Q.append(q)
        else:
            r = O.scft_run(P, wA, wB, p, n_iters=1)
            wA, wB, p = r["wA"], r["wB"], r["p"]
            F.extend(r["F"])
            Q.extend(r["Q"])
    assert len(h["F"]) == cfg.nsteps
    assert rel(h["F"], F) < 1e-10 and rel(h["Q"], Q) < 1e-10
    split = [t - 1 for t in range(1, cfg.nsteps + 1) if due.due(t)]
    assert split == [2, 6, 10] and np.isnan(h["err"][split]).all()
    assert np.isfinite(np.delete(h["err"], split)).all()
    gA, gB, _ = h["engine"].getFields()
    assert rel(gA, wA) < 1e-10 and rel(gB, wB) < 1e-10
    h["engine"].close()
