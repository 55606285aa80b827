"""Decks run on the GPU: the fixture decks' steepest-descent relaxations
against the oracle on the same configuration (FP64 to 1e-10), from the
reference's random initial fields or the project's seeded ones, with the
decks' noise drawn from the reference's random stream in both."""
import os

import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import deck, fields

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def reference_fields(cfg):
    """The reference's initial fields from the oracle's restatement of its
    random stream: N uniforms per monomerDens PhysField in deck order,
    w = 1.0 * (u - 0.5)."""
    g = O.Rng(cfg.seed)
    n = int(np.prod(cfg.shape))
    drawn = {name: (1.0 * (g.uniform(n) + (-0.5))).reshape(cfg.shape) for name in cfg.init_fields}
    return drawn[cfg.species_fields[0]], drawn[cfg.species_fields[1]], g


@pytest.mark.parametrize("init", ["reference", "gaussian"])
@pytest.mark.parametrize("name", ["melt_diblock.pre", "melt_star.pre"])
def test_deck_run_matches_oracle(name, init):
    print(f"deck {name} {init}", flush=True)
    root = deck.load(os.path.join(DECKS, name), search_path=[os.path.join(DECKS, "macros")])
    cfg = deck.configure(root)
    h = deck.run(root, init=init)
    assert h["iters"] == cfg.nsteps
    blocks = tuple(O.Block(0 if b.species == "A" else 1, b.lengthfrac, b.bSegRatio, b.head, b.tail)
                   for b in cfg.blocks)
    P = O.Params(*cfg.shape, *cfg.cell, cfg.ds, 0.5, cfg.chiN, lambda0=cfg.lambdas[0],
                 lambda1=cfg.lambdas[1], blocks=blocks)
    if init == "reference":
        wA, wB, g = reference_fields(cfg)
    else:
        wA, wB = fields.seeded_fields(cfg.shape, cfg.cell, sigma=0.1, seed=cfg.seed)
        g = O.Rng(cfg.seed)
    if cfg.noise:   # the star deck's noise comes from the same stream
        ref = O.scft_run_noise(P, wA, wB, n_iters=cfg.nsteps, strength=cfg.noise, rng=g)
    else:
        ref = O.scft_run(P, wA, wB, n_iters=cfg.nsteps)
    assert rel(h["F"], ref["F"]) < 1e-10
    assert rel(h["Q"], ref["Q"]) < 1e-10
    gA, gB, _ = h["engine"].getFields()
    assert rel(gA, ref["wA"]) < 1e-10 and rel(gB, ref["wB"]) < 1e-10


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "polyswift_plusplus_amd",
                                                    "libpsx_h5.so")), reason="libpsx_h5.so not built")
def test_dump_restore_resume_is_bitwise(tmp_path):
    """Dump the PhysFields every 4 steps (HDF5, the reference's layout), then
    restart a fresh engine from the step-4 files: steps 5..8 repeat bit for
    bit (the reference's -r restart, pstrol/PsDomain.cpp:306-325)."""
    root = deck.load(os.path.join(DECKS, "melt_diblock.pre"))
    root.attrs["dumpPeriodicity"] = 4
    prefix = str(tmp_path / "melt")
    full = deck.run(root, iters=8, prefix=prefix)
    assert os.path.exists(prefix + "_totA_1000004.h5") and os.path.exists(prefix + "_defaultPressure_1000008.h5")
    resumed = deck.run(root, iters=4, prefix=prefix, restore_step=4)
    assert resumed["step"] == 8
    assert np.array_equal(resumed["F"], full["F"][4:])
    for a, b in zip(resumed["engine"].getFields(), full["engine"].getFields()):
        assert np.array_equal(a, b)


def test_noisy_deck_matches_oracle():
    """melt_noise.pre: steepest descent with noise 0.05; the initial fields and
    every step's noise come from one MT19937 stream in the reference's order
    (init draws first, then per step w_A's N deviates, then w_B's)."""
    root = deck.load(os.path.join(DECKS, "melt_noise.pre"))
    cfg = deck.configure(root)
    assert cfg.noise == 0.05
    h = deck.run(root)
    wA, wB, g = reference_fields(cfg)
    P = O.Params(*cfg.shape, *cfg.cell, cfg.ds, 0.3, cfg.chiN, lambda0=cfg.lambdas[0],
                 lambda1=cfg.lambdas[1])
    ref = O.scft_run_noise(P, wA, wB, n_iters=cfg.nsteps, strength=cfg.noise, rng=g)
    assert rel(h["F"], ref["F"]) < 1e-10
    gA, gB, _ = h["engine"].getFields()
    assert rel(gA, ref["wA"]) < 1e-10 and rel(gB, ref["wB"]) < 1e-10
    # the noise-free run of the same deck differs
    quiet = deck.run(root, noise=False)
    assert rel(quiet["F"], ref["F"]) > 1e-6
