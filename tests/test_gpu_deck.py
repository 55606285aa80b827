"""Decks run on the GPU: the fixture decks' steepest-descent relaxations
against the oracle on the same configuration and seeded fields (FP64 to
1e-10).  Noise is off in both (noise-free parity runs)."""
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


@pytest.mark.parametrize("name", ["melt_diblock.pre", "melt_star.pre"])
def test_deck_run_matches_oracle(name):
    print(f"deck {name}", flush=True)
    root = deck.load(os.path.join(DECKS, name), search_path=[os.path.join(DECKS, "macros")])
    cfg = deck.configure(root)
    h = deck.run(root)
    assert h["iters"] == cfg.nsteps
    blocks = tuple(O.Block(0 if b.species == "A" else 1, b.lengthfrac, b.bSegRatio, b.head, b.tail)
                   for b in cfg.blocks)
    P = O.Params(*cfg.shape, *cfg.cell, cfg.ds, 0.5, cfg.chiN, lambda0=cfg.lambdas[0],
                 lambda1=cfg.lambdas[1], blocks=blocks)
    wA, wB = fields.seeded_fields(cfg.shape, cfg.cell, sigma=0.1, seed=cfg.seed)
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
