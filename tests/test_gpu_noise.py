"""Steepest descent with the reference's noise term on the GPU against the
oracle: the engine draws each step's 2N deviates from the library's host
stream (equal to the oracle's, see test_oracle_rng.py) and adds them before
the pressure and F.  FP64 to 1e-10; graph and stream paths; the step API;
FP32 sweeps; Anderson with noise refused."""
import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import DiblockSCFT, PsxError, fields
from polyswift_plusplus_amd.scft import Mixing, Rng

pytestmark = pytest.mark.gpu

SHAPE, CELL = (16, 12, 10), (0.3, 0.25, 0.35)


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _setup(seed=31):
    P = O.Params(*SHAPE, *CELL, 0.05, 0.3, 15.0, lambda0=0.1, lambda1=0.05)
    wA, wB = fields.seeded_fields(SHAPE, CELL, sigma=0.2, seed=seed)
    return P, wA, wB


@pytest.mark.parametrize("use_graph", [True, False])
def test_noisy_iterations_match_oracle(use_graph):
    print(f"noise graph={use_graph}", flush=True)
    P, wA, wB = _setup()
    ref = O.scft_run_noise(P, wA, wB, n_iters=7, strength=0.05, rng=O.Rng(123))
    e = DiblockSCFT(SHAPE, CELL, 0.3, 0.05, 15.0)
    try:
        e.setFields(wA, wB)
        e.setNoise(0.05, Rng(123))
        h = e.iterate(7, Mixing("steepest", 0.1, 0.05), use_graph=use_graph)
        assert rel(h["F"], ref["F"]) < 1e-10
        assert rel(h["err"], ref["err"]) < 1e-9
        gA, gB, gp = e.getFields()
        assert rel(gA, ref["wA"]) < 1e-10 and rel(gB, ref["wB"]) < 1e-10
        assert rel(gp, ref["p"]) < 1e-9
    finally:
        e.close()


def test_noise_through_the_step_api_and_off_again():
    """solveQ / densities / steepestDescent draws the same deviates; turning
    the noise off continues noise-free (and leaves the stream where it was)."""
    P, wA, wB = _setup(32)
    g = O.Rng(7)
    r1 = O.scft_run_noise(P, wA, wB, n_iters=2, strength=0.1, rng=g)
    r2 = O.scft_run(P, r1["wA"], r1["wB"], r1["p"], n_iters=2)
    r3 = O.scft_run_noise(P, r2["wA"], r2["wB"], r2["p"], n_iters=1, strength=0.1, rng=g)
    e = DiblockSCFT(SHAPE, CELL, 0.3, 0.05, 15.0)
    rng = Rng(7)
    try:
        e.setFields(wA, wB)
        e.setNoise(0.1, rng)
        F = []
        for _ in range(2):
            e.solveQ()
            e.densities()
            F.append(e.steepestDescent(0.1, 0.05))
        assert rel(F, r1["F"]) < 1e-10
        e.setNoise(0.0, None)
        h = e.iterate(2, Mixing("steepest", 0.1, 0.05))
        assert rel(h["F"], r2["F"]) < 1e-10
        e.setNoise(0.1, rng)
        h = e.iterate(1, Mixing("steepest", 0.1, 0.05))
        assert rel(h["F"], r3["F"]) < 1e-10
    finally:
        e.close()


def test_noise_fp32_sweeps():
    P, wA, wB = _setup(33)
    ref = O.scft_run_noise(P, wA, wB, n_iters=5, strength=0.05, rng=O.Rng(5))
    e = DiblockSCFT(SHAPE, CELL, 0.3, 0.05, 15.0, prec="fp32")
    try:
        e.setFields(wA, wB)
        e.setNoise(0.05, Rng(5))
        h = e.iterate(5, Mixing("steepest", 0.1, 0.05))
        # F is ~1e-3 here: FP32 sweeps hold it to ~1e-6 absolute
        assert np.max(np.abs(h["F"] - ref["F"])) < 1e-5
        gA, _, _ = e.getFields()
        assert rel(gA, ref["wA"]) < 1e-4
    finally:
        e.close()


def test_anderson_with_noise_is_refused():
    e = DiblockSCFT(SHAPE, CELL, 0.3, 0.05, 15.0)
    rng = Rng(1)
    try:
        e.setFields(*_setup()[1:])
        e.setNoise(0.05, rng)
        with pytest.raises(PsxError, match="steepest"):
            e.iterate(2, Mixing("anderson", 0.1, 0.0, depth=4, beta=0.5, anderson_start=1.0))
    finally:
        e.close()


def test_noise_with_contour_checkpointing():
    """The checkpointed sweep (K = 7) adds the same noise in its last chunk."""
    P, wA, wB = _setup(34)
    ref = O.scft_run_noise(P, wA, wB, n_iters=4, strength=0.05, rng=O.Rng(9))
    e = DiblockSCFT(SHAPE, CELL, 0.3, 0.05, 15.0)
    try:
        e.setFields(wA, wB)
        e.setCheckpointing(7)
        e.setNoise(0.05, Rng(9))
        h = e.iterate(4, Mixing("steepest", 0.1, 0.05))
        assert rel(h["F"], ref["F"]) < 1e-10
        gA, gB, _ = e.getFields()
        assert rel(gA, ref["wA"]) < 1e-10 and rel(gB, ref["wB"]) < 1e-10
    finally:
        e.close()
