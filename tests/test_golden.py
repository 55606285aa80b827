"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the
oracle): the oracle must keep reproducing them (CPU), and the HIP engine
must match them to the FP64 parity bar (GPU)."""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["c1_2d_picard", "small_3d_anderson", "small_3d_sd2"]


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False))


def params(g):
    shape = tuple(int(x) for x in g["shape"])
    return O.Params(*shape, *[float(x) for x in g["cell"]], ds=float(g["ds"]), fA=float(g["fA"]),
                    chiN=float(g["chiN"]), mix=int(g["mix"]), lambda0=float(g["lambda0"]),
                    lambda1=float(g["lambda1"]), depth=int(g["depth"]), beta=float(g["beta"]),
                    anderson_start=float(g["start"]))


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    g = load(name)
    P = params(g)
    n = int(g["iters"]) if name != "c1_2d_picard" else 20   # C1: first 20 of 200
    O.set_threads(4)
    r = O.scft_run(P, g["wA0"], g["wB0"], n_iters=n)
    assert rel(r["F"], g["F"][:n]) < 1e-12
    assert rel(r["Q"], g["Q"][:n]) < 1e-12
    if n == int(g["iters"]):
        assert rel(r["wA"], g["wA"]) < 1e-12 and rel(r["phiA"], g["phiA"]) < 1e-12


def test_fft_pair_fixture_oracle():
    g = load("fft_pair_6x10x8")
    assert rel(O.scaled_fft_pair(g["data"], g["kdata"]), g["res"]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_engine_matches_fixture(name):
    from polyswift_plusplus_amd import DiblockSCFT, Mixing
    g = load(name)
    print(f"golden {name}", flush=True)
    shape = tuple(int(x) for x in g["shape"])
    e = DiblockSCFT(shape, [float(x) for x in g["cell"]], float(g["fA"]), float(g["ds"]),
                    float(g["chiN"]))
    e.setFields(g["wA0"], g["wB0"])
    mix = Mixing("steepest" if int(g["mix"]) == 0 else "anderson", float(g["lambda0"]),
                 float(g["lambda1"]), int(g["depth"]), float(g["beta"]), float(g["start"]))
    h = e.iterate(int(g["iters"]), mix)
    # Anderson trajectories amplify rounding through the least-squares solve
    tol = 1e-10 if int(g["mix"]) == 0 else 1e-8
    assert rel(h["F"], g["F"]) < tol
    assert rel(h["Q"], g["Q"]) < tol
    wA, wB, p = e.getFields()
    phiA, phiB = e.getDensities()
    scale = max(np.max(np.abs(g["wA"])), 1e-3)
    assert np.max(np.abs(wA - g["wA"])) < tol * 10 * scale
    assert np.max(np.abs(phiA - g["phiA"])) < tol * 10


@pytest.mark.gpu
def test_fft_pair_fixture_engine():
    from polyswift_plusplus_amd import PsFFT
    g = load("fft_pair_6x10x8")
    assert rel(PsFFT(g["data"].shape).scaledFFTPair(g["data"], g["kdata"]), g["res"]) < 1e-12
