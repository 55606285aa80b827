"""The C++ host classes of psx_host.hpp (block graphs, blends, the adjoint
and the batch) driven by the C++ program psx_demo: every figure it prints
equals the same computation through the Python path (ctypes over the same
C ABI) on the same inputs, and the reverse mode's box part equals the box
stress (the closed form for g_phi = 0, g_Q = -1/Q)."""
import os
import subprocess

import numpy as np
import pytest

from polyswift_plusplus_amd import (BatchSCFT, BlendSCFT, Block, BlockCopolymerSCFT, DiblockSCFT,
                                    Mixing, Polymer)

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "polyswift_plusplus_amd", "psx_demo")


def demo(*args):
    print("psx_demo", *args, flush=True)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def values(out, label):
    """The numbers right after 'label =' on the line that starts with it."""
    line = next(ln for ln in out.splitlines() if ln.startswith(label))
    got = []
    for tok in line.split("=", 1)[1].split():
        try:
            got.append(float(tok))
        except ValueError:
            break
    return got


def seed(shape, a):
    """The demo's seed: w_A = cos(2 pi x / Lx) + a cos(2 pi y / Ly), w_B = -w_A."""
    nx, ny, nz = shape
    ix = np.arange(nx)[:, None, None] * np.ones(shape)
    iy = np.arange(ny)[None, :, None] * np.ones(shape)
    wA = np.cos(2 * np.pi * ix / nx) + a * np.cos(2 * np.pi * iy / ny)
    return wA, -wA


SD = Mixing("steepest", 0.1, 0.0)


def test_cpp_adjoint_and_stress():
    out = demo("adjoint", "16")
    shape, cell = (16, 16, 16), (0.2, 0.2, 0.2)
    e = DiblockSCFT(shape, cell, 0.30, 0.02, 15.0)
    try:
        e.setFields(*seed(shape, 0.5))
        e.iterate(5, SD)
        e.solveQ()
        Q = e.calcBigQ()
        st = e.boxStress()
        gA, gB, gL = e.backward(None, None, -1.0 / Q)
        assert abs(values(out, "Q")[0] / Q - 1) < 1e-13
        scale = np.max(np.abs(st))
        assert np.max(np.abs(np.array(values(out, "stress")) - st)) < 1e-12 * scale
        box = np.array(values(out, "backward box"))
        assert np.max(np.abs(box - gL)) < 1e-12 * scale
        # g_phi = 0, g_Q = -1/Q: the box part is -(1/Q) dQ/dL, the stress at fixed fields
        assert np.max(np.abs(box - st)) < 1e-10 * scale
        sums = values(out, "backward sums")
        assert abs(sums[0] - gA.sum()) < 1e-12 * np.abs(gA).sum()
        assert abs(sums[1] - gB.sum()) < 1e-12 * np.abs(gB).sum()
    finally:
        e.close()


def test_cpp_block_graph():
    out = demo("graph")
    shape, cell = (32, 32, 1), (0.125,) * 3
    blocks = [Block("A", 0.25, 1.0, (), (1,)), Block("B", 0.5, 1.0, (0,), (2,)), Block("A", 0.25, 1.0, (1,), ())]
    e = BlockCopolymerSCFT(shape, cell, blocks, 0.01, 20.0)
    try:
        e.setFields(*seed(shape, 0.3))
        h = e.iterate(10, SD)
        e.solveQ()
        line = next(ln for ln in out.splitlines() if ln.startswith("graph F"))
        F_cpp, Q_cpp = float(line.split()[3]), float(line.split()[6])
        assert abs(F_cpp - h["F"][-1]) < 1e-13 * abs(h["F"][-1])
        assert abs(Q_cpp / e.calcBigQ() - 1) < 1e-13
    finally:
        e.close()


def test_cpp_blend():
    out = demo("blend")
    shape, cell = (32, 32, 1), (0.125,) * 3
    polymers = [Polymer(0.8, 1.0, (Block("A", 0.4, 1.0, (), (1,)), Block("B", 0.6, 1.0, (0,), ()))),
                Polymer(0.2, 0.5, (Block("A", 1.0),))]
    e = BlendSCFT(shape, cell, polymers, 0.01, 18.0)
    try:
        e.setFields(*seed(shape, 0.3))
        h = e.iterate(10, SD)
        e.solveQ()
        q = e.polymerQ()
        line = next(ln for ln in out.splitlines() if ln.startswith("blend F"))
        parts = line.split()
        F_cpp, q_cpp = float(parts[3]), np.array([float(parts[6]), float(parts[7])])
        assert abs(F_cpp - h["F"][-1]) < 1e-13 * abs(h["F"][-1])
        assert np.max(np.abs(q_cpp / q - 1)) < 1e-13
    finally:
        e.close()


def test_cpp_batch():
    out = demo("batch")
    specs = [dict(shape=(16, 16, 16), cell=(0.2,) * 3, fA=0.30, ds=0.05, chiN=15.0),
             dict(shape=(12, 12, 12), cell=(0.25,) * 3, fA=0.40, ds=0.05, chiN=18.0),
             dict(shape=(16, 16, 16), cell=(0.2,) * 3, fA=0.45, ds=0.05, chiN=20.0)]
    b = BatchSCFT(specs, prec="fp32")
    try:
        assert b.nclasses == 2
        for i, sp in enumerate(specs):
            b.instances[i].setFields(*seed(sp["shape"], 0.2 * i))
        h = b.iterate(6, Mixing("anderson", lambda0=0.2, depth=5, beta=1.0, anderson_start=0.05))
        assert values(out, "batch classes")[0] == 2
        # the two programs build the seed with different cosine routines
        # (1-ulp differences), so FP32 sweeps agree to rounding, not bit for bit
        for i in range(3):
            assert abs(values(out, f"batch F {i}")[0] / h["F"][i][-1] - 1) < 1e-9, i
    finally:
        b.close()
