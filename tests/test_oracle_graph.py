"""Pins of the oracle's block-graph restatement (pspolymer/PsBlockCopolymer.cpp
189-313, pspolymer/PsBlock.cpp 146-147, 205-237, 387-472): closed forms,
invariants, and an independent numpy propagation written from the
equations (np.fft, whole-graph junction products), on CPU."""
import numpy as np
import pytest

import oracle as O
from oracle import Block

SHAPE = (8, 6, 4)
CELL = (0.35, 0.4, 0.45)


def P_of(blocks, ds=0.05, chiN=15.0, quad=0, **kw):
    return O.Params(*SHAPE, *CELL, ds, 0.5, chiN, quad=quad, blocks=tuple(blocks), **kw)


def fields(seed=3, amp=0.6):
    rng = np.random.default_rng(seed)
    return rng.normal(0, amp, SHAPE), rng.normal(0, amp, SHAPE)


# ---- the reference's architectures ---------------------------------------
def diblock(fA=0.3, bA=1.0, bB=1.0):
    return [Block(0, fA, bA, (), (1,)), Block(1, 1 - fA, bB, (0,), ())]


def triblock_aba(f1=0.25, f2=0.5):
    # pstests/tests/diblock/triblock.pre: A-B-A linear
    return [Block(0, f1, 1.0, (), (1,)), Block(1, f2, 1.0, (0,), (2,)),
            Block(0, 1 - f1 - f2, 1.0, (1,), ())]


def star3():
    # pstests/tests/diblock/star3ab.pre: arms A and B joined at their tails to
    # the head of a second A arm
    return [Block(0, 0.25, 1.0, (), (1, 2)), Block(1, 0.5, 1.0, (), (0, 2)),
            Block(0, 0.25, 1.0, (0, 1), ())]


# ---- independent numpy model ---------------------------------------------
def np_kfac(ds, b):
    k2 = np.zeros(SHAPE)
    for d, (n, h) in enumerate(zip(SHAPE, CELL)):
        kd = 2 * np.pi * np.fft.fftfreq(n, d=h)      # |k| with the Nyquist mode at +n/2
        kd = np.abs(kd)
        if n % 2 == 0:
            kd[n // 2] = 2 * np.pi * (n / 2) / (n * h)
        sh = [1, 1, 1]
        sh[d] = n
        k2 = k2 + (kd ** 2).reshape(sh)
    return np.exp(-ds * b * b * k2)


def np_propagate(q0, w, ds, b, steps):
    K = np_kfac(ds, b)
    W = np.exp(-0.5 * ds * w)
    out = [q0]
    q = q0
    for _ in range(steps):
        q = W * np.real(np.fft.ifftn(K * np.fft.fftn(W * q)))
        out.append(q)
    return np.array(out)


def trap_w(M, ds):
    c = np.full(M, ds)
    c[0] = c[-1] = 0.5 * ds
    return c


def simpson_quirk_w(M, ds):
    """pspolymer/PsBlock.cpp:256-380 as per-point weights."""
    c = np.zeros(M)
    last = M - 2 if M % 2 == 0 else M - 1
    for n in range(0, last - 1, 2):
        c[n] += 1; c[n + 1] += 4; c[n + 2] += 1
    c[:last + 1] *= 0.33333 * ds
    if M % 2 == 0:
        c[M - 2] += 0.5 * ds
        c[M - 1] += 0.5 * ds
    return c


def np_star_tail_junction(arms_free_head, center, w, ds, quad):
    """Star of arms joined at their tails to `center`'s head, center's tail
    free (star3ab.pre's topology), by direct propagation: returns Q, phi."""
    steps = [O.block_steps(b.frac, ds) for b in arms_free_head + [center]]
    ws = lambda b: w[b.species]
    ones = np.ones(SHAPE)
    arm_h = [np_propagate(ones, ws(b), ds, b.b, s) for b, s in zip(arms_free_head, steps)]
    cen_t = np_propagate(ones, ws(center), ds, center.b, steps[-1])
    # the centre's head starts at the product of the arms' finals
    cen_h = np_propagate(np.prod([a[-1] for a in arm_h], axis=0), ws(center), ds, center.b, steps[-1])
    arm_t = []
    for i, (b, s) in enumerate(zip(arms_free_head, steps)):
        init = cen_t[-1] * np.prod([a[-1] for j, a in enumerate(arm_h) if j != i], axis=0)
        arm_t.append(np_propagate(init, ws(b), ds, b.b, s))
    Q = np.mean(arm_t[0][-1])            # block 0 head is free
    wfun = trap_w if quad == 1 else simpson_quirk_w
    phi = [np.zeros(SHAPE), np.zeros(SHAPE)]
    pairs = list(zip(arms_free_head, arm_h, arm_t)) + [(center, cen_h, cen_t)]
    for b, qh, qt in pairs:
        M = qh.shape[0]
        c = wfun(M, ds)
        phi[b.species] += np.einsum("s,s...->...", c, qh * qt[::-1]) / Q
    return Q, phi


# ---- tests ------------------------------------------------------------------
def test_graph_diblock_bitwise_equals_diblock():
    P = O.Params(*SHAPE, *CELL, 0.05, 0.3, 15.0)
    wA, wB = fields()
    a = O.diblock_densities(P, wA, wB, True)
    g = O.graph_densities(P_of(diblock(0.3)), wA, wB, True)
    assert np.array_equal(a[0], g[0]) and np.array_equal(a[1], g[1]) and a[2] == g[2]


@pytest.mark.parametrize("arch", ["triblock", "star"])
def test_homogeneous_closed_form(arch):
    """w_A = a, w_B = b: only k = 0 survives, Q = exp(-a f_A - b f_B), and
    q q+ = Q at every contour point, so phi_A = sum over A blocks of the
    quadrature weights (the 0.99999 Simpson quirk)."""
    blocks = triblock_aba() if arch == "triblock" else star3()
    P = P_of(blocks)
    a, b = 0.7, -0.4
    phiA, phiB, Q = O.graph_densities(P, np.full(SHAPE, a), np.full(SHAPE, b))
    fA, fB = O.species_fracs(P)
    assert Q == pytest.approx(np.exp(-a * fA - b * fB), rel=1e-13)
    wsum = [0.0, 0.0]
    for blk in blocks:
        M = O.block_steps(blk.frac, P.ds) + 1
        wsum[blk.species] += simpson_quirk_w(M, P.ds).sum()
    assert np.allclose(phiA, wsum[0], rtol=1e-12) and np.allclose(phiB, wsum[1], rtol=1e-12)


@pytest.mark.parametrize("arch", ["triblock", "star"])
def test_contour_invariant_every_block(arch):
    """sum_r q_head[s] q_tail[M-1-s] = N Q at every point of every block."""
    P = P_of(triblock_aba() if arch == "triblock" else star3())
    wA, wB = fields(5)
    _, _, Q, sl = O.graph_densities(P, wA, wB, True)
    N = np.prod(SHAPE)
    for qh, qt in sl:
        inv = (qh * qt[::-1]).reshape(qh.shape[0], -1).sum(axis=1)
        assert np.allclose(inv, N * Q, rtol=1e-12)


@pytest.mark.parametrize("quad", [0, 1])
def test_star_matches_numpy_propagation(quad):
    """star3ab topology against direct numpy propagation: Q and phi."""
    P = P_of(star3(), quad=quad)
    wA, wB = fields(7)
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    b = star3()
    Qn, phin = np_star_tail_junction([b[0], b[1]], b[2], (wA, wB), P.ds, quad)
    assert Q == pytest.approx(Qn, rel=1e-12)
    assert np.allclose(phiA, phin[0], rtol=1e-11, atol=1e-13)
    assert np.allclose(phiB, phin[1], rtol=1e-11, atol=1e-13)


def test_split_block_is_the_same_chain():
    """A(0.2)-A(0.2)-B(0.6) is the diblock A(0.4)-B(0.6): same Q, and with
    the trapezoid rule (additive over sub-blocks) the same densities."""
    wA, wB = fields(9)
    d = O.graph_densities(P_of(diblock(0.4), quad=1), wA, wB)
    split = [Block(0, 0.2, 1.0, (), (1,)), Block(0, 0.2, 1.0, (0,), (2,)), Block(1, 0.6, 1.0, (1,), ())]
    s = O.graph_densities(P_of(split, quad=1), wA, wB)
    assert s[2] == pytest.approx(d[2], rel=1e-13)
    assert np.allclose(s[0], d[0], rtol=1e-12) and np.allclose(s[1], d[1], rtol=1e-12)


def test_tail_to_tail_pair_is_a_diblock():
    """Blocks joined tail to tail (B's direction reversed) form the same
    chain: Q equal; trapezoid densities equal."""
    wA, wB = fields(11)
    d = O.graph_densities(P_of(diblock(0.3), quad=1), wA, wB)
    tt = [Block(0, 0.3, 1.0, (), (1,)), Block(1, 0.7, 1.0, (), (0,))]
    s = O.graph_densities(P_of(tt, quad=1), wA, wB)
    assert s[2] == pytest.approx(d[2], rel=1e-13)
    assert np.allclose(s[0], d[0], rtol=1e-12) and np.allclose(s[1], d[1], rtol=1e-12)


def test_bad_graphs_rejected():
    wA, wB = fields()
    with pytest.raises(ValueError):   # index out of range
        O.graph_densities(P_of([Block(0, 0.5, 1.0, (), (3,)), Block(1, 0.5, 1.0, (0,), ())]), wA, wB)
    with pytest.raises(ValueError):   # non-integral steps
        O.graph_densities(P_of([Block(0, 0.333, 1.0, (), (1,)), Block(1, 0.667, 1.0, (0,), ())]), wA, wB)
    ring = [Block(0, 0.5, 1.0, (1,), (1,)), Block(1, 0.5, 1.0, (0,), (0,))]
    with pytest.raises((ValueError, RuntimeError)):   # a ring: no free end to start from
        O.graph_densities(P_of(ring), wA, wB)


def test_graph_run_uses_species_totals():
    """An SD iteration of the triblock: F equals the closed-form pieces
    evaluated with f_A, f_B summed over blocks."""
    P = P_of(triblock_aba(), chiN=12.0)
    wA, wB = fields(13, 0.3)
    r = O.scft_run(P, wA, wB, n_iters=1)
    phiA, phiB, Q = O.graph_densities(P, wA, wB)
    fA, fB = O.species_fracs(P)
    assert (fA, fB) == pytest.approx((0.5, 0.5))
    wA1, wB1 = r["wA"], r["wB"]
    F = (-np.log(Q) + np.mean(-(phiA * wA1) - (phiB * wB1))
         + np.mean(P.chiN * phiA * phiB - P.chiN * fA * fB))
    assert r["F"][0] == pytest.approx(F, rel=1e-12)


def test_graph_stress_matches_finite_differences():
    """Box stress of the star against central differences of -ln Q in L."""
    P = P_of(star3())
    wA, wB = fields(17, 0.4)
    s = O.box_stress(P, wA, wB)
    h = 1e-5
    for d in range(3):
        vals = []
        for sgn in (1, -1):
            cell = list(CELL)
            L = SHAPE[d] * cell[d]
            cell[d] = (L + sgn * h) / SHAPE[d]
            Pd = O.Params(*SHAPE, *cell, P.ds, 0.5, P.chiN, blocks=P.blocks)
            vals.append(-np.log(O.graph_densities(Pd, wA, wB)[2]))
        assert s[d] == pytest.approx((vals[0] - vals[1]) / (2 * h), rel=1e-6, abs=1e-10)
