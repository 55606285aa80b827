"""Pins of the CPU oracle (no GPU): every oracle function is checked against
something it does not share code with -- numpy's pocketfft, brute-force
DFTs, closed forms, invariants -- chosen so that a dropped term, a wrong sign
or index, or a transposed operand fails at least one check.

Reference lines each pin restates are cited in oracle/psx_oracle.cpp.
"""
import numpy as np
import pytest

import oracle as O


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# --------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(4, 6, 8), (3, 5, 1), (7, 1, 1), (2, 9, 10), (1, 1, 1),
                                   (11, 13, 2), (16, 12, 10), (32, 48, 30)])
def test_fft3_matches_numpy(shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel(O.fft3(a, -1), np.fft.fftn(a)) < 1e-14
    assert rel(O.fft3(a, +1), np.fft.ifftn(a) * a.size) < 1e-14


@pytest.mark.parametrize("shape", [(4, 6, 8), (3, 5, 1), (5, 3, 7)])
def test_fft3_matches_bruteforce(shape):
    rng = np.random.default_rng(2)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel(O.fft3(a, -1), O.dft3_brute(a, -1)) < 1e-14
    assert rel(O.fft3(a, +1), O.dft3_brute(a, +1)) < 1e-14
    # the brute force itself against numpy (it pins the sign convention)
    assert rel(O.dft3_brute(a, -1), np.fft.fftn(a)) < 1e-14


def test_fft3_round_trip_linearity_parseval():
    rng = np.random.default_rng(3)
    shape = (12, 10, 6)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    b = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    N = a.size
    assert rel(O.fft3(O.fft3(a, -1), +1), N * a) < 1e-14
    assert rel(O.fft3(2.0 * a - 3.0 * b, -1), 2.0 * O.fft3(a, -1) - 3.0 * O.fft3(b, -1)) < 1e-14
    assert abs(np.sum(np.abs(O.fft3(a, -1)) ** 2) / (N * np.sum(np.abs(a) ** 2)) - 1) < 1e-13


# --------------------------------------------------------------------- k^2
def k2_closed_form(shape, cell):
    ks = [2 * np.pi * np.fft.fftfreq(n, d=h) for n, h in zip(shape, cell)]
    KX, KY, KZ = np.meshgrid(*ks, indexing="ij")
    return KX ** 2 + KY ** 2 + KZ ** 2


@pytest.mark.parametrize("shape,cell", [((8, 6, 5), (0.3, 0.2, 0.7)), ((64, 64, 1), (0.0625,) * 3),
                                        ((7, 4, 1), (1.0, 0.5, 2.0))])
def test_kfac_closed_form(shape, cell):
    ds, b = 0.01, 1.3
    kf = O.kfac(shape, cell, ds, b)
    assert rel(kf, np.exp(-ds * b * b * k2_closed_form(shape, cell))) < 1e-14
    assert kf[0, 0, 0] == 1.0


# --------------------------------------------------------------- FFT pair
def test_scaled_fft_pair_matches_numpy():
    rng = np.random.default_rng(4)
    d = rng.standard_normal((8, 12, 10))
    k = rng.random((8, 12, 10))
    ref = np.real(np.fft.ifftn(k * np.fft.fftn(d))) * d.size
    assert rel(O.scaled_fft_pair(d, k), ref) < 1e-13


def test_scaled_fft_pair_single_mode():
    # cos(2 pi m x / L) is an eigenvector: result = N kfac(m) cos(...)
    shape, cell, m = (16, 8, 4), (0.25, 0.3, 0.5), 3
    x = np.arange(shape[0])[:, None, None] * cell[0]
    L = shape[0] * cell[0]
    d = np.cos(2 * np.pi * m * x / L) * np.ones(shape)
    kf = O.kfac(shape, cell, 0.01)
    km = (2 * np.pi * m / L) ** 2
    out = O.scaled_fft_pair(d, kf)
    assert rel(out, d.size * np.exp(-0.01 * km) * d) < 1e-13


# ------------------------------------------------------------ block steps
@pytest.mark.parametrize("f,ds,want", [(0.30, 0.01, 30), (0.70, 0.01, 70), (0.35, 0.01, 35),
                                       (0.32, 0.005, 64), (0.5, 0.05, 10), (0.99999, 0.1, 10)])
def test_block_steps(f, ds, want):
    assert O.block_steps(f, ds) == want


@pytest.mark.parametrize("f,ds", [(0.333, 0.01), (0.03, 0.01), (0.5, 0.3)])
def test_block_steps_errors(f, ds):
    with pytest.raises(ValueError):
        O.block_steps(f, ds)


# ------------------------------------------------------------- propagator
def test_solve_q_homogeneous_decay():
    shape, a, ds, steps = (6, 4, 5), 0.7, 0.02, 9
    kf = O.kfac(shape, (0.3, 0.3, 0.3), ds)
    wf = np.full(shape, np.exp(-0.5 * ds * a))
    q = O.solve_q(kf, wf, np.ones(shape), steps)
    for s in range(steps + 1):
        assert rel(q[s], np.full(shape, np.exp(-a * s * ds))) < 1e-13


def test_solve_q_single_mode_diffusion():
    shape, cell, ds, steps, m, eps = (16, 4, 2), (0.2, 0.2, 0.2), 0.01, 25, 2, 0.3
    L = shape[0] * cell[0]
    x = np.arange(shape[0])[:, None, None] * cell[0]
    mode = np.cos(2 * np.pi * m * x / L) * np.ones(shape)
    kf = O.kfac(shape, cell, ds)
    q = O.solve_q(kf, np.ones(shape), 1.0 + eps * mode, steps)
    km = (2 * np.pi * m / L) ** 2
    for s in (0, 1, 7, steps):
        assert rel(q[s], 1.0 + eps * np.exp(-s * ds * km) * mode) < 1e-13


def _params(shape=(8, 6, 4), cell=(0.4, 0.3, 0.5), fA=0.3, ds=0.05, chiN=15.0):
    return O.Params(*shape, *cell, ds=ds, fA=fA, chiN=chiN)


def _fields(P, sigma=0.5, seed=5):
    rng = np.random.default_rng(seed)
    return rng.normal(0, sigma, P.shape), rng.normal(0, sigma, P.shape)


def test_homogeneous_Q_closed_form():
    P = _params()
    a, b = 0.9, -0.4
    phiA, phiB, Q = O.diblock_densities(P, np.full(P.shape, a), np.full(P.shape, b))
    assert abs(Q / np.exp(-a * P.fA - b * (1 - P.fA)) - 1) < 1e-13


@pytest.mark.parametrize("fA,ds,wantA,wantB", [
    (0.30, 0.01, 0.299997, 0.699993),      # M = 31 and 71 points (odd): 0.99999 f
    (0.35, 0.01, 0.3499966, 0.6499936),    # M = 36 and 66 (even): Simpson + trapezoid tail
    (0.32, 0.005, 0.3199968, 0.6799932),
])
def test_simpson_quirk_kat(fA, ds, wantA, wantB):
    P = _params(shape=(4, 2, 2), fA=fA, ds=ds)
    phiA, phiB, Q = O.diblock_densities(P, np.zeros(P.shape), np.zeros(P.shape))
    assert abs(Q - 1) < 1e-14
    assert np.allclose(phiA, wantA, rtol=0, atol=1e-12)
    assert np.allclose(phiB, wantB, rtol=0, atol=1e-12)


def test_quadrature_weights_by_hand():
    ds = 0.1
    for M, w in [(5, 0.33333 * ds * np.array([1, 4, 2, 4, 1.0])),
                 (4, 0.33333 * ds * np.array([1, 4, 1, 0.0]) + 0.5 * ds * np.array([0, 0, 1, 1.0])),
                 (7, 0.33333 * ds * np.array([1, 4, 2, 4, 2, 4, 1.0]))]:
        rng = np.random.default_rng(M)
        qs = rng.random((M, 3))
        qts = rng.random((M, 3))
        out = O.quadrature(qs, qts, ds, Q=2.0)
        ref = sum(w[j] * qs[j] * qts[M - 1 - j] for j in range(M)) / 2.0
        assert rel(out, ref) < 1e-14


def test_contour_invariant_and_slice_order():
    P = _params()
    wA, wB = _fields(P)
    phiA, phiB, Q, (qA, qB, qtB, qtA) = O.diblock_densities(P, wA, wB, want_slices=True)
    N = P.N
    sA, sB = qA.shape[0] - 1, qB.shape[0] - 1
    for s in range(sA + 1):
        assert abs(np.sum(qA[s] * qtA[sA - s]) / (N * Q) - 1) < 1e-12
    for s in range(sB + 1):
        assert abs(np.sum(qB[s] * qtB[sB - s]) / (N * Q) - 1) < 1e-12
    # junctions: q_B starts where q_A ends, q+_A starts where q+_B ends
    assert np.array_equal(qB[0], qA[sA]) and np.array_equal(qtA[0], qtB[sB])
    # Q is the mean of q+ at the head of A and also of q at the tail of B
    assert abs(qtA[sA].mean() / Q - 1) < 1e-14
    assert abs(qB[sB].mean() / Q - 1) < 1e-12


def test_symmetric_diblock_mirror():
    # fA = 1/2 and wA = wB: forward and backward sweeps coincide
    P = _params(fA=0.5, ds=0.05)
    w, _ = _fields(P)
    phiA, phiB, Q, (qA, qB, qtB, qtA) = O.diblock_densities(P, w, w, want_slices=True)
    assert rel(qA, qtB) < 1e-14 and rel(qB, qtA) < 1e-14
    assert rel(phiA, phiB) < 1e-13


def test_translation_equivariance():
    P = _params()
    wA, wB = _fields(P)
    phiA, phiB, Q = O.diblock_densities(P, wA, wB)
    sh = (3, 1, 2)
    phiA2, phiB2, Q2 = O.diblock_densities(P, np.roll(wA, sh, (0, 1, 2)), np.roll(wB, sh, (0, 1, 2)))
    assert abs(Q2 / Q - 1) < 1e-13
    assert rel(phiA2, np.roll(phiA, sh, (0, 1, 2))) < 1e-12


def test_ds_convergence_second_order():
    # Richardson: the split-step propagator error is O(ds^2)
    # (smooth fields, so that ds k^2 is small for every populated mode)
    cell = (0.4, 0.3, 0.5)
    x = np.arange(8)[:, None, None] / 8.0
    y = np.arange(6)[None, :, None] / 6.0
    wA = 1.5 * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.ones((8, 6, 4))
    wB = -0.7 * np.sin(2 * np.pi * y) * np.ones((8, 6, 4))
    Qs = []
    for ds in (0.02, 0.01, 0.005):
        P = O.Params(8, 6, 4, *cell, ds=ds, fA=0.5, chiN=10.0)
        Qs.append(O.diblock_densities(P, wA, wB)[2])
    ratio = (Qs[0] - Qs[1]) / (Qs[1] - Qs[2])
    assert 3.6 < ratio < 4.4


# ------------------------------------------------------- update and F
def test_sd_update_and_free_energy_by_hand():
    P = _params(shape=(2, 2, 1), cell=(0.5, 0.5, 0.5), fA=0.3, ds=0.1, chiN=12.0)
    P.lambda0, P.lambda1 = 0.2, 0.1
    wA = np.array([0.3, -0.1, 0.2, 0.4]).reshape(2, 2, 1)
    wB = np.array([-0.2, 0.5, 0.1, 0.0]).reshape(2, 2, 1)
    p = np.array([0.05, -0.02, 0.01, -0.04]).reshape(2, 2, 1)
    phiA = np.array([0.25, 0.35, 0.30, 0.28]).reshape(2, 2, 1)
    phiB = np.array([0.74, 0.66, 0.70, 0.71]).reshape(2, 2, 1)
    Q = 1.37
    fA, fB, X = 0.3, 0.7, 12.0
    dHA = p - wA + X * (phiB - fB)
    dHB = p - wB + X * (phiA - fA)
    nA = wA + 0.2 * (dHA - 0.1 * dHB)
    nB = wB + 0.2 * (dHB - 0.1 * dHA)
    pn = 0.5 * (nA + nB)
    pn -= pn.mean()
    F_ref = -np.log(Q) + np.mean(-phiA * nA - phiB * nB) + np.mean(X * phiA * phiB) - X * fA * fB
    a, b, pp = wA.copy(), wB.copy(), p.copy()
    F = O.sd_update(P, a, b, pp, phiA, phiB, Q)
    assert rel(a, nA) < 1e-15 and rel(b, nB) < 1e-15 and rel(pp, pn) < 1e-14
    assert abs(F - F_ref) < 1e-14
    assert abs(pp.mean()) < 1e-16


def test_free_energy_order():
    # F(reference order) - F(consistent) = -(1/N) sum(phiA dwA + phiB dwB)
    P = _params()
    P.lambda0 = 0.1
    wA, wB = _fields(P)
    r1 = O.scft_run(P, wA, wB, n_iters=1)
    P.f_order = 1
    r2 = O.scft_run(P, wA, wB, n_iters=1)
    dwA, dwB = r1["wA"] - wA, r1["wB"] - wB
    want = -np.mean(r1["phiA"] * dwA + r1["phiB"] * dwB)
    assert abs((r1["F"][0] - r2["F"][0]) - want) < 1e-13


def test_homogeneous_free_energy_closed_form():
    P = _params(fA=0.3, ds=0.01)
    P.lambda0 = 0.0          # no update: F of the given fields
    a, b = 0.6, -0.2
    r = O.scft_run(P, np.full(P.shape, a), np.full(P.shape, b), n_iters=1)
    phA, phB = 0.299997, 0.699993   # 0.99999 f (odd point counts)
    want = -(-a * 0.3 - b * 0.7) + (-phA * a - phB * b) + 15.0 * (phA * phB - 0.3 * 0.7)
    assert abs(r["F"][0] - want) < 1e-12


def test_picard_is_steepest_descent_lambda1_zero():
    P = _params()
    wA, wB = _fields(P)
    P.lambda0, P.lambda1 = 0.15, 0.0
    r_sd = O.scft_run(P, wA, wB, n_iters=4)
    P.mix, P.depth = 1, 0
    r_pi = O.scft_run(P, wA, wB, n_iters=4)
    assert np.array_equal(r_sd["F"], r_pi["F"])
    assert np.array_equal(r_sd["wA"], r_pi["wA"])


def test_anderson_reaches_picard_fixed_point():
    P = _params(shape=(8, 4, 2), fA=0.4, ds=0.05, chiN=8.0)
    wA, wB = _fields(P, sigma=0.3)
    P.lambda0 = 0.5
    r_pi = O.scft_run(P, wA, wB, n_iters=400, tol=1e-10)
    P.mix, P.depth, P.beta, P.anderson_start = 1, 5, 0.5, 1e-1
    r_an = O.scft_run(P, wA, wB, n_iters=400, tol=1e-10)
    assert r_pi["err"][-1] < 1e-10 and r_an["err"][-1] < 1e-10
    assert r_an["iters"] < r_pi["iters"]
    assert abs(r_an["F"][-1] - r_pi["F"][-1]) < 1e-10
    # (this fixed point is the disordered melt, w nearly uniform: compare absolutely)
    assert np.max(np.abs(r_an["wA"] - r_pi["wA"])) < 1e-8
    assert np.max(np.abs(r_an["wB"] - r_pi["wB"])) < 1e-8


def test_anderson_coefficients_match_least_squares():
    # two Anderson iterations by hand: c = argmin |f_k - sum c_i (f_k - f_{k-i})|
    P = _params(shape=(4, 4, 2), fA=0.4, ds=0.05, chiN=8.0)
    wA, wB = _fields(P, sigma=0.3)
    P.mix, P.depth, P.beta, P.lambda0 = 1, 2, 0.3, 0.2
    r1 = O.scft_run(P, wA, wB, n_iters=1)          # Picard step (no history)
    r2 = O.scft_run(P, wA, wB, n_iters=2)
    r3 = O.scft_run(P, wA, wB, n_iters=3)
    # rebuild iterate 3 by hand from the residuals of iterates 1..3
    def resid(w_a, w_b, p):
        phiA, phiB, Q = O.diblock_densities(P, w_a, w_b)
        return np.concatenate([(p - w_a + 8.0 * (phiB - 0.6)).ravel(),
                               (p - w_b + 8.0 * (phiA - 0.4)).ravel()])
    x0 = np.concatenate([wA.ravel(), wB.ravel()])
    x1 = np.concatenate([r1["wA"].ravel(), r1["wB"].ravel()])
    x2 = np.concatenate([r2["wA"].ravel(), r2["wB"].ravel()])
    f0 = resid(wA, wB, np.zeros(P.shape))
    f1 = resid(r1["wA"], r1["wB"], r1["p"])
    f2 = resid(r2["wA"], r2["wB"], r2["p"])
    D = np.stack([f2 - f1, f2 - f0], axis=1)
    c, *_ = np.linalg.lstsq(D, f2, rcond=None)
    xb = x2 + c[0] * (x1 - x2) + c[1] * (x0 - x2)
    fb = f2 + c[0] * (f1 - f2) + c[1] * (f0 - f2)
    x3 = xb + 0.3 * fb
    got = np.concatenate([r3["wA"].ravel(), r3["wB"].ravel()])
    assert rel(got, x3) < 1e-10


# ------------------------------------------------------------- stress
def test_box_stress_matches_finite_differences():
    shape = (8, 6, 4)
    cell = np.array([0.45, 0.4, 0.5])
    P = O.Params(*shape, *cell, ds=0.05, fA=0.3, chiN=12.0)
    rng = np.random.default_rng(21)
    wA, wB = rng.normal(0, 0.8, shape), rng.normal(0, 0.8, shape)
    g = O.box_stress(P, wA, wB)
    h = 1e-5
    for d in range(3):
        c1, c2 = cell.copy(), cell.copy()
        c1[d] += h / shape[d]
        c2[d] -= h / shape[d]
        Q1 = O.diblock_densities(O.Params(*shape, *c1, ds=0.05, fA=0.3, chiN=12.0), wA, wB)[2]
        Q2 = O.diblock_densities(O.Params(*shape, *c2, ds=0.05, fA=0.3, chiN=12.0), wA, wB)[2]
        fd = -(np.log(Q1) - np.log(Q2)) / (2 * h)
        assert abs(g[d] - fd) < 1e-7 * max(1.0, abs(fd)), (d, g[d], fd)


def test_box_stress_zero_for_homogeneous_fields():
    P = _params()
    g = O.box_stress(P, np.full(P.shape, 0.3), np.full(P.shape, -0.1))
    assert np.max(np.abs(g)) < 1e-14
