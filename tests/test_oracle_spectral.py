"""The oracle's restatement of the FFT plugin's other consumers, pinned on
the CPU: the rest of PsFFTBase against numpy's FFT (pocketfft) and the brute
DFT; the spectral filters and the Poisson updater against closed forms on
single-mode fields, invariants (idempotence, one region = the simple
filter), and a numpy restatement of the filter on a square grid, where the
reference's transposed region map is a relabelling."""
import numpy as np
import pytest

import oracle as O

SHAPES = [(4, 6, 8), (8, 8, 1), (5, 1, 1), (6, 5, 3)]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("shape", SHAPES)
def test_fft_family_against_numpy(shape):
    rng = np.random.default_rng(1)
    x, y, k = (rng.standard_normal(shape) for _ in range(3))
    F = np.fft.fftn(x)
    N = x.size
    assert rel(O.forward_fft_abs(x), np.abs(F) ** 2) < 1e-13
    assert rel(O.forward_fft_re(x), F.real) < 1e-13
    assert rel(O.backward_fft_re(x), (np.fft.ifftn(x) * N).real) < 1e-13
    assert rel(O.convolve_re(x, y), (np.fft.ifftn(F * np.fft.fftn(y)) * N).real) < 1e-13
    assert rel(O.scaled_fft_pair_im(x, k), (np.fft.ifftn(k * np.fft.fftn(1j * x)) * N).real) < 1e-13


def test_fft_family_against_brute_dft():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 4, 2))
    B = O.dft3_brute(x.astype(complex), -1)
    assert rel(O.forward_fft_abs(x), np.abs(B) ** 2) < 1e-13
    assert rel(O.forward_fft_re(x), B.real) < 1e-13


def test_k2_closed_form():
    shape, cell = (6, 4, 1), (0.3, 0.5, 0.7)
    k2 = O.k2(shape, cell)
    n = [np.minimum(np.arange(m), m - np.arange(m)) for m in shape]   # wrapped index
    kk = [2 * np.pi * n[d] / (shape[d] * cell[d]) for d in range(3)]
    ref = kk[0][:, None, None] ** 2 + kk[1][None, :, None] ** 2 + kk[2][None, None, :] ** 2
    assert rel(k2, ref) < 1e-14


def _two_modes(shape, a, b, c0=0.7):
    x = np.arange(shape[0])[:, None, None] * np.ones(shape)
    return c0 + a * np.cos(2 * np.pi * x / shape[0]) + b * np.cos(2 * np.pi * 3 * x / shape[0])


def test_simple_filter_keeps_the_strong_mode():
    """w = c + a cos(k1 x) + b cos(k3 x): |F|^2 at k1 is (a N/2)^2, at k3
    (b N/2)^2; a cutoff between (b/a)^2 and 1 with strength 0 leaves
    a cos(k1 x) (the mean goes first)."""
    shape = (16, 4, 2)
    w = _two_modes(shape, 1.0, 0.5)
    out = O.spec_filter(w, 0.5, 0.0)
    assert rel(out, _two_modes(shape, 1.0, 0.0, 0.0)) < 1e-13
    # strength 0.3 scales the weak mode by 0.3
    assert rel(O.spec_filter(w, 0.5, 0.3), _two_modes(shape, 1.0, 0.15, 0.0)) < 1e-13
    # strength 1: only the mean goes
    assert rel(O.spec_filter(w, 0.5, 1.0), w - w.mean()) < 1e-13
    # cutoff <= 0: untouched, mean included
    assert np.array_equal(O.spec_filter(w, 0.0, 0.0), w)


def test_filter_wall_volume_and_idempotence():
    rng = np.random.default_rng(3)
    w = rng.standard_normal((8, 6, 4))
    once = O.spec_filter(w, 0.3, 0.0)
    assert rel(O.spec_filter(once, 0.3, 0.0), once) < 1e-12
    # a wall volume V changes the shift to sum / (N - V)
    V = 10.0
    strength1 = O.spec_filter(w, 0.3, 1.0, vwall=V)
    assert rel(strength1, w - w.sum() / (w.size - V)) < 1e-13


def test_multi_filter_one_region_is_the_simple_filter():
    rng = np.random.default_rng(4)
    w = rng.standard_normal((8, 6, 4))
    assert rel(O.spec_filter(w, 0.4, 0.2, cells=(8, 6, 4)), O.spec_filter(w, 0.4, 0.2)) < 1e-14


def test_multi_filter_regions():
    """Two modes in different spectral cells each are their cell's maximum and
    survive a cutoff that the simple filter applies to the weaker one."""
    shape = (16, 16, 1)
    x = np.arange(16)[:, None, None] * np.ones(shape)
    y = np.arange(16)[None, :, None] * np.ones(shape)
    w = np.cos(2 * np.pi * x / 16) + 0.4 * np.cos(2 * np.pi * 6 * y / 16)
    simple = O.spec_filter(w, 0.5, 0.0)
    assert rel(simple, np.cos(2 * np.pi * x / 16)) < 1e-13
    multi = O.spec_filter(w, 0.5, 0.0, cells=(4, 4, 1))
    assert rel(multi, w) < 1e-13
    with pytest.raises(ValueError):
        O.spec_filter(w, 0.5, 0.0, cells=(3, 4, 1))


def test_multi_filter_square_grid_against_numpy():
    """On a square grid with equal x, y spectral cells the reference's region
    map is a relabelling of the (i / sx, j / sy, k / sz) blocks."""
    rng = np.random.default_rng(5)
    shape, cells, c, s = (8, 8, 4), (2, 2, 2), 0.3, 0.25
    w = rng.standard_normal(shape)
    v = w - w.sum() / w.size
    P = np.abs(np.fft.fftn(v)) ** 2
    nb = [shape[d] // cells[d] for d in range(3)]
    blocks = P.reshape(nb[0], cells[0], nb[1], cells[1], nb[2], cells[2])
    bmax = blocks.max(axis=(1, 3, 5))
    cut = c * np.repeat(np.repeat(np.repeat(bmax, cells[0], 0), cells[1], 1), cells[2], 2)
    mask = np.where(P < cut, s, 1.0)
    ref = np.fft.ifftn(mask * np.fft.fftn(v)).real
    assert rel(O.spec_filter(w, c, s, cells=cells), ref) < 1e-12


def test_poisson_closed_form_and_limiter():
    """rho = q0 + A cos(k x): the charge loses q0 and psi becomes
    4 * 3.1415 * lB / k^2 * A cos(k x) where the old psi is within 20%; where
    it is not, 1.2 psi_old / N (the reference's limiter as written); psi_old
    = 0 stays 0."""
    shape, cell, lB, A = (16, 2, 1), (0.25, 0.5, 1.0), 0.7, 0.3
    N = 32
    x = np.arange(16)[:, None, None] * np.ones(shape)
    mode = np.cos(2 * np.pi * x / 16)
    k1 = 2 * np.pi / (16 * 0.25)
    exact = 4 * 3.1415 * lB / k1 ** 2 * A * mode
    charge = 0.9 + A * mode
    c2, psi = O.poisson_update(charge, exact * 1.1, lB, cell)
    assert rel(c2, A * mode) < 1e-13
    assert rel(psi, exact) < 1e-12
    _, psi = O.poisson_update(charge, exact * 2.0, lB, cell)
    assert rel(psi, 1.2 * 2.0 * exact / N) < 1e-13
    _, psi = O.poisson_update(charge, np.zeros(shape), lB, cell)
    assert np.all(psi == 0.0)
