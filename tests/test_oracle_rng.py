"""The reference's random stream (PsRandom) on the CPU: the oracle's MT19937
against the C++ standard's known answer and numpy's legacy generator (the
same algorithm and seeding), its Gaussians against Marsaglia's polar method
computed here from the raw words, and the library's host stream (the one the
engine's noise and the deck runner's initial fields use) equal to the
oracle's, bit for bit.  Host code only: no GPU."""
import math

import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd.scft import Rng


def test_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th output of std::mt19937 (seed 5489) is 4123659995
    assert O.Rng(5489).raw32(10000)[-1] == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 4242, 12122, 2 ** 32 + 7])
def test_mt19937_against_numpy_legacy(seed):
    """numpy's RandomState(seed) seeds MT19937 with init_genrand too; its
    random_sample is (a * 2^26 + b) / 2^53 of two words (a >> 5, b >> 6)."""
    raw = O.Rng(seed).raw32(4000).astype(np.uint64)
    a, b = raw[0::2] >> 5, raw[1::2] >> 6
    res53 = (a * 67108864.0 + b) / 9007199254740992.0
    assert np.array_equal(res53, np.random.RandomState(seed % 2 ** 32).random_sample(2000))


def test_uniform_is_word_over_2_32():
    raw = O.Rng(77).raw32(1000)
    assert np.array_equal(O.Rng(77).uniform(1000), raw.astype(np.float64) / 4294967296.0)


def test_gaussian_is_marsaglia_polar_on_the_words():
    """psRandomGaussian: pairs (v1, v2) = 2u - 1 until 0 < r < 1, then v2 fac
    and, at the next call, v1 fac, fac = sqrt(-2 ln r / r)."""
    u = O.Rng(99).uniform(4000)
    want, i = [], 0
    while len(want) < 1000:
        v1, v2 = 2.0 * u[i] - 1.0, 2.0 * u[i + 1] - 1.0
        i += 2
        r = v1 * v1 + v2 * v2
        if r >= 1.0 or r == 0.0:
            continue
        fac = math.sqrt(-2.0 * math.log(r) / r)
        want += [v2 * fac, v1 * fac]
    got = O.Rng(99).gaussian(1000)
    assert np.array_equal(got, np.array(want[:1000]))
    g = O.Rng(5).gaussian(200000)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.01


@pytest.mark.parametrize("seed", [0, 4242, 2 ** 32 + 7])
def test_library_stream_equals_oracle(seed):
    lib, ref = Rng(seed), O.Rng(seed)
    # interleave as a run does: initial uniforms, then Gaussians with a
    # pending spare across calls of odd length
    for n_u, n_g in [(1001, 333), (0, 7), (5, 1000), (0, 1)]:
        assert np.array_equal(lib.uniform(n_u), ref.uniform(n_u))
        assert np.array_equal(lib.gaussian(n_g), ref.gaussian(n_g))
    lib.close()
