// psx_rng.hpp -- the reference's host random stream (host code only).
//
// PsRandom (psstd/PsRandom.h:147-170,305-373, psstd/PsRandom.cpp:58-95):
// a 32-bit Mersenne Twister MT19937 (Matsumoto and Nishimura, ACM TOMACS 8
// (1998) 3-30) seeded by setSeed -> init_genrand, uniform reals y / 2^32 in
// [0, 1) (real() on LP64), and Gaussian deviates by Marsaglia's polar
// method returning v2 * fac and keeping v1 * fac for the next call
// (psRandomGaussian).  One stream serves the random initial conjugate
// fields (PsGridField::resetRandom, psbase/PsGridField.cpp:389-402) and the
// steepest-descent noise (PsGridField::addRandom, 404-418).
#pragma once
#include <cmath>
#include <cstdint>

struct psx_rng {
  static constexpr int kN = 624, kM = 397;
  uint32_t mt[kN];
  int idx = kN;
  bool have_spare = false;   // iset
  double spare = 0.0;        // gset

  void seed(uint32_t s) {
    mt[0] = s;
    for (int i = 1; i < kN; ++i)
      mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    idx = kN;
    have_spare = false;
    spare = 0.0;
  }

  void refill() {
    for (int k = 0; k < kN; ++k) {
      const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % kN] & 0x7fffffffu);
      mt[k] = mt[(k + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    idx = 0;
  }

  uint32_t next32() {
    if (idx >= kN) refill();
    uint32_t y = mt[idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }

  double uniform() { return (double)next32() * (1.0 / 4294967296.0); }

  double gaussian() {
    if (have_spare) {
      have_spare = false;
      return spare;
    }
    double v1, v2, r;
    do {
      v1 = (2.0 * uniform()) - 1.0;
      v2 = (2.0 * uniform()) - 1.0;
      r = (v1 * v1) + (v2 * v2);
    } while (r >= 1.0 || r == 0.0);
    const double fac = std::sqrt(-2.0 * std::log(r) / r);
    spare = v1 * fac;
    have_spare = true;
    return v2 * fac;
  }
};
