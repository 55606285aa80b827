// psx_demo.cpp -- C++ host program over the C ABI: the steepest-descent
// diblock run of the reference's own regression deck shape (2D melt, Picard
// relaxation), printing F every 10 iterations like PsFreeEnergy's history,
// then a simple spectral filter of the final w_A through the FFT plugin
// (its sum of squares and the largest |F|^2 of the filtered field).
//
//   psx_demo [nx ny nz fA chiN ds iters]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "psx_host.hpp"

int main(int argc, char** argv) {
  int nx = argc > 1 ? std::atoi(argv[1]) : 64;
  int ny = argc > 2 ? std::atoi(argv[2]) : 64;
  int nz = argc > 3 ? std::atoi(argv[3]) : 1;
  double fA = argc > 4 ? std::atof(argv[4]) : 0.30;
  double chiN = argc > 5 ? std::atof(argv[5]) : 20.0;
  double ds = argc > 6 ? std::atof(argv[6]) : 0.01;
  int iters = argc > 7 ? std::atoi(argv[7]) : 200;
  try {
    psx_grid g{nx, ny, nz, 4.0 / nx, 4.0 / ny, 4.0 / nx};
    psx_diblock c{fA, ds, 1.0, 1.0, chiN};
    psx::DiblockSolver s(g, c, PSX_FP64);
    std::vector<double> wA(s.size()), wB(s.size());
    for (size_t i = 0; i < s.size(); ++i) {  // lamellar seed along x
      const double x = (double)(i / ((size_t)ny * nz)) / nx;
      wA[i] = std::cos(2.0 * M_PI * x);
      wB[i] = -wA[i];
    }
    s.setFields(wA, wB);
    psx_mixing mix{0, 0.1, 0.0, 0, 0.0, 1e300, 0};
    std::vector<double> F, err;
    int done = s.iterate(mix, iters, 0.0, &F, &err);
    for (int i = 0; i < done; i += 10) std::printf("step %4d  F = %.12f  err = %.3e\n", i, F[i], err[i]);
    std::vector<double> p;
    s.getFields(wA, wB, p);
    psx::HipFFT fft(nx, ny, nz);
    fft.specFilter(wA.data(), 0.5, 0.0);
    std::vector<double> spec(s.size());
    fft.forwardFFTAbs(wA.data(), spec.data());
    double ss = 0.0, mx = 0.0;
    for (size_t i = 0; i < s.size(); ++i) { ss += wA[i] * wA[i]; mx = std::fmax(mx, spec[i]); }
    std::printf("filter sumsq = %.15e  max|F|^2 = %.15e\n", ss, mx);
    std::printf("final F = %.12f after %d iterations\n", F.back(), done);
  } catch (const psx::Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
