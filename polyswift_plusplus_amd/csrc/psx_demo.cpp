// psx_demo.cpp -- C++ host program over the C ABI through psx_host.hpp.
//
//   psx_demo [nx ny nz fA chiN ds iters]
//     the steepest-descent diblock run of the reference's own regression
//     deck shape (2D melt, Picard relaxation), printing F every 10
//     iterations like PsFreeEnergy's history, then a simple spectral filter
//     of the final w_A through the FFT plugin (its sum of squares and the
//     largest |F|^2 of the filtered field).
//   psx_demo adjoint n
//     an n^3 diblock after 5 steepest-descent steps: Q, the box stress, and
//     the reverse mode of the densities op with g_phi = 0, g_Q = -1/Q (whose
//     box part is the box stress and whose field part sums to -N Q'/Q).
//   psx_demo graph
//     an ABA triblock (a block graph) on a 2D grid, 10 steepest-descent steps.
//   psx_demo blend
//     a diblock with an A homopolymer of half its length (volfracs 0.8 /
//     0.2) on a 2D grid, 10 steepest-descent steps, both polymers' Q.
//   psx_demo batch
//     three diblocks of two grid sizes advanced together in FP32 with
//     Anderson mixing, 6 iterations, the last F of each.
// Every figure is printed with its label so a test can read it back.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "psx_host.hpp"

namespace {

// lamellar + a weaker y mode: w_A = cos(2 pi x / Lx) + a cos(2 pi y / Ly), w_B = -w_A
void seed(const psx_grid& g, double a, std::vector<double>& wA, std::vector<double>& wB) {
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  wA.assign(n, 0.0);
  wB.assign(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const size_t ix = i / ((size_t)g.ny * g.nz), iy = (i / g.nz) % g.ny;
    wA[i] = std::cos(2.0 * M_PI * ix / g.nx) + a * std::cos(2.0 * M_PI * iy / g.ny);
    wB[i] = -wA[i];
  }
}

psx_block block(int species, double frac, std::vector<int> head, std::vector<int> tail, double b = 1.0) {
  psx_block k;
  std::memset(&k, 0, sizeof(k));
  k.species = species;
  k.lengthfrac = frac;
  k.bSegRatio = b;
  k.nhead = (int)head.size();
  k.ntail = (int)tail.size();
  for (size_t i = 0; i < head.size(); ++i) k.head[i] = head[i];
  for (size_t i = 0; i < tail.size(); ++i) k.tail[i] = tail[i];
  return k;
}

const psx_mixing kSteepest{0, 0.1, 0.0, 0, 0.0, 1e300, 0};

int melt(int argc, char** argv) {
  int nx = argc > 1 ? std::atoi(argv[1]) : 64;
  int ny = argc > 2 ? std::atoi(argv[2]) : 64;
  int nz = argc > 3 ? std::atoi(argv[3]) : 1;
  double fA = argc > 4 ? std::atof(argv[4]) : 0.30;
  double chiN = argc > 5 ? std::atof(argv[5]) : 20.0;
  double ds = argc > 6 ? std::atof(argv[6]) : 0.01;
  int iters = argc > 7 ? std::atoi(argv[7]) : 200;
  psx_grid g{nx, ny, nz, 4.0 / nx, 4.0 / ny, 4.0 / nx};
  psx_diblock c{fA, ds, 1.0, 1.0, chiN};
  psx::DiblockSolver s(g, c, PSX_FP64);
  std::vector<double> wA, wB;
  seed(g, 0.0, wA, wB);
  s.setFields(wA, wB);
  std::vector<double> F, err;
  int done = s.iterate(kSteepest, iters, 0.0, &F, &err);
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
  return 0;
}

int adjoint(int argc, char** argv) {
  const int n = argc > 2 ? std::atoi(argv[2]) : 16;
  psx_grid g{n, n, n, 3.2 / n, 3.2 / n, 3.2 / n};
  psx_diblock c{0.30, 0.02, 1.0, 1.0, 15.0};
  psx::DiblockSolver s(g, c, PSX_FP64);
  std::vector<double> wA, wB;
  seed(g, 0.5, wA, wB);
  s.setFields(wA, wB);
  s.iterate(kSteepest, 5, 0.0);
  s.solveQ();
  const double Q = s.calcBigQ();
  const std::vector<double> st = s.boxStress();
  std::vector<double> gwA, gwB, gL;
  s.backward({}, {}, -1.0 / Q, gwA, gwB, &gL);
  double sA = 0.0, sB = 0.0;
  for (size_t i = 0; i < s.size(); ++i) { sA += gwA[i]; sB += gwB[i]; }
  std::printf("Q = %.15e\n", Q);
  std::printf("stress = %.15e %.15e %.15e\n", st[0], st[1], st[2]);
  std::printf("backward box = %.15e %.15e %.15e\n", gL[0], gL[1], gL[2]);
  std::printf("backward sums = %.15e %.15e\n", sA, sB);
  return 0;
}

int graph(int, char**) {
  psx_grid g{32, 32, 1, 0.125, 0.125, 0.125};
  // A(0.25) - B(0.5) - A(0.25): block 1's head joins block 0's tail, its tail block 2's head
  std::vector<psx_block> b{block(0, 0.25, {}, {1}), block(1, 0.5, {0}, {2}), block(0, 0.25, {1}, {})};
  psx::BlockCopolymerSolver s(g, b, 0.01, 20.0, PSX_FP64);
  std::vector<double> wA, wB;
  seed(g, 0.3, wA, wB);
  s.setFields(wA, wB);
  std::vector<double> F;
  s.iterate(kSteepest, 10, 0.0, &F);
  s.solveQ();
  std::printf("graph F = %.15e  Q = %.15e\n", F.back(), s.calcBigQ());
  return 0;
}

int blend(int, char**) {
  psx_grid g{32, 32, 1, 0.125, 0.125, 0.125};
  std::vector<psx_polymer> p{{0.8, 1.0, 2}, {0.2, 0.5, 1}};
  std::vector<psx_block> b{block(0, 0.4, {}, {1}), block(1, 0.6, {0}, {}), block(0, 1.0, {}, {})};
  psx::BlendSolver s(g, p, b, 0.01, 18.0, PSX_FP64);
  std::vector<double> wA, wB;
  seed(g, 0.3, wA, wB);
  s.setFields(wA, wB);
  std::vector<double> F;
  s.iterate(kSteepest, 10, 0.0, &F);
  s.solveQ();
  const std::vector<double> q = s.polymerQ();
  std::printf("blend F = %.15e  Q = %.15e %.15e\n", F.back(), q.at(0), q.at(1));
  return 0;
}

int batch(int, char**) {
  std::vector<psx_grid> g{{16, 16, 16, 0.2, 0.2, 0.2}, {12, 12, 12, 0.25, 0.25, 0.25}, {16, 16, 16, 0.2, 0.2, 0.2}};
  std::vector<psx_diblock> c{{0.30, 0.05, 1.0, 1.0, 15.0}, {0.40, 0.05, 1.0, 1.0, 18.0}, {0.45, 0.05, 1.0, 1.0, 20.0}};
  psx::Batch b(g, c, PSX_FP32);
  for (size_t i = 0; i < b.size(); ++i) {
    std::vector<double> wA, wB;
    seed(g[i], 0.2 * (double)i, wA, wB);
    b.instance(i).setFields(wA, wB);
  }
  const psx_mixing mix{1, 0.2, 0.0, 5, 1.0, 0.05, 0};
  std::vector<std::vector<double>> F;
  const int done = b.iterate(mix, 6, &F);
  std::printf("batch classes = %d  iterations = %d\n", b.sizeClasses(), done);
  for (size_t i = 0; i < b.size(); ++i) std::printf("batch F %zu = %.15e\n", i, F[i].back());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const char* mode = argc > 1 ? argv[1] : "";
    if (!std::strcmp(mode, "adjoint")) return adjoint(argc, argv);
    if (!std::strcmp(mode, "graph")) return graph(argc, argv);
    if (!std::strcmp(mode, "blend")) return blend(argc, argv);
    if (!std::strcmp(mode, "batch")) return batch(argc, argv);
    return melt(argc, argv);
  } catch (const psx::Error& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
