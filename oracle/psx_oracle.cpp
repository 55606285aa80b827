/*
 * psx_oracle.cpp -- CPU oracle for the SCFT hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * "Faithful" pass structure: every element-wise operation is its own loop and
 * the FFT pair is a full-grid complex-to-complex transform, as the reference
 * does it (psfft/PsFFTW.cpp:196-224, pspolymer/PsFlexPseudoSpec.cpp:267-289).
 * OpenMP threads (orc_set_threads) parallelise the FFT lines and the loops;
 * with one thread it is the single-rank reference pass structure.
 */
#include "psx_oracle.h"

#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;

static int g_threads = 1;

extern "C" void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
extern "C" int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* 1D mixed-radix FFT: recursive decimation in time, generic radix combine. */
/* ------------------------------------------------------------------------ */
namespace {

struct Plan1D {
  int n = 0;
  std::vector<int> fac;   // radices, product = n
  std::vector<cd> tw;     // tw[j] = exp(-2 pi i j / n)
};

std::vector<int> factorize(int n) {
  std::vector<int> f;
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  for (int p = 3; n > 1; p += 2) {
    while (n % p == 0) { f.push_back(p); n /= p; }
    if (p * p > n && n > 1) { f.push_back(n); break; }
  }
  return f;
}

const Plan1D& get_plan(int n) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<Plan1D>> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return *it->second;
  auto p = std::make_unique<Plan1D>();
  p->n = n;
  p->fac = factorize(n);
  p->tw.resize(n);
  const long double twopi = 6.283185307179586476925286766559L;
  for (int j = 0; j < n; ++j) {
    long double a = twopi * (long double)j / (long double)n;
    p->tw[j] = cd((double)cosl(a), (double)-sinl(a));
  }
  const Plan1D& ref = *p;
  cache[n] = std::move(p);
  return ref;
}

// out[0..n) = DFT of in[j*is], j < n.  fs = stride into the size-N0 twiddles.
void fft_work(cd* out, const cd* in, long is, int fs, const int* fac, int n,
              const Plan1D& P, bool inverse) {
  const int p = fac[0];
  const int m = n / p;
  if (m == 1) {
    for (int r = 0; r < p; ++r) out[r] = in[r * is];
  } else {
    for (int r = 0; r < p; ++r)
      fft_work(out + r * m, in + r * is, is * p, fs * p, fac + 1, m, P, inverse);
  }
  // out[r*m + k] holds Y_r[k]; X[k + q m] = sum_r W_n^{r(k + q m)} Y_r[k]
  const int N0 = P.n;
  cd t[64];
  std::vector<cd> tbig;
  cd* tp = t;
  if (p > 64) { tbig.resize(p); tp = tbig.data(); }
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      cd w = P.tw[((long)fs * r * k) % N0];
      if (inverse) w = std::conj(w);
      tp[r] = out[r * m + k] * w;
    }
    for (int q = 0; q < p; ++q) {
      cd acc = 0.0;
      for (int r = 0; r < p; ++r) {
        cd w = P.tw[((long)fs * m * ((r * q) % p)) % N0];
        if (inverse) w = std::conj(w);
        acc += tp[r] * w;
      }
      out[k + q * m] = acc;
    }
  }
}

void fft1d(cd* out, const cd* in, int n, bool inverse) {
  if (n == 1) { out[0] = in[0]; return; }
  const Plan1D& P = get_plan(n);
  fft_work(out, in, 1, 1, P.fac.data(), n, P, inverse);
}

// Transform every line along `axis` of a C-order complex nx*ny*nz array.
void fft_axis(cd* a, int nx, int ny, int nz, int axis, bool inverse) {
  const int dims[3] = {nx, ny, nz};
  const int n = dims[axis];
  if (n == 1) return;
  long inner = 1;
  for (int d = axis + 1; d < 3; ++d) inner *= dims[d];
  long outer = 1;
  for (int d = 0; d < axis; ++d) outer *= dims[d];
  const long nlines = outer * inner;
  get_plan(n);  // build outside the parallel region
#pragma omp parallel num_threads(g_threads)
  {
    std::vector<cd> buf(n), res(n);
#pragma omp for schedule(static)
    for (long l = 0; l < nlines; ++l) {
      const long o = l / inner, i = l % inner;
      cd* base = a + o * (long)n * inner + i;
      for (int k = 0; k < n; ++k) buf[k] = base[(long)k * inner];
      fft1d(res.data(), buf.data(), n, inverse);
      for (int k = 0; k < n; ++k) base[(long)k * inner] = res[k];
    }
  }
}

void fft3(cd* a, int nx, int ny, int nz, bool inverse) {
  fft_axis(a, nx, ny, nz, 2, inverse);
  fft_axis(a, nx, ny, nz, 1, inverse);
  fft_axis(a, nx, ny, nz, 0, inverse);
}

long ncells(const orc_params* P) { return (long)P->nx * P->ny * P->nz; }

}  // namespace

/* ------------------------------------------------------------------------ */
extern "C" int orc_block_steps(double bs_len, double ds) {
  // psbase/PsBlockBase.cpp:252-326 (forceBlockSteps = false, minBlockSteps 4)
  const double bs = bs_len / ds;
  const double intPart = (double)(long)bs;
  const double dx = bs - intPart;
  if (dx > 0.0001 && dx < 0.9999) return -1;
  const long c = (long)std::ceil(bs), f = (long)std::floor(bs);
  long r;
  if (c == f) r = c;
  else if (std::fabs(bs - (double)c) < std::fabs(bs - (double)f)) r = c;
  else r = f;
  if (r < 4) return -1;
  return (int)r;
}

extern "C" void orc_kfac(int nx, int ny, int nz, double dx, double dy,
                         double dz, double ds, double b, double* kfac) {
  // psbase/PsGridField.cpp:66-104; pspolymer/PsFlexPseudoSpec.cpp:188-206
  const double twopi = 2.0 * std::acos(-1.0);
  const int g[3] = {nx, ny, nz};
  const double h[3] = {dx, dy, dz};
  double n2[3];
  for (int d = 0; d < 3; ++d) n2[d] = (double)g[d] / 2.0;
  long n = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const int idx[3] = {i, j, k};
        double k2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          double nk = n2[d] - std::fabs((double)idx[d] - n2[d]);
          double kv = twopi * nk / (double)g[d];
          kv = kv / h[d];
          k2 += kv * kv;
        }
        kfac[n++] = std::exp(-1.0 * ds * k2 * b * b);
      }
}

extern "C" void orc_k2(int nx, int ny, int nz, double dx, double dy, double dz, double* k2) {
  // psbase/PsGridField.cpp:66-104 (PsGridField::calck2), C order
  const double twopi = 2.0 * std::acos(-1.0);
  const int g[3] = {nx, ny, nz};
  const double h[3] = {dx, dy, dz};
  double n2[3];
  for (int d = 0; d < 3; ++d) n2[d] = (double)g[d] / 2.0;
  long n = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const int idx[3] = {i, j, k};
        double kv[3];
        for (int d = 2; d >= 0; --d) {
          const double nk = n2[d] - std::fabs((double)idx[d] - n2[d]);
          kv[d] = twopi * nk / (double)g[d];
          kv[d] = kv[d] / h[d];
        }
        k2[n++] = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
      }
}

extern "C" int orc_fft3(int nx, int ny, int nz, double* re, double* im,
                        int sign) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const long N = (long)nx * ny * nz;
  std::vector<cd> a(N);
  for (long i = 0; i < N; ++i) a[i] = cd(re[i], im[i]);
  fft3(a.data(), nx, ny, nz, sign > 0);
  for (long i = 0; i < N; ++i) { re[i] = a[i].real(); im[i] = a[i].imag(); }
  return 0;
}

extern "C" int orc_dft3_brute(int nx, int ny, int nz, const double* re,
                              const double* im, double* ore, double* oim,
                              int sign) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const long double twopi = 6.283185307179586476925286766559L;
  for (int a = 0; a < nx; ++a)
    for (int b = 0; b < ny; ++b)
      for (int c = 0; c < nz; ++c) {
        long double sr = 0, si = 0;
        for (int i = 0; i < nx; ++i)
          for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
              // exact phase from integer residues
              long double ph = twopi * (((long)a * i % nx) / (long double)nx +
                                        ((long)b * j % ny) / (long double)ny +
                                        ((long)c * k % nz) / (long double)nz);
              long double cr = cosl(ph), ci = (sign > 0 ? 1 : -1) * sinl(ph);
              long idx = ((long)i * ny + j) * nz + k;
              sr += re[idx] * cr - im[idx] * ci;
              si += re[idx] * ci + im[idx] * cr;
            }
        long o = ((long)a * ny + b) * nz + c;
        ore[o] = (double)sr;
        oim[o] = (double)si;
      }
  return 0;
}

extern "C" int orc_scaled_fft_pair(int nx, int ny, int nz, const double* data,
                                   const double* kdata, double* res) {
  // psfft/PsFFTW.cpp:196-224: pack, forward, scale Re/Im by kdata, backward,
  // keep the real part; no 1/N.
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const long N = (long)nx * ny * nz;
  std::vector<cd> in(N);
#pragma omp parallel for num_threads(g_threads)
  for (long n = 0; n < N; ++n) in[n] = cd(data[n], 0.0);
  fft3(in.data(), nx, ny, nz, false);
#pragma omp parallel for num_threads(g_threads)
  for (long n = 0; n < N; ++n)
    in[n] = cd(in[n].real() * kdata[n], in[n].imag() * kdata[n]);
  fft3(in.data(), nx, ny, nz, true);
#pragma omp parallel for num_threads(g_threads)
  for (long n = 0; n < N; ++n) res[n] = in[n].real();
  return 0;
}

extern "C" int orc_solve_q(int nx, int ny, int nz, const double* kfac,
                           const double* wfac, const double* q0, int steps,
                           double* slices) {
  // pspolymer/PsFlexPseudoSpec.cpp:244-293
  const long N = (long)nx * ny * nz;
  const double scaleFFT = 1.0 / (double)N;
  std::vector<double> qX(q0, q0 + N), qw(N), res(N);
  std::memcpy(slices, qX.data(), N * sizeof(double));
  for (int ss = 1; ss <= steps; ++ss) {
#pragma omp parallel for num_threads(g_threads)
    for (long n = 0; n < N; ++n) qw[n] = qX[n] * wfac[n];
    if (orc_scaled_fft_pair(nx, ny, nz, qw.data(), kfac, res.data())) return -1;
#pragma omp parallel for num_threads(g_threads)
    for (long n = 0; n < N; ++n) qX[n] = res[n] * wfac[n];
#pragma omp parallel for num_threads(g_threads)
    for (long n = 0; n < N; ++n) qX[n] = qX[n] * scaleFFT;
    std::memcpy(slices + (long)ss * N, qX.data(), N * sizeof(double));
  }
  return 0;
}

extern "C" void orc_quadrature(long N, int steps, double ds, const double* qs,
                               const double* qts, double Q, double* out) {
  // pspolymer/PsBlock.cpp:256-380.  M = steps + 1 points; Simpson with the
  // literal 0.33333 over the first M (odd) or M-1 (even) points, then a
  // trapezoid on the last interval when M is even, then the 1/Q scale.
  const long M = steps + 1;
  const bool even = (M % 2) == 0;
  const long numStail = even ? M - 4 : M - 3;
  for (long r = 0; r < N; ++r) out[r] = 0.0;
  for (long n = 0; n <= numStail; n += 2) {
    const double* q1 = qs + n * N;
    const double* q2 = qs + (n + 1) * N;
    const double* q3 = qs + (n + 2) * N;
    const double* t1 = qts + (M - n - 1) * N;
    const double* t2 = qts + (M - n - 2) * N;
    const double* t3 = qts + (M - n - 3) * N;
    for (long r = 0; r < N; ++r)
      out[r] += q1[r] * t1[r] + 4.0 * (q2[r] * t2[r]) + q3[r] * t3[r];
  }
  const double simp = 0.33333 * ds;
  for (long r = 0; r < N; ++r) out[r] *= simp;
  if (even) {
    const double* q1 = qs + (M - 2) * N;
    const double* q2 = qs + (M - 1) * N;
    const double* t1 = qts + 1 * N;
    const double* t2 = qts + 0 * N;
    for (long r = 0; r < N; ++r)
      out[r] += (q1[r] * t1[r] + q2[r] * t2[r]) * (ds * 0.5);
  }
  const double iq = 1.0 / Q;
  for (long r = 0; r < N; ++r) out[r] *= iq;
}

extern "C" void orc_quadrature_trap(long N, int steps, double ds, const double* qs,
                                    const double* qts, double Q, double* out) {
  // d(-ln Q)/dw(r) of the split-step sweep: each step applies exp(-ds w/2)
  // at both of its ends, so contour point j gets ds/2 from each adjacent step.
  const long M = steps + 1;
  for (long r = 0; r < N; ++r) out[r] = 0.0;
  for (long j = 0; j < M; ++j) {
    const double c = (j == 0 || j == M - 1) ? 0.5 * ds : ds;
    const double* q = qs + j * N;
    const double* t = qts + (M - 1 - j) * N;
    for (long r = 0; r < N; ++r) out[r] += c * q[r] * t[r];
  }
  for (long r = 0; r < N; ++r) out[r] /= Q;
}

extern "C" int orc_diblock_densities(const orc_params* P, const double* wA,
                                     const double* wB, double* phiA,
                                     double* phiB, double* Qout,
                                     double* slices) {
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const long N = ncells(P);
  const int sA = orc_block_steps(P->fA, P->ds);
  const int sB = orc_block_steps(1.0 - P->fA, P->ds);
  if (sA < 0 || sB < 0) return -1;
  // exp(-ds b^2 k^2) per block (pspolymer/PsFlexPseudoSpec.cpp:188-206)
  std::vector<double> kA(N), kB(N), fa(N), fb(N), ones(N, 1.0);
  orc_kfac(nx, ny, nz, P->dx, P->dy, P->dz, P->ds, P->bA, kA.data());
  orc_kfac(nx, ny, nz, P->dx, P->dy, P->dz, P->ds, P->bB, kB.data());
  // w-factors (pspolymer/PsFlexPseudoSpec.cpp:159-181)
  const double ds2 = -0.5 * P->ds;
  for (long n = 0; n < N; ++n) { fa[n] = std::exp(ds2 * wA[n]); fb[n] = std::exp(ds2 * wB[n]); }
  std::vector<double> own;
  double* S = slices;
  const long total = (long)(sA + 1) * 2 + (long)(sB + 1) * 2;
  if (!S) { own.resize(total * N); S = own.data(); }
  double* qA = S;
  double* qB = qA + (long)(sA + 1) * N;
  double* qtB = qB + (long)(sB + 1) * N;
  double* qtA = qtB + (long)(sB + 1) * N;
  // diblock order A-fwd, B-fwd, B-bwd, A-bwd (pspolymer/PsBlockCopolymer.cpp:189-272)
  if (orc_solve_q(nx, ny, nz, kA.data(), fa.data(), ones.data(), sA, qA)) return -1;
  if (orc_solve_q(nx, ny, nz, kB.data(), fb.data(), qA + (long)sA * N, sB, qB)) return -1;
  if (orc_solve_q(nx, ny, nz, kB.data(), fb.data(), ones.data(), sB, qtB)) return -1;
  if (orc_solve_q(nx, ny, nz, kA.data(), fa.data(), qtB + (long)sB * N, sA, qtA)) return -1;
  // Q from block 0 head: qHeadInitial (= 1) * qHeadFinal (pspolymer/PsBlock.cpp:205-237)
  double sum = 0.0;
  const double* qhf = qtA + (long)sA * N;
  for (long n = 0; n < N; ++n) sum += 1.0 * qhf[n];
  const double Q = sum / (double)N;
  if (Q != Q) return -2;
  *Qout = Q;
  // densities: vf = 1, lengthRatio = 1, densWt = 1 (pspolymer/PsBlock.cpp:154-196)
  if (P->quad == 1) {
    orc_quadrature_trap(N, sA, P->ds, qA, qtA, Q, phiA);
    orc_quadrature_trap(N, sB, P->ds, qB, qtB, Q, phiB);
  } else {
    orc_quadrature(N, sA, P->ds, qA, qtA, Q, phiA);
    orc_quadrature(N, sB, P->ds, qB, qtB, Q, phiB);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* General block graphs (pspolymer/PsBlockCopolymer.cpp:132-313,
 * pspolymer/PsBlock.cpp:136-147,154-237,387-472, psbase/PsBlockBase.cpp:205-240,
 * 352-420).  Each block has a propagator from its head and one from its
 * tail.  A free end starts at q0 = 1; a joined end starts at the product of
 * the final propagators that the blocks joined there publish to it, taken in
 * ascending block index (the std::map order of qHeadJnts/qTailJnts).  Blocks
 * are swept in index order, head end then tail end, until every block has
 * both finals. */
/* ------------------------------------------------------------------------ */
namespace {

enum { kHead = 0, kTail = 1 };

struct GBlock {
  int species = 0;
  double frac = 0.0, b = 1.0;
  int steps = 0;
  int poly = 0;                                 // polymer of the block
  std::vector<std::pair<int, int>> cnt[2];      // (block, its end) joined at each end
  double* q[2] = {nullptr, nullptr};            // M slices propagated from end e
  bool initSet[2] = {false, false};             // initial q at end e
  bool finalSet[2] = {false, false};            // final q arriving at end e
  std::map<int, const double*> jnt[2];          // published finals at end e, by block
};

std::vector<orc_block> graph_desc(const orc_params* P) {
  if (P->nblocks > 0) return std::vector<orc_block>(P->blocks, P->blocks + P->nblocks);
  orc_block a{}, b{};
  a.species = 0; a.frac = P->fA; a.b = P->bA; a.nhead = 0; a.ntail = 1; a.tail[0] = 1;
  b.species = 1; b.frac = 1.0 - P->fA; b.b = P->bB; b.nhead = 1; b.head[0] = 0; b.ntail = 0;
  return {a, b};
}

// Polymers: (volfrac, lengthRatio, first block, block count).  Without a
// blend description, one polymer of volfrac = lengthRatio = 1 holds every block.
struct GPoly { double vf, lr; int first, n; };

std::vector<GPoly> poly_desc(const orc_params* P) {
  const int nb = (int)graph_desc(P).size();
  if (P->npoly <= 0 || !P->polys) return {GPoly{1.0, 1.0, 0, nb}};
  std::vector<GPoly> out;
  int first = 0;
  for (int p = 0; p < P->npoly; ++p) {
    out.push_back(GPoly{P->polys[p].volfrac, P->polys[p].lratio, first, P->polys[p].nblocks});
    first += P->polys[p].nblocks;
  }
  return out;
}

// Blocks, steps and resolved junctions; false on a bad description.  A
// blend's head/tail indices count within each polymer.
bool build_graph(const orc_params* P, std::vector<GBlock>& B) {
  const std::vector<orc_block> d = graph_desc(P);
  const std::vector<GPoly> pl = poly_desc(P);
  const int nb = (int)d.size();
  if (pl.size() > ORC_MAX_POLY) return false;
  int covered = 0;
  for (const GPoly& g : pl) {
    if (g.n < 1 || !(g.vf > 0) || !(g.lr > 0)) return false;
    covered += g.n;
  }
  if (covered != nb) return false;
  B.assign(nb, GBlock());
  for (int p = 0; p < (int)pl.size(); ++p)
    for (int i = pl[p].first; i < pl[p].first + pl[p].n; ++i) B[i].poly = p;
  for (int i = 0; i < nb; ++i) {
    if (d[i].species != 0 && d[i].species != 1) return false;
    if (d[i].nhead < 0 || d[i].ntail < 0 || d[i].nhead > ORC_MAX_JOIN || d[i].ntail > ORC_MAX_JOIN)
      return false;
    B[i].species = d[i].species;
    B[i].frac = d[i].frac;
    B[i].b = d[i].b;
    // blockSteps = lengthRatio * lengthFrac / ds (psbase/PsBlockBase.cpp:192-197)
    B[i].steps = orc_block_steps(pl[B[i].poly].lr * d[i].frac, P->ds);
    if (B[i].steps < 0) return false;
  }
  // A block j named at end e of block i is joined at j's tail when j's tail
  // list names i, else at j's head (PsBlockCopolymer::buildCntBlockType).
  for (int i = 0; i < nb; ++i)
    for (int e = 0; e < 2; ++e) {
      const int n = e == kHead ? d[i].nhead : d[i].ntail;
      const int* lst = e == kHead ? d[i].head : d[i].tail;
      const GPoly& g = pl[B[i].poly];
      for (int k = 0; k < n; ++k) {
        const int jl = lst[k];
        if (jl < 0 || jl >= g.n) return false;
        const int j = g.first + jl;
        if (j == i) return false;
        for (const auto& c : B[i].cnt[e])
          if (c.first == j) return false;
        bool at_tail = false;
        for (int m = 0; m < d[j].ntail; ++m)
          if (g.first + d[j].tail[m] == i) at_tail = true;
        B[i].cnt[e].push_back({j, at_tail ? kTail : kHead});
      }
    }
  return true;
}

}  // namespace

extern "C" void orc_species_fracs(const orc_params* P, double* fA, double* fB) {
  // f_alpha = sum of lengthWeight * lengthFrac * volfrac over the species'
  // blocks (psbase/PsBlockBase.cpp:185-189, lengthWeight = 1)
  double a = 0.0, b = 0.0;
  const std::vector<orc_block> d = graph_desc(P);
  for (const GPoly& g : poly_desc(P))
    for (int i = g.first; i < g.first + g.n && i < (int)d.size(); ++i)
      (d[i].species == 0 ? a : b) += 1.0 * d[i].frac * g.vf;
  if (fA) *fA = a;
  if (fB) *fB = b;
}

extern "C" int orc_polymer_count(const orc_params* P) { return (int)poly_desc(P).size(); }

extern "C" long orc_graph_slice_count(const orc_params* P) {
  std::vector<GBlock> B;
  if (!build_graph(P, B)) return -1;
  long t = 0;
  for (const GBlock& g : B) t += 2L * (g.steps + 1);
  return t;
}

extern "C" int orc_graph_densities(const orc_params* P, const double* wA,
                                   const double* wB, double* phiA, double* phiB,
                                   double* Qout, double* slices) {
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const long N = ncells(P);
  std::vector<GBlock> B;
  if (!build_graph(P, B)) return -1;
  const int nb = (int)B.size();
  long total = 0;
  for (const GBlock& g : B) total += 2L * (g.steps + 1);
  std::vector<double> own;
  double* S = slices;
  if (!S) { own.resize(total * N); S = own.data(); }
  for (int i = 0; i < nb; ++i) {
    const long M = B[i].steps + 1;
    B[i].q[kHead] = S;
    B[i].q[kTail] = S + M * N;
    S += 2 * M * N;
  }
  // kfac per block (pspolymer/PsFlexPseudoSpec.cpp:188-206), w-factor per species
  std::vector<std::vector<double>> kf(nb, std::vector<double>(N));
  for (int i = 0; i < nb; ++i)
    orc_kfac(nx, ny, nz, P->dx, P->dy, P->dz, P->ds, B[i].b, kf[i].data());
  std::vector<double> wf[2] = {std::vector<double>(N), std::vector<double>(N)};
  const double ds2 = -0.5 * P->ds;
  for (long n = 0; n < N; ++n) { wf[0][n] = std::exp(ds2 * wA[n]); wf[1][n] = std::exp(ds2 * wB[n]); }
  // reset: free ends start at q0 = 1 (pspolymer/PsBlock.cpp:136-147)
  for (GBlock& g : B)
    for (int e = 0; e < 2; ++e)
      if (g.cnt[e].empty()) {
        for (long r = 0; r < N; ++r) g.q[e][r] = 1.0;
        g.initSet[e] = true;
      }
  // PsBlockCopolymer::update / updateBlockQ / publishQFrom
  for (;;) {
    bool all = true, progress = false;
    for (int n = 0; n < nb; ++n)
      for (int e = 0; e < 2; ++e) {
        GBlock& g = B[n];
        const int o = 1 - e;
        if (!g.initSet[e]) {
          bool ready = true;
          for (const auto& c : g.cnt[e])
            if (!g.jnt[e].count(c.first)) ready = false;
          if (ready) {
            // combineSetJnt: 1 * q_j1 * q_j2 * ... in block-index order
            double* q0 = g.q[e];
            for (long r = 0; r < N; ++r) q0[r] = 1.0;
            for (const auto& kv : g.jnt[e])
              for (long r = 0; r < N; ++r) q0[r] *= kv.second[r];
            g.initSet[e] = true;
          }
        }
        if (g.initSet[e] && !g.finalSet[o]) {
          if (orc_solve_q(nx, ny, nz, kf[n].data(), wf[g.species].data(), g.q[e], g.steps, g.q[e]))
            return -1;
          g.finalSet[o] = true;
          progress = true;
          const double* fin = g.q[e] + (long)g.steps * N;
          for (const auto& c : g.cnt[o]) B[c.first].jnt[c.second][n] = fin;
        }
      }
    for (const GBlock& g : B)
      if (!g.finalSet[0] || !g.finalSet[1]) all = false;
    if (all) break;
    if (!progress) return -1;   // a cycle or a junction nothing reaches
  }
  // Q of each polymer from its first block's head: qHeadInitial * qHeadFinal
  // (pspolymer/PsBlock.cpp:205-237)
  const std::vector<GPoly> pl = poly_desc(P);
  std::vector<double> Qp(pl.size());
  for (size_t p = 0; p < pl.size(); ++p) {
    const GBlock& b0 = B[pl[p].first];
    const double* qhi = b0.q[kHead];
    const double* qhf = b0.q[kTail] + (long)b0.steps * N;
    double sum = 0.0;
    for (long n = 0; n < N; ++n) sum += qhi[n] * qhf[n];
    Qp[p] = sum / (double)N;
    if (Qp[p] != Qp[p]) return -2;
    Qout[p] = Qp[p];
  }
  // densities: each block adds its qqt, scaled by densFac = densWt * vf /
  // lengthRatio, into its species field (pspolymer/PsBlock.cpp:154-196,
  // densWt = 1)
  for (long r = 0; r < N; ++r) { phiA[r] = 0.0; phiB[r] = 0.0; }
  std::vector<double> tmp(N);
  for (const GBlock& g : B) {
    const double Q = Qp[g.poly];
    if (P->quad == 1) orc_quadrature_trap(N, g.steps, P->ds, g.q[kHead], g.q[kTail], Q, tmp.data());
    else orc_quadrature(N, g.steps, P->ds, g.q[kHead], g.q[kTail], Q, tmp.data());
    const double fac = 1.0 * pl[g.poly].vf / pl[g.poly].lr;
    double* phi = g.species == 0 ? phiA : phiB;
    for (long r = 0; r < N; ++r) phi[r] += tmp[r] * fac;
  }
  return 0;
}

extern "C" double orc_free_energy_poly(const orc_params* P, long N,
                                       const double* wA, const double* wB,
                                       const double* phiA, const double* phiB,
                                       const double* Qs) {
  // pseffhamil/PsPolymerUpdater.cpp:173-226, calcUpdateFieldProd 303-341,
  // pseffhamil/PsFloryInteraction.cpp:275-321 (calcDisorder = true), bulk.
  double fA, fB;
  orc_species_fracs(P, &fA, &fB);
  const double vol = (double)N;
  double prod = 0.0, inter = 0.0;
  for (long r = 0; r < N; ++r) {
    prod += (-(phiA[r] * wA[r]) - (phiB[r] * wB[r])) / vol;
    inter += (P->chiN * phiA[r] * phiB[r] - P->chiN * fA * fB) / vol;
  }
  // single-chain terms: sum_p -(volfrac_p ln Q_p) / lengthRatio_p (198-207)
  const std::vector<GPoly> pl = poly_desc(P);
  double chains = 0.0;
  for (size_t p = 0; p < pl.size(); ++p) chains = chains - ((pl[p].vf * std::log(Qs[p])) / pl[p].lr);
  return chains + (prod + inter);
}

extern "C" double orc_free_energy(const orc_params* P, long N,
                                  const double* wA, const double* wB,
                                  const double* phiA, const double* phiB,
                                  double Q) {
  return orc_free_energy_poly(P, N, wA, wB, phiA, phiB, &Q);
}

static double sd_update_q(const orc_params* P, long N, double* wA,
                          double* wB, double* p, const double* phiA,
                          const double* phiB, const double* Qs) {
  // pseffhamil/PsSteepDUpdater.cpp:159-203 with update_set2Fields 250-333
  double fA, fB;
  orc_species_fracs(P, &fA, &fB);
  std::vector<double> dHA(N), dHB(N);
  for (long r = 0; r < N; ++r) {
    dHA[r] = p[r] - wA[r] + P->chiN * (phiB[r] - fB);
    dHB[r] = p[r] - wB[r] + P->chiN * (phiA[r] - fA);
  }
  std::vector<double> wA0, wB0;
  if (P->f_order == 1) { wA0.assign(wA, wA + N); wB0.assign(wB, wB + N); }
  for (long r = 0; r < N; ++r) {
    const double xa = (dHA[r] + (-1.0 * P->lambda1) * dHB[r]) * P->lambda0;
    const double xb = (dHB[r] + (-1.0 * P->lambda1) * dHA[r]) * P->lambda0;
    wA[r] += xa;
    wB[r] += xb;
  }
  // pressure: p = 0.5 (wA + wB), then minus its mean (psphysf/PsConstraint.cpp:184-262)
  double sum = 0.0;
  for (long r = 0; r < N; ++r) { p[r] = (wA[r] + wB[r]) * 0.5; sum += p[r]; }
  const double avg = sum / (double)N;
  for (long r = 0; r < N; ++r) p[r] -= avg;
  if (P->f_order == 1) return orc_free_energy_poly(P, N, wA0.data(), wB0.data(), phiA, phiB, Qs);
  return orc_free_energy_poly(P, N, wA, wB, phiA, phiB, Qs);
}

extern "C" double orc_sd_update(const orc_params* P, long N, double* wA,
                                double* wB, double* p, const double* phiA,
                                const double* phiB, double Q) {
  return sd_update_q(P, N, wA, wB, p, phiA, phiB, &Q);
}

/* ------------------------------------------------------------------------ */
/* Anderson mixing (no reference counterpart).  Type-II Anderson acceleration
 * (Walker & Ni, SIAM J. Numer. Anal. 49 (2011) 1715; the SCFT form of
 * Matsen, Eur. Phys. J. E 30 (2009) 361):
 *   U_ij = <f_k - f_{k-i}, f_k - f_{k-j}>,  V_i = <f_k - f_{k-i}, f_k>,
 *   c = U^{-1} V,  xbar = x_k + sum_i c_i (x_{k-i} - x_k),
 *   fbar = f_k + sum_i c_i (f_{k-i} - f_k),  x_{k+1} = xbar + beta fbar,
 * with x = (wA, wB), f = (dH_A, dH_B) and <a,b> = sum_r (a_A b_A + a_B b_B).
 * The final step couples A and B like the reference's relaxlambdas[1]:
 * x_{k+1} = xbar + beta (fbar_A - lambda1 fbar_B, fbar_B - lambda1 fbar_A).
 * Depth 0 (or residual RMS above anderson_start) is that step with lambda0
 * and no history: steepest descent [lambda0, lambda1]. */
/* ------------------------------------------------------------------------ */
namespace {

// Solve the small dense system A x = b (n <= 32) by Gaussian elimination
// with partial pivoting; returns false if singular.
// A pivot at or below 1e-13 times the largest |A_ii| counts as singular.
bool solve_dense(int n, std::vector<double> A, std::vector<double> b,
                 std::vector<double>& x) {
  double dmax = 0.0;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, std::fabs(A[i * n + i]));
  const double thr = 1e-13 * dmax;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (!(std::fabs(A[piv * n + c]) > thr)) return false;
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  x.assign(n, 0.0);
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= A[r * n + k] * x[k];
    x[r] = s / A[r * n + r];
  }
  return true;
}

}  // namespace

static int scft_run_impl(const orc_params* P, double* wA, double* wB,
                         double* p, int n_iters, double tol,
                         double* F_hist, double* Q_hist, double* err_hist,
                         double* phiA_out, double* phiB_out, double* t_hist) {
  const long N = ncells(P);
  std::vector<double> phiA(N), phiB(N);
  double fA, fB;
  orc_species_fracs(P, &fA, &fB);
  const int m = P->mix == 1 ? P->depth : 0;
  // history ring (x, f) of the last m iterates
  std::vector<std::vector<double>> hx, hf;
  int done = 0;
  for (int it = 0; it < n_iters; ++it) {
    const auto t_it = std::chrono::steady_clock::now();
    double Qs[ORC_MAX_POLY];
    // pstrol/PsDomain.cpp:332-372: densities reset, propagate, Q, densities
    int rc = P->nblocks > 0 ? orc_graph_densities(P, wA, wB, phiA.data(), phiB.data(), Qs, nullptr)
                            : orc_diblock_densities(P, wA, wB, phiA.data(), phiB.data(), Qs, nullptr);
    if (rc) return -1;
    const double Q = Qs[0];
    // residual f = dH (pseffhamil/PsSteepDUpdater.cpp:174-184, Flory calcDfD)
    std::vector<double> f(2 * N), x(2 * N);
    double ff = 0.0;
    for (long r = 0; r < N; ++r) {
      f[r] = p[r] - wA[r] + P->chiN * (phiB[r] - fB);
      f[N + r] = p[r] - wB[r] + P->chiN * (phiA[r] - fA);
      x[r] = wA[r];
      x[N + r] = wB[r];
      ff += f[r] * f[r] + f[N + r] * f[N + r];
    }
    const double err = std::sqrt(ff / (double)N);
    double F;
    if (P->mix == 0) {
      F = sd_update_q(P, N, wA, wB, p, phiA.data(), phiB.data(), Qs);
    } else {
      std::vector<double> xn(2 * N);
      const int k = (int)hx.size();
      const int mk = (err > P->anderson_start) ? 0 : std::min(m, k);
      std::vector<double> c;
      bool ok = false;
      if (mk > 0) {
        std::vector<double> U(mk * mk), V(mk);
        std::vector<std::vector<double>> D(mk, std::vector<double>(2 * N));
        for (int i = 0; i < mk; ++i) {
          const std::vector<double>& fo = hf[k - 1 - i];
          for (long r = 0; r < 2 * N; ++r) D[i][r] = f[r] - fo[r];
        }
        for (int i = 0; i < mk; ++i) {
          double v = 0.0;
          for (long r = 0; r < 2 * N; ++r) v += D[i][r] * f[r];
          V[i] = v;
          for (int j = 0; j <= i; ++j) {
            double u = 0.0;
            for (long r = 0; r < 2 * N; ++r) u += D[i][r] * D[j][r];
            U[i * mk + j] = U[j * mk + i] = u;
          }
        }
        ok = solve_dense(mk, U, V, c);
      }
      // x_bar, f_bar (Anderson combination; plain x, f when no history is used)
      std::vector<double> xb(x), fb(f);
      if (ok) {
        for (long r = 0; r < 2 * N; ++r)
          for (int i = 0; i < mk; ++i) {
            xb[r] += c[i] * (hx[k - 1 - i][r] - x[r]);
            fb[r] += c[i] * (hf[k - 1 - i][r] - f[r]);
          }
      }
      // step with the reference's A/B coupling (pseffhamil/PsSteepDUpdater.cpp:265-278):
      // dwA = s (fA - lambda1 fB), dwB = s (fB - lambda1 fA), s = beta or lambda0
      const double step = ok ? P->beta : P->lambda0;
      for (long r = 0; r < N; ++r) {
        xn[r] = xb[r] + (fb[r] + (-P->lambda1) * fb[N + r]) * step;
        xn[N + r] = xb[N + r] + (fb[N + r] + (-P->lambda1) * fb[r]) * step;
      }
      if (m > 0) {
        hx.push_back(x);
        hf.push_back(f);
        if ((int)hx.size() > m) { hx.erase(hx.begin()); hf.erase(hf.begin()); }
      }
      std::vector<double> wA0(wA, wA + N), wB0(wB, wB + N);
      double sum = 0.0;
      for (long r = 0; r < N; ++r) {
        wA[r] = xn[r];
        wB[r] = xn[N + r];
        p[r] = (wA[r] + wB[r]) * 0.5;
        sum += p[r];
      }
      const double avg = sum / (double)N;
      for (long r = 0; r < N; ++r) p[r] -= avg;
      if (P->f_order == 1)
        F = orc_free_energy_poly(P, N, wA0.data(), wB0.data(), phiA.data(), phiB.data(), Qs);
      else
        F = orc_free_energy_poly(P, N, wA, wB, phiA.data(), phiB.data(), Qs);
    }
    if (F_hist) F_hist[it] = F;
    if (Q_hist) Q_hist[it] = Q;
    if (err_hist) err_hist[it] = err;
    if (t_hist)
      t_hist[it] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_it).count();
    done = it + 1;
    if (tol > 0.0 && err < tol) break;
  }
  if (phiA_out) std::memcpy(phiA_out, phiA.data(), N * sizeof(double));
  if (phiB_out) std::memcpy(phiB_out, phiB.data(), N * sizeof(double));
  return done;
}

extern "C" int orc_scft_run(const orc_params* P, double* wA, double* wB,
                            double* p, int n_iters, double tol,
                            double* F_hist, double* Q_hist, double* err_hist,
                            double* phiA_out, double* phiB_out) {
  return scft_run_impl(P, wA, wB, p, n_iters, tol, F_hist, Q_hist, err_hist, phiA_out, phiB_out,
                       nullptr);
}

extern "C" int orc_scft_run_timed(const orc_params* P, double* wA, double* wB, double* p,
                                  int n_iters, double* F_hist, double* t_hist) {
  return scft_run_impl(P, wA, wB, p, n_iters, 0.0, F_hist, nullptr, nullptr, nullptr, nullptr,
                       t_hist);
}

/* ------------------------------------------------------------------------ */
/* Box stress (no reference counterpart; the reference has no analytic
 * stress).  With T_t = W_t F^-1 K_t F W_t the step operator of contour step t
 * and K = exp(-ds b^2 k^2), k_d = 2 pi n_d / L_d, the partition function
 * N Q = <q+_{Ns-1-t}, T_t q_t> for every t, so
 *   dQ/dL_d = (1/N) sum_t <q+_{Ns-1-t}, W_t F^-1 (dK/dL_d) F W_t q_t>,
 *   dK/dL_d = K * ds b^2 * 2 k_d^2 / L_d,
 * and at fixed fields dF/dL_d = -(1/Q) dQ/dL_d. */
/* ------------------------------------------------------------------------ */
extern "C" int orc_box_stress(const orc_params* P, const double* wA,
                              const double* wB, double* dFdL) {
  const int nx = P->nx, ny = P->ny, nz = P->nz;
  const long N = ncells(P);
  std::vector<GBlock> B;
  if (!build_graph(P, B)) return -1;
  const long total = orc_graph_slice_count(P);
  std::vector<double> S(total * N), phiA(N), phiB(N);
  double Qs[ORC_MAX_POLY];
  if (orc_graph_densities(P, wA, wB, phiA.data(), phiB.data(), Qs, S.data())) return -1;
  const std::vector<GPoly> pl = poly_desc(P);
  const double twopi = 2.0 * std::acos(-1.0);
  const int g[3] = {nx, ny, nz};
  const double h[3] = {P->dx, P->dy, P->dz};
  std::vector<double> kd2[3];
  for (int d = 0; d < 3; ++d) kd2[d].resize(N);
  long n = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k, ++n) {
        const int idx[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
          double n2 = g[d] / 2.0;
          double nk = n2 - std::fabs((double)idx[d] - n2);
          double kv = twopi * nk / (double)g[d] / h[d];
          kd2[d][n] = kv * kv;
        }
      }
  std::vector<double> wf[2] = {std::vector<double>(N), std::vector<double>(N)};
  for (long r = 0; r < N; ++r) {
    wf[0][r] = std::exp(-0.5 * P->ds * wA[r]);
    wf[1][r] = std::exp(-0.5 * P->ds * wB[r]);
  }
  std::vector<std::array<double, 3>> accp(pl.size(), std::array<double, 3>{0.0, 0.0, 0.0});
  std::vector<cd> a(N), b(N);
  std::vector<double> K(N);
  const double* base = S.data();
  for (const GBlock& blk : B) {
    double* acc = accp[blk.poly].data();
    // step t of the block: q_t from the head, q+ after steps-1-t steps from the tail
    const long M = blk.steps + 1;
    const double* qh = base;
    const double* qt = base + M * N;
    base += 2 * M * N;
    orc_kfac(nx, ny, nz, P->dx, P->dy, P->dz, P->ds, blk.b, K.data());
    const double* w = wf[blk.species].data();
    const double bb = blk.b * blk.b;
    for (int t = 0; t < blk.steps; ++t) {
      const double* q1 = qh + (long)t * N;
      const double* q2 = qt + (long)(blk.steps - 1 - t) * N;
      for (long r = 0; r < N; ++r) { a[r] = cd(w[r] * q1[r], 0.0); b[r] = cd(w[r] * q2[r], 0.0); }
      fft3(a.data(), nx, ny, nz, false);
      fft3(b.data(), nx, ny, nz, false);
      for (long r = 0; r < N; ++r) {
        const double s = (std::conj(b[r]) * a[r]).real() * K[r] * P->ds * bb * 2.0;
        for (int d = 0; d < 3; ++d) acc[d] += s * kd2[d][r];
      }
    }
  }
  // F = -sum_p (vf_p / lr_p) ln Q_p + (terms without L at fixed fields)
  for (int d = 0; d < 3; ++d) {
    const double L = g[d] * h[d];
    double sum = 0.0;
    for (size_t p = 0; p < pl.size(); ++p) {
      // <x, F^-1 y> = (1/N) <F x, y>, then Q = (N Q) / N
      const double dQ = accp[p][d] / L / (double)N / (double)N;
      sum += (pl[p].vf / pl[p].lr) * (dQ / Qs[p]);
    }
    dFdL[d] = g[d] > 1 ? -sum : 0.0;
  }
  return 0;
}

extern "C" double orc_time_steps(int nx, int ny, int nz, int steps) {
  const long N = (long)nx * ny * nz;
  std::vector<double> k(N), w(N), q0(N, 1.0), sl((long)(steps + 1) * N);
  orc_kfac(nx, ny, nz, 0.0625, 0.0625, 0.0625, 0.01, 1.0, k.data());
  for (long r = 0; r < N; ++r) w[r] = std::exp(-0.005 * std::sin(0.001 * r));
  auto t0 = std::chrono::steady_clock::now();
  orc_solve_q(nx, ny, nz, k.data(), w.data(), q0.data(), steps, sl.data());
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

/* ------------------------------------------------------------------------ */
/* The rest of PsFFTBase and its consumers (spectral filters, Poisson).      */
/* ------------------------------------------------------------------------ */
namespace {
std::vector<cd> forward_of_real(int nx, int ny, int nz, const double* data, bool imag) {
  const long N = (long)nx * ny * nz;
  std::vector<cd> a(N);
  for (long n = 0; n < N; ++n) a[n] = imag ? cd(0.0, data[n]) : cd(data[n], 0.0);
  fft3(a.data(), nx, ny, nz, false);
  return a;
}
}  // namespace

extern "C" int orc_fft_forward_abs(int nx, int ny, int nz, const double* data, double* res) {
  // psfft/PsFFTW.cpp:121-146: forward, then re^2 + im^2 (the squared modulus)
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  std::vector<cd> a = forward_of_real(nx, ny, nz, data, false);
  for (size_t n = 0; n < a.size(); ++n)
    res[n] = (a[n].real() * a[n].real()) + (a[n].imag() * a[n].imag());
  return 0;
}

extern "C" int orc_fft_convolve_re(int nx, int ny, int nz, const double* d1, const double* d2,
                                   double* res) {
  // psfft/PsFFTW.cpp:148-188
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  std::vector<cd> a = forward_of_real(nx, ny, nz, d1, false);
  std::vector<cd> b = forward_of_real(nx, ny, nz, d2, false);
  for (size_t n = 0; n < a.size(); ++n) {
    const double re = (a[n].real() * b[n].real()) - (a[n].imag() * b[n].imag());
    const double im = (a[n].imag() * b[n].real()) + (a[n].real() * b[n].imag());
    a[n] = cd(re, im);
  }
  fft3(a.data(), nx, ny, nz, true);
  for (size_t n = 0; n < a.size(); ++n) res[n] = a[n].real();
  return 0;
}

extern "C" int orc_fft_scaled_pair_im(int nx, int ny, int nz, const double* data,
                                      const double* kdata, double* res) {
  // psfft/PsFFTW.cpp:226-256: the input is 0 + i data
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  std::vector<cd> a = forward_of_real(nx, ny, nz, data, true);
  for (size_t n = 0; n < a.size(); ++n) a[n] = cd(a[n].real() * kdata[n], a[n].imag() * kdata[n]);
  fft3(a.data(), nx, ny, nz, true);
  for (size_t n = 0; n < a.size(); ++n) res[n] = a[n].real();
  return 0;
}

extern "C" int orc_fft_forward_re(int nx, int ny, int nz, const double* data, double* res) {
  // psfft/PsFFTW.cpp:258-276 (calcForwardFFT)
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  std::vector<cd> a = forward_of_real(nx, ny, nz, data, false);
  for (size_t n = 0; n < a.size(); ++n) res[n] = a[n].real();
  return 0;
}

extern "C" int orc_fft_backward_re(int nx, int ny, int nz, const double* data, double* res) {
  // psfft/PsFFTW.cpp:278-296 (calcBackwardFFT)
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const long N = (long)nx * ny * nz;
  std::vector<cd> a(N);
  for (long n = 0; n < N; ++n) a[n] = cd(data[n], 0.0);
  fft3(a.data(), nx, ny, nz, true);
  for (long n = 0; n < N; ++n) res[n] = a[n].real();
  return 0;
}

extern "C" int orc_spec_filter(int nx, int ny, int nz, double* w, double cutoff, double strength,
                               const int* cells, double vwall) {
  // pseffhamil/PsSimpleSpecFilter.cpp:74-144, PsMultiSpecFilter.cpp:123-267,
  // PsSpecFilterUpdater.cpp:222-239 (one rank, FLOATTYPE double)
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  if (cutoff <= 0.0) return 0;
  const long N = (long)nx * ny * nz;
  // subtractAverage: sequential sum (getSumAll), then w += -(sum / vol)
  double tot = 0.0;
  for (long n = 0; n < N; ++n) tot = tot + w[n];
  const double localVol = (double)N - vwall;
  const double wtot = -1.0 * (tot / (double)(1 * localVol));
  for (long n = 0; n < N; ++n) w[n] += wtot;
  std::vector<double> res(N), mask(N);
  orc_fft_forward_abs(nx, ny, nz, w, res.data());
  if (!cells) {
    double maxVal = -1000000.00;
    for (long n = 0; n < N; ++n) if (res[n] > maxVal) maxVal = res[n];
    const double maxCutoff = cutoff * maxVal;
    for (long n = 0; n < N; ++n) mask[n] = res[n] < maxCutoff ? strength : 1.0;
  } else {
    const int g[3] = {nx, ny, nz};
    long nsc[3];
    for (int d = 0; d < 3; ++d) {
      if (cells[d] < 1 || g[d] % cells[d] != 0) return -2;
      nsc[d] = g[d] / cells[d];
    }
    const long tot_cells = nsc[0] * nsc[1] * nsc[2];
    // build_specCells_transpose on one rank: flat index in (y, x, z) order
    std::vector<long> kcellMap(N);
    for (long jloc = 0; jloc < ny; ++jloc)
      for (long i = 0; i < nx; ++i)
        for (long k = 0; k < nz; ++k) {
          const long j = jloc;
          const long ic0 = i / cells[0], ic1 = j / cells[1], ic2 = k / cells[2];
          const long nn = ((jloc * nx) + i) * nz + k;
          kcellMap[nn] = ic2 * (nsc[0] * nsc[1]) + ic1 * nsc[0] + ic0;
        }
    std::vector<double> cutTmp(tot_cells, 0.0);
    for (long n = 0; n < N; ++n) {
      const long r = kcellMap[n];
      const double kCut = cutoff * res[n];
      if (kCut > cutTmp[r]) cutTmp[r] = kCut;
    }
    for (long n = 0; n < N; ++n) mask[n] = res[n] < cutTmp[kcellMap[n]] ? strength : 1.0;
  }
  std::vector<double> out(N);
  orc_scaled_fft_pair(nx, ny, nz, w, mask.data(), out.data());
  const double scaleFFT = 1.0 / (double)N;
  for (long n = 0; n < N; ++n) w[n] = scaleFFT * out[n];
  return 0;
}

extern "C" int orc_poisson_update(int nx, int ny, int nz, double dx, double dy, double dz,
                                  double lB, double* charge, double* psi) {
  // pseffhamil/PsPoissonUpdater.cpp:171-289 (one rank, FLOATTYPE double)
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const long N = (long)nx * ny * nz;
  std::vector<double> k2(N), kf(N);
  orc_k2(nx, ny, nz, dx, dy, dz, k2.data());
  for (long n = 0; n < N; ++n) kf[n] = k2[n] < 0.000000001 ? 0.0 : (4.0 * 3.1415 * lB) / k2[n];
  double localC = 0.0;
  for (long n = 0; n < N; ++n) localC = localC + charge[n];
  localC = localC / ((double)1 * (double)N);
  const double totalC = -1.0 * localC;
  for (long n = 0; n < N; ++n) charge[n] += totalC;
  std::vector<double> res(N);
  orc_scaled_fft_pair(nx, ny, nz, charge, kf.data(), res.data());
  const double scaleFFT = 1.0 / (double)N;
  for (long n = 0; n < N; ++n) {
    const double psi0 = psi[n];
    const double psi1 = res[n] * scaleFFT;
    const double frac = std::fabs((psi1 - psi0) / psi0);
    if (frac > 0.2) res[n] = psi0 * 1.2;
  }
  for (long n = 0; n < N; ++n) {
    const double tmp = (1.0 * scaleFFT) * res[n];
    psi[n] = tmp;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The reference's random stream and the noisy steepest descent.            */
/* ------------------------------------------------------------------------ */
// MT19937 (Matsumoto & Nishimura 1998) as PsRandom runs it
// (psstd/PsRandom.cpp:58-95, psstd/PsRandom.h:147-170,305-373): state
// words, one block regeneration per 624 outputs, tempering; real() = y/2^32;
// psRandomGaussian = Marsaglia polar, v2*fac returned, v1*fac kept.
struct orc_rng {
  unsigned long st[624];
  int left = 1;
  int pos = 0;
  int iset = 0;
  double gset = 0.0;
};

namespace {
void orc_mt_regen(orc_rng* g) {
  const unsigned long UP = 0x80000000UL, LO = 0x7fffffffUL, A = 0x9908b0dfUL;
  for (int i = 0; i < 624; ++i) {
    const unsigned long u = g->st[i], v = g->st[(i + 1) % 624];
    const unsigned long mix = (u & UP) | (v & LO);
    g->st[i] = g->st[(i + 397) % 624] ^ (mix >> 1) ^ ((v & 1UL) ? A : 0UL);
  }
  g->pos = 0;
  g->left = 624;
}

unsigned long orc_mt_next(orc_rng* g) {
  if (--g->left == 0) orc_mt_regen(g);
  unsigned long y = g->st[g->pos++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680UL;
  y ^= (y << 15) & 0xefc60000UL;
  y ^= (y >> 18);
  return y & 0xffffffffUL;
}

double orc_mt_real(orc_rng* g) { return (double)(long)orc_mt_next(g) * (1.0 / 4294967296.0); }

double orc_mt_gauss(orc_rng* g) {
  if (g->iset != 0) {
    g->iset = 0;
    return g->gset;
  }
  double v1, v2, r;
  do {
    double u = orc_mt_real(g);
    v1 = (2.0 * u) - 1.0;
    u = orc_mt_real(g);
    v2 = (2.0 * u) - 1.0;
    r = (v1 * v1) + (v2 * v2);
  } while ((r >= 1.0) || (r == 0.0));
  const double fac = sqrt(-2.0 * log(r) / r);
  g->gset = v1 * fac;
  g->iset = 1;
  return v2 * fac;
}
}  // namespace

extern "C" orc_rng* orc_rng_create(unsigned long seed) {
  orc_rng* g = new orc_rng();
  g->st[0] = seed & 0xffffffffUL;
  for (int j = 1; j < 624; ++j)
    g->st[j] = (1812433253UL * (g->st[j - 1] ^ (g->st[j - 1] >> 30)) + (unsigned long)j) & 0xffffffffUL;
  g->left = 1;
  g->pos = 0;
  return g;
}

extern "C" void orc_rng_destroy(orc_rng* g) { delete g; }

extern "C" void orc_rng_raw32(orc_rng* g, long n, unsigned int* out) {
  for (long i = 0; i < n; ++i) out[i] = (unsigned int)orc_mt_next(g);
}

extern "C" void orc_rng_uniform(orc_rng* g, long n, double* out) {
  for (long i = 0; i < n; ++i) out[i] = orc_mt_real(g);
}

extern "C" void orc_rng_gaussian(orc_rng* g, long n, double* out) {
  for (long i = 0; i < n; ++i) out[i] = orc_mt_gauss(g);
}

extern "C" int orc_scft_run_noise(const orc_params* P, double* wA, double* wB, double* p,
                                  int n_iters, double* F_hist, double* Q_hist, double* err_hist,
                                  double strength, orc_rng* g) {
  // Steepest descent with PsPolymerUpdater::addRandUpdateConjgFields
  // (pseffhamil/PsPolymerUpdater.cpp:264-295) after the lambda update
  // (pseffhamil/PsSteepDUpdater.cpp:255-283): per step N deviates for w_A's
  // noise field, then N for w_B's, each value (0 + strength g) * 1.
  if (P->mix != 0) return -1;
  const long N = ncells(P);
  std::vector<double> phiA(N), phiB(N), nA(N), nB(N);
  double fA, fB;
  orc_species_fracs(P, &fA, &fB);
  for (int it = 0; it < n_iters; ++it) {
    double Qs[ORC_MAX_POLY];
    int rc = P->nblocks > 0 ? orc_graph_densities(P, wA, wB, phiA.data(), phiB.data(), Qs, nullptr)
                            : orc_diblock_densities(P, wA, wB, phiA.data(), phiB.data(), Qs, nullptr);
    if (rc) return -1;
    const double Q = Qs[0];
    double ff = 0.0;
    std::vector<double> dHA(N), dHB(N);
    for (long r = 0; r < N; ++r) {
      dHA[r] = p[r] - wA[r] + P->chiN * (phiB[r] - fB);
      dHB[r] = p[r] - wB[r] + P->chiN * (phiA[r] - fA);
      ff += dHA[r] * dHA[r] + dHB[r] * dHB[r];
    }
    for (long r = 0; r < N; ++r) {
      wA[r] += (dHA[r] + (-1.0 * P->lambda1) * dHB[r]) * P->lambda0;
      wB[r] += (dHB[r] + (-1.0 * P->lambda1) * dHA[r]) * P->lambda0;
    }
    for (long r = 0; r < N; ++r) nA[r] = (0.0 + strength * orc_mt_gauss(g)) * 1.0;
    for (long r = 0; r < N; ++r) nB[r] = (0.0 + strength * orc_mt_gauss(g)) * 1.0;
    for (long r = 0; r < N; ++r) { wA[r] += nA[r]; wB[r] += nB[r]; }
    double sum = 0.0;
    for (long r = 0; r < N; ++r) { p[r] = (wA[r] + wB[r]) * 0.5; sum += p[r]; }
    const double avg = sum / (double)N;
    for (long r = 0; r < N; ++r) p[r] -= avg;
    const double F = orc_free_energy_poly(P, N, wA, wB, phiA.data(), phiB.data(), Qs);
    if (F_hist) F_hist[it] = F;
    if (Q_hist) Q_hist[it] = Q;
    if (err_hist) err_hist[it] = std::sqrt(ff / (double)N);
  }
  return n_iters;
}
