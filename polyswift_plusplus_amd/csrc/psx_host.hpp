// psx_host.hpp -- C++ host wrapper over the psx C ABI (include/psx_scft.h).
//
// The classes sit where the reference's plugins sit and keep their method
// names, so a PolySwift++ build can call them in place of its own:
//   psx::HipFFT          ~ PsFFTBase<double,NDIM> / PsFFTW (psbase/PsFFTBase.h:39-147)
//   psx::DiblockSolver   ~ PsBlockCopolymer + two PsFlexPseudoSpec blocks
//                          (pspolymer/PsBlockCopolymer.cpp:189-239) with the
//                          flory interaction and steepestDescent updater
//                          (pseffhamil/PsSteepDUpdater.cpp:159-203)
//   psx::BlockCopolymerSolver, psx::BlendSolver
//                        ~ the same for any tree of blocks, and for several
//                          Polymers of a PsPolymerHldr (pspolymer/PsPolymerHldr.cpp:35-83)
//   psx::Batch           ~ many diblocks advanced together (no counterpart)
// Every non-zero psx_status becomes a psx::Error exception carrying the
// library's message, the way the reference throws TxDebugExcept.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/psx_scft.h"

namespace psx {

class Error : public std::runtime_error {
 public:
  Error(psx_status s, const std::string& m) : std::runtime_error(m), status(s) {}
  psx_status status;
};

inline void check(psx_status s) {
  if (s == PSX_OK) return;
  char buf[1024];
  psx_last_error(buf, sizeof(buf));
  throw Error(s, std::string("psx status ") + std::to_string(s) + ": " + buf);
}

// A device buffer (psx_buffer_*), for host-array forms of device calls.
class DevBuf {
 public:
  explicit DevBuf(size_t bytes) : n_(bytes) { check(psx_buffer_create(bytes, &d_)); }
  ~DevBuf() { psx_buffer_destroy(d_); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void* get() const { return d_; }
  void upload(const void* h) { check(psx_buffer_upload(d_, h, n_)); }
  void download(void* h) const { check(psx_buffer_download(h, d_, n_)); }

 private:
  void* d_ = nullptr;
  size_t n_ = 0;
};

// FFT plugin: every PsFFTBase transform on host arrays of getFFTSize()
// elements of the plan's type, unnormalised (psfft/PsFFTW.cpp:121-296).
class HipFFT {
 public:
  HipFFT(int nx, int ny, int nz, psx_prec prec = PSX_FP64) : prec_(prec) {
    check(psx_fft_plan_create(nx, ny, nz, prec, nullptr, &p_));
  }
  ~HipFFT() { psx_fft_plan_destroy(p_); }
  HipFFT(const HipFFT&) = delete;
  HipFFT& operator=(const HipFFT&) = delete;
  size_t getFFTSize() const { return psx_fft_size(p_); }
  void scaledFFTPair(const void* data, const void* kdata, void* resPtr) {
    check(psx_fft_scaled_pair_host(p_, data, kdata, resPtr));
  }
  // device-pointer form (asynchronous on the plan's stream)
  void scaledFFTPairDevice(const void* d_data, const void* d_kdata, void* d_res) {
    check(psx_fft_scaled_pair(p_, d_data, d_kdata, d_res));
  }
  void forwardFFTAbs(const void* data, void* resPtr) { run1(psx_fft_forward_abs, data, resPtr); }
  void calcForwardFFT(const void* data, void* resPtr) { run1(psx_fft_forward_re, data, resPtr); }
  void calcBackwardFFT(const void* data, void* resPtr) { run1(psx_fft_backward_re, data, resPtr); }
  void convolveRe(const void* data1, const void* data2, void* resPtr) {
    run2(psx_fft_convolve_re, data1, data2, resPtr);
  }
  void scaledFFTPairIm(const void* data, const void* kdata, void* resPtr) {
    run2(psx_fft_scaled_pair_im, data, kdata, resPtr);
  }
  // PsSimpleSpecFilter (specCells null) / PsMultiSpecFilter::applyFilter
  // after subtractAverage, in place on a host field
  void specFilter(void* w, double cutoffFactor, double filterStrength, const int* specCells = nullptr,
                  double wallVolume = 0.0) {
    DevBuf d(bytes());
    d.upload(w);
    check(psx_fft_spec_filter(p_, d.get(), cutoffFactor, filterStrength, specCells, wallVolume));
    check(psx_fft_synchronize(p_));
    d.download(w);
  }
  // PsPoissonUpdater::update, in place on host charge and psi fields
  void poissonUpdate(void* charge, void* psi, double bjerrumLen, double dx, double dy, double dz) {
    DevBuf c(bytes()), q(bytes());
    c.upload(charge);
    q.upload(psi);
    check(psx_fft_poisson_update(p_, c.get(), q.get(), bjerrumLen, dx, dy, dz));
    check(psx_fft_synchronize(p_));
    c.download(charge);
    q.download(psi);
  }
  psx_fft* handle() { return p_; }

 private:
  size_t bytes() const { return getFFTSize() * (prec_ == PSX_FP64 ? 8 : 4); }
  void run1(psx_status (*fn)(psx_fft*, const void*, void*), const void* in, void* out) {
    DevBuf a(bytes()), r(bytes());
    a.upload(in);
    check(fn(p_, a.get(), r.get()));
    check(psx_fft_synchronize(p_));
    r.download(out);
  }
  void run2(psx_status (*fn)(psx_fft*, const void*, const void*, void*), const void* in1,
            const void* in2, void* out) {
    DevBuf a(bytes()), b(bytes()), r(bytes());
    a.upload(in1);
    b.upload(in2);
    check(fn(p_, a.get(), b.get(), r.get()));
    check(psx_fft_synchronize(p_));
    r.download(out);
  }
  psx_fft* p_ = nullptr;
  psx_prec prec_;
};

// The reference's random stream theRng (psstd/PsRandom.h).
class Rng {
 public:
  explicit Rng(unsigned long long seed) { check(psx_rng_create(seed, &r_)); }
  ~Rng() { psx_rng_destroy(r_); }
  Rng(const Rng&) = delete;
  Rng& operator=(const Rng&) = delete;
  void uniform(long long n, double* out) { check(psx_rng_uniform(r_, n, out)); }
  void gaussian(long long n, double* out) { check(psx_rng_gaussian(r_, n, out)); }
  psx_rng* handle() { return r_; }

 private:
  psx_rng* r_ = nullptr;
};

// One SCFT system on one HIP device: the PsBlockCopolymer / PsPolymerHldr
// side of the reference (reset + solveQ of every block, calcBigQ, the
// quadrature) with the flory interaction and the field updaters.  The
// concrete solvers below differ only in how they build the chain.
class Engine {
 public:
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;
  Engine(Engine&& o) noexcept : e_(o.e_), n_(o.n_), own_(o.own_) { o.e_ = nullptr; }
  virtual ~Engine() {
    if (own_ && e_) psx_scft_destroy(e_);
  }

  size_t size() const { return n_; }
  void setFields(const std::vector<double>& wA, const std::vector<double>& wB,
                 const std::vector<double>* p = nullptr) {
    if (wA.size() != n_ || wB.size() != n_ || (p && p->size() != n_))
      throw Error(PSX_ERR_SHAPE, "field size differs from the grid");
    check(psx_scft_set_fields_host(e_, wA.data(), wB.data(), p ? p->data() : nullptr));
  }
  void getFields(std::vector<double>& wA, std::vector<double>& wB, std::vector<double>& p) {
    wA.resize(n_); wB.resize(n_); p.resize(n_);
    check(psx_scft_get_fields_host(e_, wA.data(), wB.data(), p.data()));
  }
  void solveQ() { check(psx_scft_propagate(e_)); }
  double calcBigQ() { double q; check(psx_scft_Q(e_, &q)); return q; }
  // Q of every polymer of a blend (one entry for a single copolymer)
  std::vector<double> polymerQ() {
    double q[PSX_MAX_POLYMERS];
    int n = 0;
    check(psx_scft_polymer_Q(e_, q, &n));
    return std::vector<double>(q, q + n);
  }
  void setPhysFields(std::vector<double>& phiA, std::vector<double>& phiB) {
    phiA.resize(n_); phiB.resize(n_);
    check(psx_scft_densities(e_, nullptr, nullptr));
    check(psx_scft_get_densities_host(e_, phiA.data(), phiB.data()));
  }
  double freeEnergy() { double F; check(psx_scft_free_energy(e_, &F)); return F; }
  // steepest-descent noise from the reference's stream (nullptr: off)
  void setNoise(double strength, Rng* rng) {
    check(psx_scft_set_noise(e_, strength, rng ? rng->handle() : nullptr));
  }
  // spectral filter updater on w_A (fields & 1) and w_B (fields & 2)
  void specFilter(int fields, double cutoffFactor, double filterStrength, const int* specCells = nullptr) {
    check(psx_scft_spec_filter(e_, fields, cutoffFactor, filterStrength, specCells));
  }
  // sweep launches: 0 two line launches per contour step, 1 plane launches
  // (one per step; 3D grids with ny == nz in {32, 48, 64, 96, 128}, the default there)
  void setSweepMode(int mode) { check(psx_scft_set_sweep_mode(e_, mode)); }
  int sweepMode() const {
    int m = 0;
    check(psx_scft_sweep_mode(e_, &m, nullptr));
    return m;
  }
  // time the line kernels and both sweep modes, keep the fastest
  void autotune(int reps = 5) { check(psx_scft_autotune(e_, reps, nullptr)); }
  double steepestDescent(double lambda0, double lambda1) {
    double F;
    check(psx_scft_update_steepest(e_, lambda0, lambda1, &F));
    return F;
  }
  // n_iters SCFT iterations (PsDomain::update) until the residual RMS falls
  // below tol (0: all of them); returns the iterations run.
  int iterate(const psx_mixing& mix, int n_iters, double tol, std::vector<double>* F = nullptr,
              std::vector<double>* err = nullptr, std::vector<double>* Q = nullptr) {
    std::vector<double> f(n_iters), q(n_iters), e(n_iters);
    int done = 0;
    check(psx_scft_iterate(e_, &mix, n_iters, tol, 1, f.data(), q.data(), e.data(), &done));
    if (F) F->assign(f.begin(), f.begin() + done);
    if (err) err->assign(e.begin(), e.begin() + done);
    if (Q) Q->assign(q.begin(), q.begin() + done);
    return done;
  }
  void resetHistory() { check(psx_scft_reset_history(e_)); }
  // FP32 sweeps with FP64 reductions and updates, or plain FP64
  void setPrecision(psx_prec prec) { check(psx_scft_set_precision(e_, prec)); }
  // 0 the reference's Simpson rule (0.33333), 1 the trapezoid rule
  void setQuadrature(int mode) { check(psx_scft_set_quadrature(e_, mode)); }
  // keep every K-th backward slice and recompute the rest (0: store all)
  void setCheckpointing(int every) { check(psx_scft_set_checkpointing(e_, every)); }
  void setCellSizes(double dx, double dy, double dz) { check(psx_scft_set_cell(e_, dx, dy, dz)); }
  // dF/dL_x,y,z at fixed fields (after solveQ)
  std::vector<double> boxStress() {
    std::vector<double> g(3);
    check(psx_scft_box_stress(e_, g.data()));
    return g;
  }
  // dF/dchiN and, for the AB diblock, dF/dfA at a converged state
  void paramGrad(double* dFdchiN, double* dFdfA) { check(psx_scft_param_grad(e_, dFdchiN, dFdfA)); }
  // Reverse mode of (w_A, w_B, box) -> (phi_A, phi_B, Q) for
  // L = <gA, phi_A> + <gB, phi_B> + gQ Q: dL/dw_A, dL/dw_B and, when
  // dLdbox is not null, dL/dL_x,y,z (3 values).  Empty gA / gB count as 0.
  void backward(const std::vector<double>& gA, const std::vector<double>& gB, double gQ,
                std::vector<double>& dLdwA, std::vector<double>& dLdwB, std::vector<double>* dLdbox = nullptr) {
    if ((!gA.empty() && gA.size() != n_) || (!gB.empty() && gB.size() != n_))
      throw Error(PSX_ERR_SHAPE, "cotangent size differs from the grid");
    dLdwA.resize(n_); dLdwB.resize(n_);
    if (dLdbox) dLdbox->resize(3);
    check(psx_scft_backward_host(e_, gA.empty() ? nullptr : gA.data(), gB.empty() ? nullptr : gB.data(), gQ,
                                 dLdwA.data(), dLdwB.data(), dLdbox ? dLdbox->data() : nullptr));
  }
  psx_scft* handle() { return e_; }

 protected:
  Engine(psx_scft* e, size_t n, bool own) : e_(e), n_(n), own_(own) {}
  psx_scft* e_ = nullptr;
  size_t n_ = 0;
  bool own_ = true;
  static size_t cells(const psx_grid& g) { return (size_t)g.nx * g.ny * g.nz; }
  friend class Batch;
};

// Bulk AB diblock: PsBlockCopolymer with two PsFlexPseudoSpec blocks.
class DiblockSolver : public Engine {
 public:
  DiblockSolver(const psx_grid& g, const psx_diblock& c, psx_prec prec = PSX_FP64)
      : Engine(create(g, c, prec), cells(g), true) {}

 private:
  static psx_scft* create(const psx_grid& g, const psx_diblock& c, psx_prec prec) {
    psx_scft* e = nullptr;
    check(psx_scft_create(&g, &c, prec, nullptr, &e));
    return e;
  }
};

// Any tree of A and B blocks sharing one ds (linear multiblocks, stars):
// a blockCopolymer Polymer with its headjoined / tailjoined Blocks
// (pspolymer/PsBlockCopolymer.cpp:60-182).
class BlockCopolymerSolver : public Engine {
 public:
  BlockCopolymerSolver(const psx_grid& g, const std::vector<psx_block>& blocks, double ds, double chiN,
                       psx_prec prec = PSX_FP64)
      : Engine(create(g, blocks, ds, chiN, prec), cells(g), true) {}

 private:
  static psx_scft* create(const psx_grid& g, const std::vector<psx_block>& b, double ds, double chiN,
                          psx_prec prec) {
    psx_scft* e = nullptr;
    check(psx_scft_create_blocks(&g, (int)b.size(), b.data(), ds, chiN, prec, nullptr, &e));
    return e;
  }
};

// A blend of block copolymers on the same A and B fields (PsPolymerHldr with
// several Polymers, pspolymer/PsPolymerHldr.cpp:35-83): each polymer takes
// the next nblocks entries of the block list.
class BlendSolver : public Engine {
 public:
  BlendSolver(const psx_grid& g, const std::vector<psx_polymer>& polymers, const std::vector<psx_block>& blocks,
              double ds, double chiN, psx_prec prec = PSX_FP64)
      : Engine(create(g, polymers, blocks, ds, chiN, prec), cells(g), true) {}

 private:
  static psx_scft* create(const psx_grid& g, const std::vector<psx_polymer>& p, const std::vector<psx_block>& b,
                          double ds, double chiN, psx_prec prec) {
    size_t nb = 0;
    for (const auto& q : p) nb += (size_t)q.nblocks;
    if (nb != b.size()) throw Error(PSX_ERR_INVALID, "the polymers' block counts differ from the block list");
    psx_scft* e = nullptr;
    check(psx_scft_create_blend(&g, (int)p.size(), p.data(), b.data(), ds, chiN, prec, nullptr, &e));
    return e;
  }
};

// Diblock parameter sweeps advanced together, one hipGraph per iteration
// (no reference counterpart: the reference runs one system per process).
class Batch {
 public:
  Batch(const std::vector<psx_grid>& grids, const std::vector<psx_diblock>& chains, psx_prec prec = PSX_FP32) {
    if (grids.size() != chains.size() || grids.empty())
      throw Error(PSX_ERR_INVALID, "one grid per chain, at least one instance");
    check(psx_batch_create((int)grids.size(), grids.data(), chains.data(), prec, nullptr, &b_));
    try {
      for (size_t i = 0; i < grids.size(); ++i) {
        psx_scft* e = nullptr;
        check(psx_batch_instance(b_, (int)i, &e));
        inst_.emplace_back(Engine(e, Engine::cells(grids[i]), false));
      }
    } catch (...) {
      inst_.clear();
      psx_batch_destroy(b_);
      throw;
    }
  }
  ~Batch() { inst_.clear(); psx_batch_destroy(b_); }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  size_t size() const { return inst_.size(); }
  int sizeClasses() const {
    int n = 0, c = 0;
    check(psx_batch_size(b_, &n, &c));
    return c;
  }
  // instance i (owned by the batch): fields, densities, stress
  Engine& instance(size_t i) { return inst_.at(i); }
  void setPrecision(psx_prec prec) { check(psx_batch_set_precision(b_, prec)); }
  void autotune(int reps = 5) { check(psx_batch_autotune(b_, reps, nullptr)); }
  // n_iters iterations of every instance; F[i] holds instance i's history
  int iterate(const psx_mixing& mix, int n_iters, std::vector<std::vector<double>>* F = nullptr) {
    std::vector<double> f(inst_.size() * (size_t)n_iters), e(f.size());
    int done = 0;
    check(psx_batch_iterate(b_, &mix, n_iters, 1, f.data(), e.data(), &done));
    if (F) {
      F->assign(inst_.size(), {});
      for (size_t i = 0; i < inst_.size(); ++i)
        (*F)[i].assign(f.begin() + i * n_iters, f.begin() + i * n_iters + done);
    }
    return done;
  }
  psx_batch* handle() { return b_; }

 private:
  psx_batch* b_ = nullptr;
  std::vector<Engine> inst_;
};

}  // namespace psx
