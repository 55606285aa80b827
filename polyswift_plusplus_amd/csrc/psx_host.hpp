// psx_host.hpp -- C++ host wrapper over the psx C ABI (include/psx_scft.h).
//
// The classes sit where the reference's plugins sit and keep their method
// names, so a PolySwift++ build can call them in place of its own:
//   psx::HipFFT          ~ PsFFTBase<double,NDIM> / PsFFTW (psbase/PsFFTBase.h:39-147)
//   psx::DiblockSolver   ~ PsBlockCopolymer + two PsFlexPseudoSpec blocks
//                          (pspolymer/PsBlockCopolymer.cpp:189-239) with the
//                          flory interaction and steepestDescent updater
//                          (pseffhamil/PsSteepDUpdater.cpp:159-203)
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

// FFT pair plugin: scaledFFTPair(data, kdata, res) on host arrays of
// getFFTSize() elements, unnormalised, real part kept.
class HipFFT {
 public:
  HipFFT(int nx, int ny, int nz, psx_prec prec = PSX_FP64) { check(psx_fft_plan_create(nx, ny, nz, prec, nullptr, &p_)); }
  ~HipFFT() { psx_fft_plan_destroy(p_); }
  HipFFT(const HipFFT&) = delete;
  HipFFT& operator=(const HipFFT&) = delete;
  size_t getFFTSize() const { return psx_fft_size(p_); }
  void scaledFFTPair(const double* data, const double* kdata, double* resPtr) {
    check(psx_fft_scaled_pair_host(p_, data, kdata, resPtr));
  }
  // device-pointer form (asynchronous on the plan's stream)
  void scaledFFTPairDevice(const void* d_data, const void* d_kdata, void* d_res) {
    check(psx_fft_scaled_pair(p_, d_data, d_kdata, d_res));
  }

 private:
  psx_fft* p_ = nullptr;
};

// Bulk AB diblock: reset + solveQ of both blocks, calcBigQ, the quadrature,
// and the field updaters, on one HIP device.
class DiblockSolver {
 public:
  DiblockSolver(const psx_grid& g, const psx_diblock& c, psx_prec prec = PSX_FP64) {
    check(psx_scft_create(&g, &c, prec, nullptr, &e_));
    n_ = (size_t)g.nx * g.ny * g.nz;
  }
  ~DiblockSolver() { psx_scft_destroy(e_); }
  DiblockSolver(const DiblockSolver&) = delete;
  DiblockSolver& operator=(const DiblockSolver&) = delete;

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
  void setPhysFields(std::vector<double>& phiA, std::vector<double>& phiB) {
    phiA.resize(n_); phiB.resize(n_);
    check(psx_scft_densities(e_, nullptr, nullptr));
    check(psx_scft_get_densities_host(e_, phiA.data(), phiB.data()));
  }
  double steepestDescent(double lambda0, double lambda1) {
    double F;
    check(psx_scft_update_steepest(e_, lambda0, lambda1, &F));
    return F;
  }
  // n_iters SCFT iterations (PsDomain::update); returns iterations run.
  int iterate(const psx_mixing& mix, int n_iters, double tol, std::vector<double>* F = nullptr,
              std::vector<double>* err = nullptr) {
    std::vector<double> f(n_iters), q(n_iters), e(n_iters);
    int done = 0;
    check(psx_scft_iterate(e_, &mix, n_iters, tol, 1, f.data(), q.data(), e.data(), &done));
    if (F) F->assign(f.begin(), f.begin() + done);
    if (err) err->assign(e.begin(), e.begin() + done);
    return done;
  }
  psx_scft* handle() { return e_; }

 private:
  psx_scft* e_ = nullptr;
  size_t n_ = 0;
};

}  // namespace psx
