// psx_filter.hpp -- element-wise and reduction kernels of the FFT plugin's
// other consumers: the spectral filters (pseffhamil/PsSpecFilterUpdater.cpp,
// PsSimpleSpecFilter.cpp, PsMultiSpecFilter.cpp) and the Poisson updater
// (pseffhamil/PsPoissonUpdater.cpp).  The transforms themselves are the
// plugin's line passes (k_cplx_lines); these kernels are the O(N) steps
// between them.  All grid-stride, kThreads per block.
#pragma once
#include "psx_kernels.hpp"

namespace psx {

// a <- a * b on interleaved complex spectra (psfft/PsFFTW.cpp:164-167).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_spec_cmul(T* __restrict__ a, const T* __restrict__ b, long long N) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x) {
    const T ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
    a[2 * i] = ar * br - ai * bi;
    a[2 * i + 1] = ai * br + ar * bi;
  }
}

// out[0] = sum of the nb partial sums (one block).
__global__ void __launch_bounds__(kThreads)
k_sum_partials(const double* __restrict__ part, int nb, double* __restrict__ out) {
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < nb; i += blockDim.x) v[0] += part[i];
  block_sum<1>(v);
  if (threadIdx.x == 0) out[0] = v[0];
}

// x <- x - mean, mean = sum / vol in T: the S(0) removal of
// PsSpecFilterUpdater::subtractAverage (PsSpecFilterUpdater.cpp:225-239) and
// the charge shift of PsPoissonUpdater::update (PsPoissonUpdater.cpp:230-262).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_shift_mean(T* __restrict__ x, long long N, const double* __restrict__ sum, double vol) {
  const T mean = (T)sum[0] / (T)vol;
  const T shift = (T)-1.0 * mean;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x)
    x[i] = x[i] + shift;
}

// Fill n words with v (region maxima start at +0).
__global__ void __launch_bounds__(kThreads)
k_fill_u64(unsigned long long* __restrict__ p, long long n, unsigned long long v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// rmax[r] = max of res over the cells of spectral region r.  res >= 0 (a
// squared modulus), and for non-negative doubles the IEEE bit patterns order
// like the values, so the maximum is an integer atomicMax on the bits.
// region == nullptr: one region (block maximum, one atomic per block).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_region_max(const T* __restrict__ res, const int* __restrict__ region, long long N,
             unsigned long long* __restrict__ rmax) {
  if (region) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x)
      atomicMax(&rmax[region[i]], (unsigned long long)__double_as_longlong((double)res[i]));
    return;
  }
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x)
    m = fmax(m, (double)res[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&rmax[0], (unsigned long long)__double_as_longlong(m));
}

// mask = res < cutoffFactor * max(region) ? filterStrength : 1
// (PsSimpleSpecFilter.cpp:121-131; PsMultiSpecFilter.cpp:243-259: the
// per-region maximum of cutoffFactor * res equals cutoffFactor times the
// maximum of res, rounding being monotone).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_filter_mask(const T* __restrict__ res, const int* __restrict__ region,
              const unsigned long long* __restrict__ rmax, T cutoff, T strength,
              T* __restrict__ mask, long long N) {
  const T gcut = cutoff * (T)__longlong_as_double((long long)rmax[0]);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x) {
    const T cut = region ? cutoff * (T)__longlong_as_double((long long)rmax[region[i]]) : gcut;
    mask[i] = res[i] < cut ? strength : (T)1;
  }
}

// w = scale * res (PsSimpleSpecFilter.cpp:138-142).
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_scale_store(const T* __restrict__ res, T scale, T* __restrict__ w, long long N) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x)
    w[i] = scale * res[i];
}

// The potential update of PsPoissonUpdater::update (PsPoissonUpdater.cpp:
// 267-284), including its limiter: where the new value psi1 = res * scale
// changes by more than 20% of the old psi0, res is replaced by 1.2 psi0,
// and the stored field is then scale * res (the field's scale() goes through
// a double temporary) -- so a limited cell holds 1.2 psi0 / N, as the
// reference computes it.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_poisson_psi(const T* __restrict__ res, T* __restrict__ psi, T scale, long long N) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x) {
    const T psi0 = psi[i];
    const T psi1 = res[i] * scale;
    const T frac = fabs((psi1 - psi0) / psi0);
    T r = res[i];
    if ((double)frac > 0.2) r = (T)((double)psi0 * 1.2);
    psi[i] = (T)((double)scale * (double)r);
  }
}

}  // namespace psx
