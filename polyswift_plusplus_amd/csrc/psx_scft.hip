// psx_scft.hip -- host side of the C ABI declared in include/psx_scft.h.
//
// One engine = one bulk AB block-copolymer melt (a diblock, or any tree of A
// and B blocks) on one HIP device.  All fields stay resident in HBM.  The
// block propagators are grouped into paths (contiguous slice arrays); a
// contour step of a path is one round-trip line FFT launch per active axis
// (psx_fft_reg.hpp), with the exp(-ds w / 2) multiplies fused in and the
// boundary launches of consecutive steps fused in pairs.  Independent paths
// (the forward and backward sweeps of a chain, the arms of a star) run on
// their own streams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psx_scft.h"
#include "psx_kernels.hpp"
#include "psx_fft_reg.hpp"
#include "psx_fft_cline.hpp"
#include "psx_fft_plane.hpp"
#include "psx_filter.hpp"
#include "psx_rng.hpp"

using namespace psx;

namespace {

thread_local std::string g_err;

psx_status fail(psx_status code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define PSX_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return fail(PSX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PSX_TRY(call)                                                         \
  do {                                                                        \
    psx_status s_ = (call);                                                   \
    if (s_ != PSX_OK) return s_;                                              \
  } while (0)

size_t prec_bytes(int prec) { return prec == PSX_FP64 ? 8 : 4; }

// Radix plan: 8s first, then one 4 and one 2, then odd radices <= 13.
bool radix_plan(int n, FftPlanDev* pl) {
  pl->n = n;
  pl->nst = 0;
  int m = n;
  auto push = [&](int r) {
    if (pl->nst >= 8) return false;
    pl->rad[pl->nst++] = r;
    m /= r;
    return true;
  };
  while (m % 8 == 0) if (!push(8)) return false;
  if (m % 4 == 0) if (!push(4)) return false;
  if (m % 2 == 0) if (!push(2)) return false;
  for (int p : {3, 5, 7, 11, 13})
    while (m % p == 0) if (!push(p)) return false;
  return m == 1;
}

// Twiddles exp(-2 pi i j / n) in long double, rounded once.
std::vector<double> twiddles(int n) {
  std::vector<double> t(2 * (size_t)n);
  const long double twopi = 6.283185307179586476925286766559L;
  for (int j = 0; j < n; ++j) {
    long double a = twopi * (long double)j / (long double)n;
    t[2 * j] = (double)cosl(a);
    t[2 * j + 1] = (double)-sinl(a);
  }
  return t;
}

template <typename T>
std::vector<T> cast_vec(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

psx_status upload(void** dptr, const void* h, size_t bytes) {
  if (hipMalloc(dptr, bytes) != hipSuccess) return fail(PSX_ERR_NOMEM, "hipMalloc failed");
  PSX_HIP(hipMemcpy(*dptr, h, bytes, hipMemcpyHostToDevice));
  return PSX_OK;
}

psx_status dalloc(void** dptr, size_t bytes) {
  *dptr = nullptr;
  if (bytes == 0) bytes = 16;
  if (hipMalloc(dptr, bytes) != hipSuccess) {
    *dptr = nullptr;
    return fail(PSX_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return PSX_OK;
}

void dfree(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// One line pass along one grid axis.
struct Axis {
  int axis = 2;       // 0 = x, 1 = y, 2 = z
  int n = 1;
  int inner = 1;
  int nlines = 1;
  bool contig = true;
  FftPlanDev pl{};
  void* tw[2] = {nullptr, nullptr};  // per precision
};

psx_status make_axes(const int dims[3], std::vector<Axis>* axes) {
  axes->clear();
  const long long N = (long long)dims[0] * dims[1] * dims[2];
  for (int d = 2; d >= 0; --d) {
    if (dims[d] == 1) continue;
    Axis a;
    a.axis = d;
    a.n = dims[d];
    long long inner = 1;
    for (int e = d + 1; e < 3; ++e) inner *= dims[e];
    a.inner = (int)inner;
    a.nlines = (int)(N / dims[d]);
    a.contig = inner == 1;
    if (!radix_plan(a.n, &a.pl))
      return fail(PSX_ERR_SHAPE, "grid size " + std::to_string(a.n) +
                                     " has a prime factor above 13 (unsupported)");
    axes->push_back(a);
  }
  if (axes->empty()) {  // 1x1x1: one identity pass of length 1
    Axis a;
    a.axis = 2; a.n = 1; a.inner = 1; a.nlines = 1; a.contig = true;
    radix_plan(1, &a.pl);
    axes->push_back(a);
  }
  return PSX_OK;
}

psx_status axis_tw(Axis& a, int prec) {
  if (a.tw[prec]) return PSX_OK;
  std::vector<double> t = twiddles(a.n);
  if (prec == PSX_FP64) return upload(&a.tw[prec], t.data(), t.size() * sizeof(double));
  std::vector<float> f = cast_vec<float>(t);
  return upload(&a.tw[prec], f.data(), f.size() * sizeof(float));
}

// Tile: real lines per block of the pair kernel (even).
int pair_tile(const Axis& a, int prec) {
  const int target = prec == PSX_FP64 ? 2048 : 4096;
  int L = target / a.n;
  if (L < 2) L = 2;
  if (L > 64) L = 64;
  L &= ~1;
  int cap = a.nlines + (a.nlines & 1);
  if (L > cap) L = cap < 2 ? 2 : cap;
  return L;
}

int cplx_tile(const Axis& a) {
  int P = 1024 / a.n;
  if (P < 1) P = 1;
  if (P > 32) P = 32;
  if (P > a.nlines) P = a.nlines;
  return P;
}

LineGeom geom(const Axis& a, int prec, int L) {
  LineGeom g;
  g.pl = a.pl;
  g.inner = a.inner;
  g.nlines = a.nlines;
  g.L = L;
  g.ls = a.n + 1;
  g.tw = a.tw[prec];
  g.ntiles = 1;
  g.nfields = 1;
  return g;
}

template <typename T, int N1, int N2>
void launch_reg(const Axis& a, int prec, const PairIO* fio, int nfields,
                hipStream_t s) {
  constexpr int RL = 2 * 64 / N2;
  LineGeom g = geom(a, prec, RL);
  g.ntiles = (a.nlines + RL - 1) / RL;
  g.nfields = nfields;
  // transpose buffer (64 (N1 + 1) complex) or the staged rows (RL (n + N2/2) reals)
  const size_t tbuf = (size_t)TBuf<N1, N2>::wave * 2 * sizeof(T);
  const size_t sbuf = (size_t)RL * (N1 * N2 + N2 / 2) * sizeof(T);
  // FP32: + LDS copies of the twiddles (n complex) and two spectral multipliers
  // (pair tables for lane-pair splits)
  const size_t smem = (tbuf > sbuf ? tbuf : sbuf) + (sizeof(T) == 4 ? (size_t)table_bytes<T, N1, N2>(2) : 0);
  dim3 grid(g.ntiles * nfields);
  if (a.contig)
    hipLaunchKernelGGL((k_line_reg<T, N1, N2, true, false>), grid, dim3(64), smem, s, g, fio);
  else if (a.inner % RL == 0)
    hipLaunchKernelGGL((k_line_reg<T, N1, N2, false, true>), grid, dim3(64), smem, s, g, fio);
  else
    hipLaunchKernelGGL((k_line_reg<T, N1, N2, false, false>), grid, dim3(64), smem, s, g, fio);
}

// The tiled strided kernel applies at n = 64, 96, 128 when whole tile rows
// of lt lines fit the line stride (128-byte rows: 32 FP32 / 16 FP64 lines).
template <typename T>
bool tile_fits(const Axis& a, int lt = sizeof(T) == 4 ? 32 : 16) {
  return (a.n == 128 || a.n == 96 || a.n == 64) && !a.contig && a.inner % lt == 0;
}

// Strided axis through 128-byte tile rows (k_line_tile): W waves of RL lines.
template <typename T, int N1, int N2, int W>
void launch_tile(const Axis& a, int prec, const PairIO* fio, int nfields, hipStream_t s) {
  constexpr int n = N1 * N2, RL = 2 * 64 / N2, LT = W * RL;
  LineGeom g = geom(a, prec, LT);
  g.ntiles = a.nlines / LT;
  g.nfields = nfields;
  const size_t sb = (size_t)n * (LT + 2) * sizeof(T);
  const size_t tb = (size_t)W * TBuf<N1, N2>::wave * 2 * sizeof(T);
  const size_t buf = ((sb > tb ? sb : tb) + 15) / 16 * 16;
  const size_t smem = buf + (sizeof(T) == 4 ? (size_t)table_bytes<T, N1, N2>(2) : 0);
  hipLaunchKernelGGL((k_line_tile<T, N1, N2, W>), dim3(g.ntiles * nfields), dim3(64 * W), smem, s, g, fio);
}

template <typename T, int NN>
void launch_lane(const Axis& a, int prec, const PairIO* fio, int nfields,
                 hipStream_t s) {
  LineGeom g = geom(a, prec, 2);
  g.ntiles = (a.nlines + 127) / 128;
  g.nfields = nfields;
  dim3 grid(g.ntiles * nfields);
  if (a.contig)
    hipLaunchKernelGGL((k_line_lane<T, NN, true>), grid, dim3(64), 0, s, g, fio);
  else
    hipLaunchKernelGGL((k_line_lane<T, NN, false>), grid, dim3(64), 0, s, g, fio);
}

// Generic Stockham LDS kernel: any n with prime factors <= 13.
template <typename T>
void launch_stockham(const Axis& a, int prec, const PairIO* fio, int nfields,
                     hipStream_t s) {
  const int L = pair_tile(a, prec);
  LineGeom g = geom(a, prec, L);
  g.ntiles = (a.nlines + L - 1) / L;
  g.nfields = nfields;
  const int P = L / 2;
  const size_t smem = 2 * (size_t)P * g.ls * 2 * sizeof(T);
  dim3 grid(g.ntiles * nfields);
  if (a.contig)
    hipLaunchKernelGGL((k_pair_roundtrip<T, true>), grid, dim3(kThreads), smem, s, g, fio);
  else
    hipLaunchKernelGGL((k_pair_roundtrip<T, false>), grid, dim3(kThreads), smem, s, g, fio);
}

// Round-trip pass along one axis: register-resident kernels for the line
// lengths of the workloads, the LDS Stockham kernel for everything else.
// PSX_FORCE_STOCKHAM=1 in the environment selects the Stockham kernel for
// every length (used to cross-check the two kernel families).
int g_force_stockham = -1;
// PSX_LINE_VARIANT=<v> makes variant v the default for every pass (kernel
// cross-checks and timing; an engine's tuned choice still wins).
int g_default_variant = [] {
  const char* v = getenv("PSX_LINE_VARIANT");
  return v ? atoi(v) : 0;
}();

template <typename T>
psx_status launch_pair(const Axis& a, int prec, const PairIO* fio,
                       int nfields, hipStream_t s, int variant = -1) {
  if (g_force_stockham < 0) {
    const char* v = getenv("PSX_FORCE_STOCKHAM");
    g_force_stockham = (v && v[0] == '1') ? 1 : 0;
  }
  if (variant < 0 && g_default_variant != 0) variant = g_default_variant;
  if (g_force_stockham == 1 || variant == 1) {
    launch_stockham<T>(a, prec, fio, nfields, s);
  } else if (variant == 2 && a.n == 64) {
    launch_reg<T, 16, 4>(a, prec, fio, nfields, s);
  } else if (variant == 2 && a.n == 128) {
    launch_reg<T, 32, 4>(a, prec, fio, nfields, s);
  } else if (variant == 2 && a.n == 96) {
    launch_reg<T, 24, 4>(a, prec, fio, nfields, s);
  } else if ((variant == 4 || (variant < 0 && sizeof(T) == 4)) && tile_fits<T>(a)) {
    // 128-byte tile rows: FP32 default on strided passes (measured at 128^3:
    // 10.7 -> 7.5 us per y pass of both sweeps); for FP64 a tuning candidate
    if (a.n == 96) {
      if (sizeof(T) == 4) launch_tile<T, 12, 8, 2>(a, prec, fio, nfields, s);
      else launch_tile<T, 12, 8, 1>(a, prec, fio, nfields, s);
    } else if (a.n == 128) {
      if (sizeof(T) == 4) launch_tile<T, 8, 16, 4>(a, prec, fio, nfields, s);
      else launch_tile<T, 8, 16, 2>(a, prec, fio, nfields, s);
    } else {
      if (sizeof(T) == 4) launch_tile<T, 8, 8, 2>(a, prec, fio, nfields, s);
      else launch_tile<T, 8, 8, 1>(a, prec, fio, nfields, s);
    }
  } else if (variant == 5 && a.n == 128) {
    launch_reg<T, 8, 16>(a, prec, fio, nfields, s);
  } else if (variant == 3 && a.n == 128) {
    launch_reg<T, 16, 8>(a, prec, fio, nfields, s);
  } else if (variant == 3 && a.n == 32) {
    launch_reg<T, 4, 8>(a, prec, fio, nfields, s);
  } else {
    switch (a.n) {
      case 1: launch_lane<T, 1>(a, prec, fio, nfields, s); break;
      case 2: launch_lane<T, 2>(a, prec, fio, nfields, s); break;
      case 3: launch_lane<T, 3>(a, prec, fio, nfields, s); break;
      case 4: launch_lane<T, 4>(a, prec, fio, nfields, s); break;
      case 5: launch_lane<T, 5>(a, prec, fio, nfields, s); break;
      case 8: launch_lane<T, 8>(a, prec, fio, nfields, s); break;
      // 8 lanes per line where the length allows (8 complex lines per wave:
      // 128-byte rows in FP64); fewer points per lane keep more waves in
      // flight, which matters more here than FP32 row width.
      case 16: launch_reg<T, 4, 4>(a, prec, fio, nfields, s); break;
      case 32: launch_reg<T, 8, 4>(a, prec, fio, nfields, s); break;
      case 48: launch_reg<T, 12, 4>(a, prec, fio, nfields, s); break;
      case 64: launch_reg<T, 8, 8>(a, prec, fio, nfields, s); break;
      // general lane pairs at 96: 8 lanes of 12 points, 16 real lines per
      // wave (measured at 96^3: FP32 step 35.3 -> 26.0 us, FP64 51.5 ->
      // 34.7 us against 4 lanes of 24 points)
      case 96: launch_reg<T, 12, 8>(a, prec, fio, nfields, s); break;
      // lane pairs at 128: 8 points per lane, 3 (FP64) / 5 (FP32) waves per
      // SIMD against 2 for 16 points per lane
      case 128: launch_reg<T, 8, 16>(a, prec, fio, nfields, s); break;
      case 256: launch_reg<T, 16, 16>(a, prec, fio, nfields, s); break;
      default: launch_stockham<T>(a, prec, fio, nfields, s); break;
    }
  }
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// Plane launches (k_plane): the z and y round trips of whole x-planes, for
// 3D grids with n_y = n_z = n of a register split below.
bool plane_length(int n) { return n == 32 || n == 48 || n == 64 || n == 96 || n == 128; }

template <typename T, int N1, int N2, int NW, bool REUSE, bool DIRECT = false, int WPE = 0>
psx_status plane_kernel(const Axis& az, int prec, const PairIO* fio, int nfields, int nplanes,
                        hipStream_t s, bool prepare) {
  typedef PlaneShape<T, N1, N2, NW, REUSE> S;
  static_assert(S::SMEM <= 160 * 1024, "plane kernel exceeds the LDS of a CU");
  if (prepare) {
    PSX_HIP(hipFuncSetAttribute((const void*)k_plane<T, N1, N2, NW, REUSE, DIRECT, WPE>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, S::SMEM));
    return PSX_OK;
  }
  LineGeom g = geom(az, prec, 0);
  g.ntiles = nplanes;
  g.nfields = nfields;
  hipLaunchKernelGGL((k_plane<T, N1, N2, NW, REUSE, DIRECT, WPE>), dim3(nplanes * nfields), dim3(64 * NW), S::SMEM, s,
                     g, fio);
  return PSX_OK;
}


// Waves per SIMD the aliased plane kernel is compiled for (0: unconstrained).
template <typename T, int N1, int N2>
constexpr int wpe_aliased() { return sizeof(T) == 4 && N1 == 12 && N2 == 8 ? 5 : 0; }

// CUs of the current device (queried once).
static int g_device_cus = 0;
static int device_cus() {
  if (g_device_cus == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      g_device_cus = v;
    else
      g_device_cus = 256;
  }
  return g_device_cus;
}

// FP32 planes at n = 64 and 96 fit either buffer layout.  Behind the plane
// the transpose buffers need no block barriers but twice the LDS; aliasing
// the plane (REUSE) fits twice the blocks per CU.  A launch with no more
// blocks than the CUs hold in the barrier-free layout takes it (single 96^3
// FP32 engine: 18.4 -> 17.5 us per dual step); a larger one aliases (the
// 96^3 x 19-instance batch class: 20.0 -> 19.5 ms per iteration).  The
// aliased (12, 8) kernel is compiled for five waves per SIMD (86 VGPRs, no
// spills; 101 unconstrained), so three of its 41 KB six-wave blocks share a
// CU instead of two: the 96^3 class 18.85 -> 18.26 ms per batch iteration.
template <typename T, int N1, int N2, int NW>
psx_status plane_kernel_fit(const Axis& az, int prec, const PairIO* fio, int nfields, int nplanes,
                            hipStream_t s, bool prepare) {
  if (prepare) {
    PSX_TRY((plane_kernel<T, N1, N2, NW, false>(az, prec, fio, nfields, nplanes, s, true)));
    return plane_kernel<T, N1, N2, NW, true, false, wpe_aliased<T, N1, N2>()>(az, prec, fio, nfields, nplanes, s,
                                                                              true);
  }
  constexpr int per_cu = 160 * 1024 / PlaneShape<T, N1, N2, NW, false>::SMEM;
  if ((long long)nfields * nplanes <= (long long)per_cu * device_cus())
    return plane_kernel<T, N1, N2, NW, false>(az, prec, fio, nfields, nplanes, s, false);
  return plane_kernel<T, N1, N2, NW, true, false, wpe_aliased<T, N1, N2>()>(az, prec, fio, nfields, nplanes, s,
                                                                            false);
}

// One plane launch over nfields fields of nplanes planes; az is the z axis
// (its twiddles serve y too).  prepare: only raise the kernel's LDS limit
// (once per length and type, outside any stream capture).  FP32 at n = 64
// and 96 picks the buffer layout by launch size (plane_kernel_fit); the
// other kernels keep the transpose buffers behind the plane where both fit
// in LDS, else alias it.  FP64 at 128 runs the lane pairs (8, 16) on 16
// waves: with the fused pair operator they fit the 128 VGPRs of a
// 1024-thread block (120, no spills; the three-step form spilled 136 bytes
// per lane), and against the (16, 8) split on 8 waves the plane pass of a
// dual step went from 18.3 to 16.2-16.5 us.
template <typename T>
psx_status launch_plane(const Axis& az, int prec, const PairIO* fio, int nfields, int nplanes,
                        hipStream_t s, bool prepare = false) {
  constexpr bool F32 = sizeof(T) == 4;
  switch (az.n) {
    case 32: PSX_TRY((plane_kernel<T, 8, 4, 1, false>(az, prec, fio, nfields, nplanes, s, prepare))); break;
    // 2 waves of 16 lines for 24: the second wave's last 8 lines are idle
    case 48: PSX_TRY((plane_kernel<T, 12, 4, 2, false>(az, prec, fio, nfields, nplanes, s, prepare))); break;
    case 64:
      if constexpr (F32) PSX_TRY((plane_kernel_fit<T, 8, 8, 4>(az, prec, fio, nfields, nplanes, s, prepare)));
      else PSX_TRY((plane_kernel<T, 8, 8, 4, false>(az, prec, fio, nfields, nplanes, s, prepare)));
      break;
    // FP64 at 96 and 128: the first and last y round trips straight from and
    // to global memory (DIRECT; a wave's column pairs are 128-byte row
    // segments at 96, 64-byte ones at 128): measured with 128-byte segments,
    // 128^3 plane launches 19.7 -> 18.4 us per dual step, 96^3 13.9 ->
    // 13.1 us.  FP32 rows of 32 or 64 bytes lost (128^3 10.4 ->
    // 11.9 us; 96^3 8.9 -> 9.7 us; 64^3 4.4 -> 5.0 us), and so did the
    // (16, 8) FP32 split at 128 with them (13.2 us aliased, 12.7 us with the
    // transposes behind the plane).
    case 96:
      if constexpr (F32) PSX_TRY((plane_kernel_fit<T, 12, 8, 6>(az, prec, fio, nfields, nplanes, s, prepare)));
      else PSX_TRY((plane_kernel<T, 12, 8, 6, false, true>(az, prec, fio, nfields, nplanes, s, prepare)));
      break;
    case 128:
      if constexpr (F32) PSX_TRY((plane_kernel<T, 8, 16, 16, false>(az, prec, fio, nfields, nplanes, s, prepare)));
      else PSX_TRY((plane_kernel<T, 8, 16, 16, true, true>(az, prec, fio, nfields, nplanes, s, prepare)));
      break;
    default: return fail(PSX_ERR_INVALID, "no plane kernel for n = " + std::to_string(az.n));
  }
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// Register-resident complex line pass (k_cline): 64 / N2 lines per wave.
template <typename T, int N1, int N2>
void launch_cline(const Axis& a, int prec, const CplxIO& f, hipStream_t s) {
  constexpr int CL = 64 / N2;
  LineGeom g = geom(a, prec, CL);
  g.ntiles = (a.nlines + CL - 1) / CL;
  const size_t smem = (size_t)TBuf<N1, N2>::wave * 2 * sizeof(T);
  const dim3 grid(g.ntiles);
  if (f.do_fwd && f.do_inv) hipLaunchKernelGGL((k_cline<T, N1, N2, 2>), grid, dim3(64), smem, s, g, f);
  else if (f.do_fwd) hipLaunchKernelGGL((k_cline<T, N1, N2, 0>), grid, dim3(64), smem, s, g, f);
  else hipLaunchKernelGGL((k_cline<T, N1, N2, 1>), grid, dim3(64), smem, s, g, f);
}

// One pass of the general FFT pair: the register kernels for the line
// lengths that have a split, the LDS Stockham kernel otherwise (and for
// every length under PSX_FORCE_STOCKHAM=1).
template <typename T>
psx_status launch_cplx(const Axis& a, int prec, const CplxIO& f, hipStream_t s) {
  if (g_force_stockham < 0) {
    const char* v = getenv("PSX_FORCE_STOCKHAM");
    g_force_stockham = (v && v[0] == '1') ? 1 : 0;
  }
  if (g_force_stockham != 1 && (f.do_fwd || f.do_inv)) {
    bool done = true;
    switch (a.n) {
      case 16: launch_cline<T, 4, 4>(a, prec, f, s); break;
      case 32: launch_cline<T, 8, 4>(a, prec, f, s); break;
      case 48: launch_cline<T, 12, 4>(a, prec, f, s); break;
      case 64: launch_cline<T, 8, 8>(a, prec, f, s); break;
      case 96: launch_cline<T, 12, 8>(a, prec, f, s); break;
      case 128: launch_cline<T, 8, 16>(a, prec, f, s); break;
      case 256: launch_cline<T, 16, 16>(a, prec, f, s); break;
      default: done = false;
    }
    if (done) {
      PSX_HIP(hipGetLastError());
      return PSX_OK;
    }
  }
  const int P = cplx_tile(a);
  LineGeom g = geom(a, prec, P);
  const size_t smem = 2 * (size_t)P * g.ls * 2 * sizeof(T);
  dim3 grid((a.nlines + P - 1) / P);
  if (a.contig)
    hipLaunchKernelGGL((k_cplx_lines<T, true>), grid, dim3(kThreads), smem, s, g, f);
  else
    hipLaunchKernelGGL((k_cplx_lines<T, false>), grid, dim3(kThreads), smem, s, g, f);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// Blocks of the grid-stride cell kernels (and of their partial sums).  At
// most 1024 blocks of 256 threads: the quadrature kernel holds 5 waves per
// SIMD, so 1024 blocks are all resident at once, where 2048 ran as one full
// and one partial round (128^3 FP32 iteration 2.47 -> 2.44 ms; 512 blocks:
// 2.54 ms).
int grid_1d(long long N) {
  long long b = (N + kThreads - 1) / kThreads;
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

psx_status make_stream(void* user, hipStream_t* s, bool* own) {
  if (user) { *s = (hipStream_t)user; *own = false; return PSX_OK; }
  PSX_HIP(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
  *own = true;
  return PSX_OK;
}

}  // namespace

// ===========================================================================
// FFT pair plugin
// ===========================================================================
struct psx_fft {
  int dims[3];
  long long N;
  int prec;
  hipStream_t stream;
  bool own_stream;
  std::vector<Axis> axes;
  void* work = nullptr;   // N complex of T
  void* hbuf[3] = {nullptr, nullptr, nullptr};  // device staging for the host API
  // the spectral consumers (allocated on first use)
  void* work2 = nullptr;           // second spectrum (convolveRe), N complex
  void* rbuf = nullptr;            // N of T: |F|^2, then the filtered field
  void* mbuf = nullptr;            // N of T: filter mask
  void* kbuf = nullptr;            // N of T: Poisson multiplier 4 pi lB / k^2
  double k2key[4] = {0, 0, 0, 0};  // (lB, dx, dy, dz) of kbuf
  bool k2valid = false;
  double* dsum = nullptr;          // reduction partials, then the sum
  unsigned long long* rmax = nullptr;  // per-region maxima (bit patterns)
  int* region = nullptr;           // spectral region of every cell
  int region_cells[3] = {0, 0, 0};
  int nregions = 0;
};

extern "C" void psx_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return;
  std::snprintf(buf, len, "%s", g_err.c_str());
}

extern "C" const char* psx_version(void) { return "psx-scft 0.1 (gfx950)"; }

extern "C" psx_status psx_fft_plan_create(int nx, int ny, int nz, psx_prec prec,
                                          void* stream, psx_fft** out) {
  if (!out) return fail(PSX_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail(PSX_ERR_SHAPE, "grid dims must be >= 1");
  if (prec != PSX_FP64 && prec != PSX_FP32) return fail(PSX_ERR_INVALID, "bad precision");
  if ((long long)nx * ny * nz > (1LL << 30)) return fail(PSX_ERR_SHAPE, "grid too large");
  psx_fft* p = new psx_fft();
  p->dims[0] = nx; p->dims[1] = ny; p->dims[2] = nz;
  p->N = (long long)nx * ny * nz;
  p->prec = prec;
  psx_status st = make_axes(p->dims, &p->axes);
  if (st == PSX_OK) {
    for (auto& a : p->axes) {
      st = axis_tw(a, prec);
      if (st != PSX_OK) break;
    }
  }
  if (st == PSX_OK) st = dalloc(&p->work, 2 * p->N * prec_bytes(prec));
  if (st == PSX_OK) st = make_stream(stream, &p->stream, &p->own_stream);
  if (st != PSX_OK) { psx_fft_plan_destroy(p); return st; }
  *out = p;
  return PSX_OK;
}

extern "C" psx_status psx_fft_plan_destroy(psx_fft* p) {
  if (!p) return PSX_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto& a : p->axes) { dfree(a.tw[0]); dfree(a.tw[1]); }
  dfree(p->work);
  dfree(p->work2); dfree(p->rbuf); dfree(p->mbuf); dfree(p->kbuf);
  dfree((void*&)p->dsum); dfree((void*&)p->rmax); dfree((void*&)p->region);
  for (auto& h : p->hbuf) dfree(h);
  if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return PSX_OK;
}

extern "C" size_t psx_fft_size(const psx_fft* p) { return p ? (size_t)p->N : 0; }

// The pair over `na` axes with `work` (N complex of T) between the passes:
// forward along z, y (in order), then x with the k multiply and its inverse,
// then the inverses back out; the last axis keeps only the real part.
// in_mode 1: data is the real part of the input, 2: its imaginary part
// (scaledFFTPairIm).
template <typename T>
static psx_status fft_pair_core(const Axis* axes, int na, int prec, void* work, hipStream_t s,
                                const void* data, const void* kdata, void* res, int in_mode) {
  for (int i = 0; i < na; ++i) {
    CplxIO f{};
    f.in = (i == 0) ? data : work;
    f.in_mode = (i == 0) ? in_mode : 0;
    f.do_fwd = 1;
    const bool last = (i == na - 1);
    f.kmul = last ? kdata : nullptr;
    f.do_inv = last ? 1 : 0;
    f.out = (last && na == 1) ? res : work;
    f.out_mode = (last && na == 1) ? 1 : 0;
    PSX_TRY(launch_cplx<T>(axes[i], prec, f, s));
  }
  for (int i = na - 2; i >= 0; --i) {
    CplxIO f{};
    f.in = work;
    f.in_mode = 0;
    f.do_fwd = 0;
    f.do_inv = 1;
    f.kmul = nullptr;
    f.out = (i == 0) ? res : work;
    f.out_mode = (i == 0) ? 1 : 0;
    PSX_TRY(launch_cplx<T>(axes[i], prec, f, s));
  }
  return PSX_OK;
}

template <typename T>
static psx_status fft_pair_impl(psx_fft* p, const void* data, const void* kdata, void* res,
                                int in_mode = 1) {
  return fft_pair_core<T>(p->axes.data(), (int)p->axes.size(), p->prec, p->work, p->stream, data,
                          kdata, res, in_mode);
}

extern "C" psx_status psx_fft_scaled_pair(psx_fft* p, const void* d_data,
                                          const void* d_kdata, void* d_res) {
  if (!p || !d_data || !d_kdata || !d_res) return fail(PSX_ERR_INVALID, "null argument");
  if (p->prec == PSX_FP64) return fft_pair_impl<double>(p, d_data, d_kdata, d_res);
  return fft_pair_impl<float>(p, d_data, d_kdata, d_res);
}

extern "C" psx_status psx_fft_scaled_pair_host(psx_fft* p, const void* h_data,
                                               const void* h_kdata, void* h_res) {
  if (!p || !h_data || !h_kdata || !h_res) return fail(PSX_ERR_INVALID, "null argument");
  const size_t bytes = p->N * prec_bytes(p->prec);
  for (auto& h : p->hbuf)
    if (!h) PSX_TRY(dalloc(&h, bytes));
  PSX_HIP(hipMemcpyAsync(p->hbuf[0], h_data, bytes, hipMemcpyHostToDevice, p->stream));
  PSX_HIP(hipMemcpyAsync(p->hbuf[1], h_kdata, bytes, hipMemcpyHostToDevice, p->stream));
  PSX_TRY(psx_fft_scaled_pair(p, p->hbuf[0], p->hbuf[1], p->hbuf[2]));
  PSX_HIP(hipMemcpyAsync(h_res, p->hbuf[2], bytes, hipMemcpyDeviceToHost, p->stream));
  PSX_HIP(hipStreamSynchronize(p->stream));
  return PSX_OK;
}

// ---------------------------------------------------------------------------
// The FFT pair without a plan and without allocation
// (psx_fft_scaled_pair_noalloc).  Its scratch -- the twiddle tables of the
// grid's axes and the N complex values between the passes -- is one
// file-scope device array per GPU.  In host code `g_noalloc_scratch` names a
// host-side stand-in, never the device memory: every use goes through the
// address hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr int kNoallocMaxLine = PSX_NOALLOC_MAX_LINE;
constexpr size_t kNoallocWorkBytes = PSX_NOALLOC_WORK_BYTES;
constexpr size_t kNoallocTwAxis = 2 * (size_t)kNoallocMaxLine * sizeof(double);   // one axis' table
constexpr size_t kNoallocTwBytes = 3 * kNoallocTwAxis;
constexpr size_t kNoallocScratchBytes = kNoallocTwBytes + kNoallocWorkBytes;
static_assert(kNoallocScratchBytes == PSX_NOALLOC_SCRATCH_BYTES, "scratch layout");
static __device__ __align__(256) unsigned char g_noalloc_scratch[kNoallocScratchBytes];

// Twiddle table of one axis, as twiddles() lays it out: exp(-2 pi i j / n)
// for j < n, in T (cosine and sine in FP64, then rounded).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_noalloc_twiddles(int n, T* tw) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    double sn, cs;
    sincospi(2.0 * (double)j / (double)n, &sn, &cs);
    tw[2 * j] = (T)cs;
    tw[2 * j + 1] = (T)(-sn);
  }
}

namespace {
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[64] = {};   // device address of g_noalloc_scratch per HIP device

// Device address of the scratch array on the current device, looked up once
// per device with that device current.
psx_status noalloc_scratch(void** out) {
  int dev = -1;
  PSX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(PSX_ERR_HIP, "device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(g_noalloc_scratch));
    if (rc != hipSuccess || !p)
      return fail(PSX_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return PSX_OK;
}
}  // namespace

template <typename T>
static psx_status noalloc_pair(const int dims[3], int prec, hipStream_t s, const void* data,
                               const void* kdata, void* res) {
  void* base = nullptr;
  PSX_TRY(noalloc_scratch(&base));
  unsigned char* sc = (unsigned char*)base;
  Axis ax[3];
  int na = 0;
  const long long N = (long long)dims[0] * dims[1] * dims[2];
  for (int d = 2; d >= 0; --d) {
    if (dims[d] == 1) continue;
    Axis& a = ax[na];
    a.axis = d;
    a.n = dims[d];
    long long inner = 1;
    for (int e = d + 1; e < 3; ++e) inner *= dims[e];
    a.inner = (int)inner;
    a.nlines = (int)(N / dims[d]);
    a.contig = inner == 1;
    if (!radix_plan(a.n, &a.pl)) return fail(PSX_ERR_SHAPE, "grid size has a prime factor above 13");
    a.tw[prec] = sc + (size_t)na * kNoallocTwAxis;
    ++na;
  }
  if (na == 0) {   // 1x1x1: one identity pass of length 1
    Axis& a = ax[0];
    a.axis = 2; a.n = 1; a.inner = 1; a.nlines = 1; a.contig = true;
    radix_plan(1, &a.pl);
    a.tw[prec] = sc;
    na = 1;
  }
  for (int i = 0; i < na; ++i) {
    const int blocks = (ax[i].n + kThreads - 1) / kThreads;
    hipLaunchKernelGGL((k_noalloc_twiddles<T>), dim3(blocks), dim3(kThreads), 0, s, ax[i].n, (T*)ax[i].tw[prec]);
    PSX_HIP(hipGetLastError());
  }
  return fft_pair_core<T>(ax, na, prec, sc + kNoallocTwBytes, s, data, kdata, res, 1);
}

extern "C" psx_status psx_fft_scaled_pair_noalloc(int nx, int ny, int nz, psx_prec prec, void* stream,
                                                  const void* d_data, const void* d_kdata, void* d_res) {
  if (!d_data || !d_kdata || !d_res) return fail(PSX_ERR_INVALID, "null argument");
  if (prec != PSX_FP64 && prec != PSX_FP32) return fail(PSX_ERR_INVALID, "bad precision");
  if (nx < 1 || ny < 1 || nz < 1) return fail(PSX_ERR_SHAPE, "grid dims must be >= 1");
  if (nx > kNoallocMaxLine || ny > kNoallocMaxLine || nz > kNoallocMaxLine)
    return fail(PSX_ERR_SHAPE, "no-alloc FFT pair: every grid dimension must be <= " +
                                   std::to_string(kNoallocMaxLine));
  const long long N = (long long)nx * ny * nz;
  if ((unsigned long long)N * 2ull * prec_bytes(prec) > kNoallocWorkBytes)
    return fail(PSX_ERR_SHAPE, "no-alloc FFT pair: N = " + std::to_string(N) + " cells need " +
                                   std::to_string(N * 2 * (long long)prec_bytes(prec)) +
                                   " bytes of complex scratch, more than its " +
                                   std::to_string((long long)kNoallocWorkBytes));
  const int dims[3] = {nx, ny, nz};
  hipStream_t s = (hipStream_t)stream;
  if (prec == PSX_FP64) return noalloc_pair<double>(dims, prec, s, d_data, d_kdata, d_res);
  return noalloc_pair<float>(dims, prec, s, d_data, d_kdata, d_res);
}

// ---------------------------------------------------------------------------
// The rest of PsFFTBase (psbase/PsFFTBase.h:83-139) and its consumers
// ---------------------------------------------------------------------------

// Unnormalised forward transform along every axis: the first pass reads `in`
// (in_mode), the last writes `out` (out_mode), the passes between work in
// place in p->work.
template <typename T>
static psx_status fft_passes(psx_fft* p, bool inverse, const void* in, int in_mode, void* out,
                             int out_mode) {
  const int na = (int)p->axes.size();
  for (int j = 0; j < na; ++j) {
    const int i = inverse ? na - 1 - j : j;
    CplxIO f{};
    f.in = j == 0 ? in : p->work;
    f.in_mode = j == 0 ? in_mode : 0;
    f.do_fwd = inverse ? 0 : 1;
    f.do_inv = inverse ? 1 : 0;
    f.kmul = nullptr;
    const bool last = j == na - 1;
    f.out = last ? out : p->work;
    f.out_mode = last ? out_mode : 0;
    PSX_TRY(launch_cplx<T>(p->axes[i], p->prec, f, p->stream));
  }
  return PSX_OK;
}

static psx_status plan_buf(psx_fft* p, void** b, size_t elems) {
  if (!*b) PSX_TRY(dalloc(b, elems * prec_bytes(p->prec)));
  return PSX_OK;
}

static int fft_red_blocks(const psx_fft* p) { return grid_1d(p->N); }

static psx_status plan_sum_buf(psx_fft* p) {
  if (!p->dsum) PSX_TRY(dalloc((void**)&p->dsum, (size_t)(fft_red_blocks(p) + 1) * sizeof(double)));
  return PSX_OK;
}

// x <- x - sum(x) / vol on the plan's stream.
template <typename T>
static psx_status shift_by_mean(psx_fft* p, T* x, double vol) {
  PSX_TRY(plan_sum_buf(p));
  const int g = fft_red_blocks(p);
  hipLaunchKernelGGL(k_partial_sum<T>, dim3(g), dim3(kThreads), 0, p->stream, (const T*)x, p->N, p->dsum);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, p->stream, (const double*)p->dsum, g,
                     p->dsum + g);
  hipLaunchKernelGGL(k_shift_mean<T>, dim3(g), dim3(kThreads), 0, p->stream, x, p->N,
                     (const double*)(p->dsum + g), vol);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// The spectral regions of PsMultiSpecFilter::build_specCells_transpose
// (pseffhamil/PsMultiSpecFilter.cpp:165-215) as a serial run builds them:
// the map is laid out in the transposed (y, x, z) order of the transpose
// plan's grid, while the spectrum it is applied to comes from the plain plan
// in (x, y, z) order, so cell n = (i ny + j) nz + k gets the region of the
// transposed position with the same flat index.  On a square grid with
// equal x and y cell sizes this only relabels the regions.
static psx_status ensure_regions(psx_fft* p, const int* cells) {
  if (p->region && cells[0] == p->region_cells[0] && cells[1] == p->region_cells[1] &&
      cells[2] == p->region_cells[2])
    return PSX_OK;
  const int nx = p->dims[0], ny = p->dims[1], nz = p->dims[2];
  const int dims[3] = {nx, ny, nz};
  int nsc[3];
  for (int d = 0; d < 3; ++d) {
    if (cells[d] < 1 || dims[d] % cells[d] != 0)
      return fail(PSX_ERR_INVALID, "specCellSizes must divide the grid (" + std::to_string(dims[d]) +
                                       " cells, spectral cell " + std::to_string(cells[d]) + ")");
    nsc[d] = dims[d] / cells[d];
  }
  std::vector<int> map((size_t)p->N);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const size_t nn = ((size_t)j * nx + i) * nz + k;
        map[nn] = (k / cells[2]) * (nsc[0] * nsc[1]) + (j / cells[1]) * nsc[0] + i / cells[0];
      }
  dfree((void*&)p->region);
  dfree((void*&)p->rmax);
  p->nregions = nsc[0] * nsc[1] * nsc[2];
  PSX_TRY(upload((void**)&p->region, map.data(), map.size() * sizeof(int)));
  PSX_TRY(dalloc((void**)&p->rmax, (size_t)p->nregions * sizeof(unsigned long long)));
  for (int d = 0; d < 3; ++d) p->region_cells[d] = cells[d];
  return PSX_OK;
}

// PsSpecFilterUpdater::subtractAverage then PsSimpleSpecFilter::applyFilter
// (cells == nullptr) or PsMultiSpecFilter::applyFilter, in place on w.
template <typename T>
static psx_status spec_filter_impl(psx_fft* p, T* w, double cutoff, double strength,
                                   const int* cells, double vwall) {
  if (cutoff <= 0.0) return PSX_OK;   // PsSimpleSpecFilter.cpp:82-86
  const long long N = p->N;
  PSX_TRY(plan_buf(p, &p->rbuf, N));
  PSX_TRY(plan_buf(p, &p->mbuf, N));
  const int* reg = nullptr;
  int nreg = 1;
  if (cells) {
    PSX_TRY(ensure_regions(p, cells));
    reg = p->region;
    nreg = p->nregions;
  } else if (!p->rmax || p->nregions < 1) {
    dfree((void*&)p->rmax);
    PSX_TRY(dalloc((void**)&p->rmax, sizeof(unsigned long long)));
    p->nregions = 1;
    p->region_cells[0] = p->region_cells[1] = p->region_cells[2] = 0;
  }
  PSX_TRY(shift_by_mean<T>(p, w, (double)N - vwall));
  PSX_TRY(fft_passes<T>(p, false, w, 1, p->rbuf, 2));   // forwardFFTAbs
  const int g = fft_red_blocks(p);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_1d(nreg)), dim3(kThreads), 0, p->stream, p->rmax,
                     (long long)nreg, 0ull);
  hipLaunchKernelGGL(k_region_max<T>, dim3(g), dim3(kThreads), 0, p->stream, (const T*)p->rbuf, reg, N,
                     p->rmax);
  hipLaunchKernelGGL(k_filter_mask<T>, dim3(g), dim3(kThreads), 0, p->stream, (const T*)p->rbuf, reg,
                     (const unsigned long long*)p->rmax, (T)cutoff, (T)strength, (T*)p->mbuf, N);
  PSX_HIP(hipGetLastError());
  PSX_TRY(fft_pair_impl<T>(p, w, p->mbuf, p->rbuf));
  const T scale = (T)(1.0 / (double)(T)N);
  hipLaunchKernelGGL(k_scale_store<T>, dim3(g), dim3(kThreads), 0, p->stream, (const T*)p->rbuf, scale, w, N);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// 4 * 3.1415 * lB / k^2 (0 where k^2 < 1e-9) in C order, k^2 as
// PsGridField::calck2 forms it (psbase/PsGridField.cpp:66-104);
// PsPoissonUpdater::build_k2 (pseffhamil/PsPoissonUpdater.cpp:171-203).
template <typename T>
static psx_status ensure_k2inv(psx_fft* p, double lB, const double* h) {
  const double key[4] = {lB, h[0], h[1], h[2]};
  if (p->k2valid && std::memcmp(key, p->k2key, sizeof(key)) == 0) return PSX_OK;
  const double twopi = 2.0 * std::acos(-1.0);
  const int* n = p->dims;
  double n2[3];
  for (int d = 0; d < 3; ++d) n2[d] = (double)n[d] / 2.0;
  std::vector<T> tab((size_t)p->N);
  size_t idx = 0;
  for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j)
      for (int k = 0; k < n[2]; ++k) {
        const int pos[3] = {i, j, k};
        double kv[3];
        for (int d = 2; d >= 0; --d) {
          const double nk = n2[d] - std::fabs((double)pos[d] - n2[d]);
          kv[d] = twopi * nk / (double)n[d];
          kv[d] = kv[d] / h[d];
        }
        const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        tab[idx++] = k2 < 0.000000001 ? (T)0 : (T)((4.0 * 3.1415 * lB) / k2);
      }
  dfree(p->kbuf);
  p->k2valid = false;
  PSX_TRY(upload(&p->kbuf, tab.data(), tab.size() * sizeof(T)));
  std::memcpy(p->k2key, key, sizeof(key));
  p->k2valid = true;
  return PSX_OK;
}

template <typename T>
static psx_status poisson_impl(psx_fft* p, T* charge, T* psi, double lB, const double* h) {
  PSX_TRY(ensure_k2inv<T>(p, lB, h));
  PSX_TRY(plan_buf(p, &p->rbuf, p->N));
  PSX_TRY(shift_by_mean<T>(p, charge, (double)p->N));
  PSX_TRY(fft_pair_impl<T>(p, charge, p->kbuf, p->rbuf));
  const T scale = (T)(1.0 / (double)(T)p->N);
  hipLaunchKernelGGL(k_poisson_psi<T>, dim3(fft_red_blocks(p)), dim3(kThreads), 0, p->stream,
                     (const T*)p->rbuf, psi, scale, p->N);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

#define PSX_FFT_DISPATCH(call_d, call_f) \
  ((p->prec == PSX_FP64) ? (call_d) : (call_f))

extern "C" psx_status psx_fft_forward_abs(psx_fft* p, const void* d_data, void* d_res) {
  if (!p || !d_data || !d_res) return fail(PSX_ERR_INVALID, "null argument");
  return PSX_FFT_DISPATCH(fft_passes<double>(p, false, d_data, 1, d_res, 2),
                          fft_passes<float>(p, false, d_data, 1, d_res, 2));
}

extern "C" psx_status psx_fft_forward_re(psx_fft* p, const void* d_data, void* d_res) {
  if (!p || !d_data || !d_res) return fail(PSX_ERR_INVALID, "null argument");
  return PSX_FFT_DISPATCH(fft_passes<double>(p, false, d_data, 1, d_res, 1),
                          fft_passes<float>(p, false, d_data, 1, d_res, 1));
}

extern "C" psx_status psx_fft_backward_re(psx_fft* p, const void* d_data, void* d_res) {
  if (!p || !d_data || !d_res) return fail(PSX_ERR_INVALID, "null argument");
  return PSX_FFT_DISPATCH(fft_passes<double>(p, true, d_data, 1, d_res, 1),
                          fft_passes<float>(p, true, d_data, 1, d_res, 1));
}

extern "C" psx_status psx_fft_scaled_pair_im(psx_fft* p, const void* d_data, const void* d_kdata,
                                             void* d_res) {
  if (!p || !d_data || !d_kdata || !d_res) return fail(PSX_ERR_INVALID, "null argument");
  return PSX_FFT_DISPATCH(fft_pair_impl<double>(p, d_data, d_kdata, d_res, 2),
                          fft_pair_impl<float>(p, d_data, d_kdata, d_res, 2));
}

template <typename T>
static psx_status convolve_impl(psx_fft* p, const void* d1, const void* d2, void* res) {
  PSX_TRY(plan_buf(p, &p->work2, 2 * p->N));
  PSX_TRY(fft_passes<T>(p, false, d2, 1, p->work2, 0));
  PSX_TRY(fft_passes<T>(p, false, d1, 1, p->work, 0));
  hipLaunchKernelGGL(k_spec_cmul<T>, dim3(fft_red_blocks(p)), dim3(kThreads), 0, p->stream, (T*)p->work,
                     (const T*)p->work2, p->N);
  PSX_HIP(hipGetLastError());
  return fft_passes<T>(p, true, p->work, 0, res, 1);
}

extern "C" psx_status psx_fft_convolve_re(psx_fft* p, const void* d_data1, const void* d_data2,
                                          void* d_res) {
  if (!p || !d_data1 || !d_data2 || !d_res) return fail(PSX_ERR_INVALID, "null argument");
  return PSX_FFT_DISPATCH(convolve_impl<double>(p, d_data1, d_data2, d_res),
                          convolve_impl<float>(p, d_data1, d_data2, d_res));
}

extern "C" psx_status psx_fft_spec_filter(psx_fft* p, void* d_w, double cutoff_factor,
                                          double filter_strength, const int* spec_cells,
                                          double wall_volume) {
  if (!p || !d_w) return fail(PSX_ERR_INVALID, "null argument");
  if (!(wall_volume >= 0.0 && wall_volume < (double)p->N))
    return fail(PSX_ERR_INVALID, "wall volume must lie in [0, N)");
  return PSX_FFT_DISPATCH(
      spec_filter_impl<double>(p, (double*)d_w, cutoff_factor, filter_strength, spec_cells, wall_volume),
      spec_filter_impl<float>(p, (float*)d_w, cutoff_factor, filter_strength, spec_cells, wall_volume));
}

extern "C" psx_status psx_fft_poisson_update(psx_fft* p, void* d_charge, void* d_psi, double bjerrum_len,
                                             double dx, double dy, double dz) {
  if (!p || !d_charge || !d_psi) return fail(PSX_ERR_INVALID, "null argument");
  if (!(dx > 0.0 && dy > 0.0 && dz > 0.0)) return fail(PSX_ERR_INVALID, "cell sizes must be > 0");
  const double h[3] = {dx, dy, dz};
  return PSX_FFT_DISPATCH(poisson_impl<double>(p, (double*)d_charge, (double*)d_psi, bjerrum_len, h),
                          poisson_impl<float>(p, (float*)d_charge, (float*)d_psi, bjerrum_len, h));
}

extern "C" psx_status psx_fft_synchronize(psx_fft* p) {
  if (!p) return fail(PSX_ERR_INVALID, "null plan");
  PSX_HIP(hipStreamSynchronize(p->stream));
  return PSX_OK;
}

extern "C" psx_status psx_probe_stream(size_t bytes, int reps, int mode, double* ms) {
  if (!ms || reps < 1 || (mode != 0 && mode != 1) || bytes < 16) return fail(PSX_ERR_INVALID, "bad argument");
  const long long n = (long long)(bytes / 16);
  void *a = nullptr, *b = nullptr, *part = nullptr;
  const int blocks = 4096;
  psx_status st = dalloc(&a, (size_t)n * 16);
  if (st == PSX_OK && mode == 1) st = dalloc(&b, (size_t)n * 16);
  if (st == PSX_OK) st = dalloc(&part, blocks * sizeof(double));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == PSX_OK && (hipMemset(a, 0, (size_t)n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                       hipEventCreate(&e1) != hipSuccess))
    st = fail(PSX_ERR_HIP, "probe setup failed");
  auto launch = [&]() {
    if (mode == 0)
      hipLaunchKernelGGL(k_probe_read, dim3(blocks), dim3(kThreads), 0, nullptr, (const double2*)a, n, (double*)part);
    else
      hipLaunchKernelGGL(k_probe_copy, dim3(blocks), dim3(kThreads), 0, nullptr, (const double2*)a, n, (double2*)b);
  };
  if (st == PSX_OK) {
    launch();   // warm
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, nullptr);
    float t = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess ||
        hipGetLastError() != hipSuccess)
      st = fail(PSX_ERR_HIP, "probe failed");
    *ms = (double)t / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(a); dfree(b); dfree(part);
  return st;
}

extern "C" psx_status psx_buffer_create(size_t bytes, void** d_out) {
  if (!d_out) return fail(PSX_ERR_INVALID, "null argument");
  return dalloc(d_out, bytes);
}

extern "C" psx_status psx_buffer_destroy(void* d) {
  dfree(d);
  return PSX_OK;
}

extern "C" psx_status psx_buffer_upload(void* d_dst, const void* h_src, size_t bytes) {
  if (!d_dst || !h_src) return fail(PSX_ERR_INVALID, "null argument");
  PSX_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return PSX_OK;
}

extern "C" psx_status psx_buffer_download(void* h_dst, const void* d_src, size_t bytes) {
  if (!h_dst || !d_src) return fail(PSX_ERR_INVALID, "null argument");
  PSX_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  return PSX_OK;
}

// ===========================================================================
// SCFT engine
// ===========================================================================
namespace {

// One block of the engine's block graph.
struct BlockInfo {
  int species = 0;                 // 0 = A, 1 = B
  double frac = 0.0, b = 1.0;
  int steps = 0;
  int poly = 0;                    // polymer of a blend
  int bclass = 0;                  // index of the spectral tables for this b
  std::vector<std::pair<int, int>> cnt[2];   // (block, its end 0 head / 1 tail) joined at head, tail
  int path[2] = {-1, -1};          // path of the propagator from the head (0) / tail (1)
  int off[2] = {0, 0};             // its slice 0 in that path
};

// A propagation path: block propagators in sequence, each starting at the
// final slice of the one before (a junction with a single partner), stored
// as one contiguous slice array.  A path starts at a free end (slice 0 = 1)
// or at a junction product of finals of paths of earlier rounds.
struct PathInfo {
  std::vector<int> step_block;     // block of every contour step
  int len = 0;
  int round = 0;
  int slot = 0;                    // stream slot within the round
  std::vector<std::pair<int, int>> init;   // (path, slice) factors of slice 0; empty = free end
  std::vector<int> launch_axis;    // axis index of every launch (see sweep_launches)
  std::vector<int> launch_nops;    // 1 plain, 2 fused
};

struct PathBufs {
  void* P = nullptr;               // (len + 1) slices of T
  void* tmp = nullptr;             // intermediate of a step
};

struct SweepBufs {
  bool ready = false;
  std::vector<PathBufs> path;
  void* wf[2] = {nullptr, nullptr};          // exp(-ds w / 2) of species A, B (T)
  std::vector<std::array<void*, 3>> emul;    // [b class][axis idx] spectral factors (T)
  std::vector<void*> epl;          // [b class] plane launches: z factors, then y factors (T)
  std::vector<PairIO*> rtab;       // per round: device table [launch][path slot]
  bool wfac_valid = false;
  // contour checkpointing (linear chains; see CkptSeg)
  void* ck = nullptr;              // backward slices 0, K, 2K, ... and the final one
  void* seg = nullptr;             // K + 1 recomputed backward slices
  void* fch[2] = {nullptr, nullptr};   // forward chunks, K + 1 slices each
  void* scr = nullptr;             // 3 rotating slices of the checkpointing sweep
  PairIO* btab = nullptr;          // backward sweep writing the checkpoints
  std::vector<PairIO*> gtab, ftab; // per chunk: backward recompute, forward advance
};

// One chunk of a checkpointed contour pass: backward slices u0..u1 are
// recomputed from checkpoint u0 = jK, and the forward propagator advances
// over t0..t1 = Ns-u1 .. Ns-u0, the points that pair with them.  Chunks are
// kept in processing order (t ascending).
struct CkptSeg {
  int u0, u1, t0, t1;
  std::vector<int> g_axis, f_axis;   // launch axes of the two range sweeps
};

struct GraphKey {
  int prec, kind, depth, f_order;
  double lam0, lam1, beta, start;
  bool operator<(const GraphKey& o) const {
    return std::memcmp(this, &o, sizeof(GraphKey)) < 0;
  }
};

}  // namespace

struct psx_scft {
  int dev = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // The forward and backward sweeps are independent: they run as two
  // branches (second stream, forked and joined with events) so that their
  // kernels drift out of phase and one sweep's loads overlap the other's
  // arithmetic.  Default priority like the main stream.
  std::vector<hipStream_t> pstream;     // stream slot k >= 1 of a round (slot 0 = stream)
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;      // per slot
  bool split_sweeps = true;
  // Quadrature under the sweeps (linear chains, PSX_OVERLAP_QUAD=1): point t
  // of the contour needs forward slice t and backward slice Ns-t, both
  // written once the sweeps pass step max(t, Ns-t).  Chunks of points run on
  // qstream as their slices appear (raw sums into phiA/phiB); the last chunk
  // finishes after the join.  Off by default: at 128^3 the sweeps already
  // fill the wave slots and the concurrent chunks slowed the iteration
  // (FP64 173.3 -> 171.0 it/s, FP32 292.1 -> 286.2).
  bool overlap_quad = false;
  hipStream_t qstream = nullptr;
  hipEvent_t ev_qdone = nullptr;
  struct QChunk {
    std::vector<std::pair<int, int>> ranges;   // forward point ranges [lo, hi]
    int launch = 0;                            // after this launch of both paths its slices exist
    hipEvent_t ev[2] = {nullptr, nullptr};
  };
  std::vector<QChunk> qchunks;                 // empty: quadrature after the sweeps
  bool quad_partial = false;                   // enqueueing an overlapped iteration
  int dims[3];
  long long N;
  double cell[3];
  double fA, fB, ds, chiN;              // fA, fB: volfrac * lengthfrac summed per species
  int Ns;                               // contour steps summed over the blocks
  // Polymers of a blend: volfrac, lengthRatio, their blocks first .. first + n - 1
  // (one polymer of volfrac = lengthRatio = 1 for a single block copolymer)
  struct PolyInfo { double vf = 1.0, lr = 1.0; int first = 0, n = 0; };
  std::vector<PolyInfo> polys;
  std::vector<BlockInfo> blocks;
  std::vector<double> bval;             // distinct bSegRatio values (b classes)
  std::vector<PathInfo> paths;
  std::vector<std::vector<int>> rounds; // path ids per round, slot order
  bool diblock = false;                 // two blocks A-B joined head to tail
  int ckpt_K = 0;                       // contour checkpoint interval (0: every slice stored)
  int nck = 0;                          // checkpoint slices
  std::vector<CkptSeg> segs;
  std::vector<int> btab_axis;
  // z and y passes as plane launches (k_plane; 3D grids with n_y = n_z of a
  // plane length); on by default where available
  bool plane = false;
  bool in_batch = false;                // owned by a psx_batch (its tables point at our slices)
  int quad_kernel = 0;                  // 0 auto, 1 always k_quad_update (psx_scft_set_quad_kernel)
  int quad_mode = 0;                    // contour weights: 0 reference Simpson, 1 trapezoid
  // reverse mode (psx_scft_backward): two adjoint slices, the step result,
  // dJ/dw and the dQ/dw sums per species, and the line-pass table
  double* adj = nullptr;                // 7 N
  PairIO* adj_tab = nullptr;
  size_t adj_tab_rows = 0;
  psx_fft* fplan = nullptr;             // FP64 plan of the spectral filters (first use)
  long long graph_gen = 0;              // bumped whenever captured graphs go stale
  // steepest-descent noise (PsPolymerUpdater::addRandUpdateConjgFields)
  psx_rng* noise_rng = nullptr;         // caller's stream; null: no noise drawn
  double noise_strength = 0.0;
  double* noise = nullptr;              // device, 2N: the current iteration's noise
  double* noise_host[2] = {nullptr, nullptr};   // pinned staging, alternating
  hipEvent_t noise_ev[2] = {nullptr, nullptr};
  int noise_k = 0;
  int prec;
  std::vector<Axis> axes;
  int variant[2][3] = {{-1, -1, -1}, {-1, -1, -1}};   // line-kernel choice per precision, pass
  SweepBufs sw[2];
  double *wA = nullptr, *wB = nullptr, *praw = nullptr, *phiA = nullptr, *phiB = nullptr;
  double* cw = nullptr;          // quadrature weights (Ns + 2)
  double* part = nullptr;        // reduction partials
  int nb = 1;                    // blocks of the 1D kernels
  int part_cap = 0;              // doubles in part
  DevState* st = nullptr;
  double* hist = nullptr;
  int hist_cap = 0;
  double* ringx = nullptr;
  double* ringf = nullptr;
  int ring_m1 = 0;
  bool swept = false;            // slices valid for the current fields
  bool dens_valid = false;       // phi valid for the current slices
  std::map<GraphKey, hipGraphExec_t> graphs;
  // box stress: FP64 complex work grid, per-axis k^2 and kfac tables
  void* zbuf = nullptr;
  double* kd2[3] = {nullptr, nullptr, nullptr};
  std::vector<std::array<double*, 3>> kf1;   // per b class: exp(-ds b^2 k_d^2)
  double* sacc = nullptr;
  std::vector<Axis> caxes;     // axes of the complex stress transforms
  // cached host-visible scratch for the host API
  void* hstage = nullptr;
};

namespace {

void drop_graphs(psx_scft* e) {
  for (auto& kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
  e->graphs.clear();
  ++e->graph_gen;   // a batch holding this engine's buffers in its graphs re-captures
}

// exp(-ds b^2 k_d^2) / n_d for every axis and block; the product over axes is
// the reference kfac = exp(-ds b^2 k^2) (pspolymer/PsFlexPseudoSpec.cpp:188-206)
// because k^2 = kx^2 + ky^2 + kz^2 (psbase/PsGridField.cpp:66-104).
std::vector<double> axis_emul(const Axis& a, double cell, double ds, double b) {
  std::vector<double> e(a.n);
  const double twopi = 2.0 * std::acos(-1.0);
  const double n2 = (double)a.n / 2.0;
  for (int k = 0; k < a.n; ++k) {
    const double nk = n2 - std::fabs((double)k - n2);
    double kv = twopi * nk / (double)a.n;
    kv = kv / cell;
    e[k] = std::exp(-1.0 * ds * kv * kv * b * b) / (double)a.n;
  }
  return e;
}

// The spectral rows as the kernels read them: the n factors of every row,
// then one (a, b) row per factor row for the fused lane-pair round trips
// (psx_fft_reg.hpp, PairAB): entry k < n/2 holds a = K(k) + K(k + n/2) and
// b = K(k) - K(k + n/2); zeros for odd n.
std::vector<double> spectral_rows(const std::vector<double>& K, int n) {
  const size_t rows = K.size() / (size_t)n;
  std::vector<double> out(K);
  out.resize(2 * K.size(), 0.0);
  if (n % 2 != 0) return out;
  for (size_t r = 0; r < rows; ++r) {
    const double* k = K.data() + r * n;
    double* ab = out.data() + K.size() + r * n;
    for (int i = 0; i < n / 2; ++i) {
      ab[2 * i] = k[i] + k[i + n / 2];
      ab[2 * i + 1] = k[i] - k[i + n / 2];
    }
  }
  return out;
}

// Launch list of one sweep.  Contour step s applies the active axes in the
// order 0..na-1 when s is even and na-1..0 when s is odd (the axis operators
// commute), so the last pass of step s and the first of step s+1 run along
// the same axis and are fused into one launch (two round trips, the step's
// result stored in between).  Both sweeps share the list.
//
// Plane mode (3D, see k_plane): every run of consecutive z and y passes is
// one launch with axis code kPlaneAxis.  Such a run is (0 first, 1) at the
// start of a sweep, (1, 0 last) at its end, and (1, 0 last | 0 first, 1)
// across each step boundary in between, so the launches alternate between
// planes and fused x passes: one launch per contour step.
struct SweepOp { int axis, step; bool first, last; };
constexpr int kPlaneAxis = 3;

std::vector<std::pair<SweepOp, int>> sweep_launches(int na, int Ns, bool plane = false) {
  std::vector<SweepOp> ops;
  for (int st = 0; st < Ns; ++st)
    for (int p = 0; p < na; ++p) {
      const int a = (st % 2 == 0) ? p : na - 1 - p;
      ops.push_back(SweepOp{a, st, p == 0, p == na - 1});
    }
  // (first op, number of ops: 1 or 2; plane launches 2 or 4).  A plane
  // launch's op has first = its first op starts a step, last = (2 ops) its
  // second op ends the step.
  std::vector<std::pair<SweepOp, int>> out;
  for (size_t i = 0; i < ops.size();) {
    if (plane && na == 3 && ops[i].axis <= 1) {
      size_t j = i;
      while (j < ops.size() && ops[j].axis <= 1) ++j;
      SweepOp o = ops[i];
      o.axis = kPlaneAxis;
      o.last = (j - i == 2) && ops[j - 1].last;
      out.push_back({o, (int)(j - i)});
      i = j;
    } else if (ops[i].last && i + 1 < ops.size() && ops[i + 1].axis == ops[i].axis) {
      out.push_back({ops[i], 2});
      i += 2;
    } else {
      out.push_back({ops[i], 1});
      i += 1;
    }
  }
  return out;
}

// Field descriptors of the launches (see sweep_launches) that advance path p
// over its contour steps s0 .. s0+nsteps-1, with slice(k) giving the device
// address of the k-th slice of that range (k = 0 .. nsteps).  The w-factor
// and spectral factors of each step are those of the step's block.
template <typename SliceFn>
std::vector<PairIO> range_fio(const psx_scft* e, int prec, int p, int s0, int nsteps, SliceFn slice) {
  const SweepBufs& s = e->sw[prec];
  const PathInfo& pi = e->paths[p];
  const int na = (int)e->axes.size();
  const auto launches = sweep_launches(na, nsteps, e->plane);
  std::vector<PairIO> tab(launches.size());
  void* tmp = s.path[p].tmp;
  auto blk = [&](int st) -> const BlockInfo& { return e->blocks[pi.step_block[s0 + st]]; };
  auto W = [&](int st) { return (const void*)s.wf[blk(st).species]; };
  auto E = [&](int a, int st) { return (const void*)s.emul[blk(st).bclass][a]; };
  auto EP = [&](int st) { return (const void*)s.epl[blk(st).bclass]; };
  for (size_t j = 0; j < launches.size(); ++j) {
    const SweepOp& o = launches[j].first;
    PairIO& f = tab[j];
    std::memset(&f, 0, sizeof(f));
    f.in = o.first ? slice(o.step) : tmp;
    f.pre = o.first ? W(o.step) : nullptr;
    if (o.axis == kPlaneAxis) {
      f.emul = EP(o.step);
      if (launches[j].second == 2) {
        f.post = o.last ? W(o.step) : nullptr;
        f.out = o.last ? slice(o.step + 1) : tmp;
      } else {   // (1, 0 last | 0 first, 1): the step's result is stored between
        f.post = W(o.step);
        f.mid = slice(o.step + 1);
        f.pre2 = W(o.step + 1);
        f.emul2 = EP(o.step + 1);
        f.out = tmp;
      }
      continue;
    }
    f.emul = E(o.axis, o.step);
    if (launches[j].second == 1) {
      f.post = o.last ? W(o.step) : nullptr;
      f.out = o.last ? slice(o.step + 1) : tmp;
    } else {
      const int s2 = o.step + 1;
      const bool last2 = na == 1;   // the second op is also the last of its step
      f.post = W(o.step);
      f.mid = slice(s2);
      f.pre2 = W(s2);
      f.emul2 = E(o.axis, s2);
      f.post2 = last2 ? W(s2) : nullptr;
      f.out = last2 ? slice(s2 + 1) : tmp;
    }
  }
  return tab;
}

// All launches of path p into its own slice array.
std::vector<PairIO> path_fio(const psx_scft* e, int prec, int p) {
  const char* base = (const char*)e->sw[prec].path[p].P;
  const size_t sb = (size_t)e->N * prec_bytes(prec);
  return range_fio(e, prec, p, 0, e->paths[p].len,
                   [&](int k) { return (void*)(base + (size_t)k * sb); });
}

// Paths of round r have identical launch lists (equal lengths).
bool round_uniform(const psx_scft* e, int r) {
  const std::vector<int>& rp = e->rounds[r];
  for (int p : rp)
    if (e->paths[p].len != e->paths[rp[0]].len) return false;
  return true;
}

// Device table of round r: entry [j * np + slot] is launch j of the path in
// that slot (rows past a shorter path's launch count stay zero, never used).
psx_status upload_round_table(psx_scft* e, int prec, int r) {
  const std::vector<int>& rp = e->rounds[r];
  const size_t np = rp.size();
  size_t rows = 0;
  std::vector<std::vector<PairIO>> per;
  for (int p : rp) {
    per.push_back(path_fio(e, prec, p));
    rows = std::max(rows, per.back().size());
  }
  std::vector<PairIO> tab(rows * np);
  std::memset(tab.data(), 0, tab.size() * sizeof(PairIO));
  for (size_t k = 0; k < np; ++k)
    for (size_t j = 0; j < per[k].size(); ++j) tab[j * np + k] = per[k][j];
  return upload((void**)&e->sw[prec].rtab[r], tab.data(), tab.size() * sizeof(PairIO));
}

// Checkpointed mode (linear chains, path 0 forward from block 0's free head,
// path 1 backward): the backward sweep stores slices u = 0, K, 2K, ... and
// its final slice; each chunk recomputes K backward slices from a checkpoint
// and advances the forward propagator over the points that pair with them.
void ckpt_plan(psx_scft* e) {
  const int Ns = e->paths[1].len, K = e->ckpt_K;
  e->nck = Ns / K + 1 + (Ns % K ? 1 : 0);
  const int na = (int)e->axes.size();
  e->btab_axis.clear();
  for (const auto& l : sweep_launches(na, Ns, e->plane)) e->btab_axis.push_back(l.first.axis);
  e->segs.clear();
  const int nseg = (Ns + K - 1) / K;
  for (int j = nseg - 1; j >= 0; --j) {
    CkptSeg c;
    c.u0 = j * K;
    c.u1 = std::min((j + 1) * K, Ns);
    c.t0 = Ns - c.u1;
    c.t1 = Ns - c.u0;
    for (const auto& l : sweep_launches(na, c.u1 - c.u0, e->plane)) {
      c.g_axis.push_back(l.first.axis);
      c.f_axis.push_back(l.first.axis);
    }
    e->segs.push_back(c);
  }
}

psx_status ensure_ckpt(psx_scft* e, int prec) {
  SweepBufs& s = e->sw[prec];
  const size_t sb = (size_t)e->N * prec_bytes(prec);
  const int K = e->ckpt_K, Ns = e->paths[1].len;
  PSX_TRY(dalloc(&s.ck, (size_t)e->nck * sb));
  PSX_TRY(dalloc(&s.seg, (size_t)(K + 1) * sb));
  PSX_TRY(dalloc(&s.fch[0], (size_t)(K + 1) * sb));
  PSX_TRY(dalloc(&s.fch[1], (size_t)(K + 1) * sb));
  PSX_TRY(dalloc(&s.scr, 3 * sb));
  char* ck = (char*)s.ck;
  char* scr = (char*)s.scr;
  auto bslice = [&](int k) -> void* {
    if (k % K == 0) return ck + (size_t)(k / K) * sb;
    if (k == Ns) return ck + (size_t)(e->nck - 1) * sb;
    return scr + (size_t)(k % 3) * sb;
  };
  std::vector<PairIO> bt = range_fio(e, prec, 1, 0, Ns, bslice);
  PSX_TRY(upload((void**)&s.btab, bt.data(), bt.size() * sizeof(PairIO)));
  s.gtab.assign(e->segs.size(), nullptr);
  s.ftab.assign(e->segs.size(), nullptr);
  for (size_t ci = 0; ci < e->segs.size(); ++ci) {
    const CkptSeg& c = e->segs[ci];
    char* seg = (char*)s.seg;
    char* fch = (char*)s.fch[ci % 2];
    std::vector<PairIO> gt = range_fio(e, prec, 1, c.u0, c.u1 - c.u0,
                                       [&](int k) { return (void*)(seg + (size_t)k * sb); });
    std::vector<PairIO> ft = range_fio(e, prec, 0, c.t0, c.t1 - c.t0,
                                       [&](int k) { return (void*)(fch + (size_t)k * sb); });
    PSX_TRY(upload((void**)&s.gtab[ci], gt.data(), gt.size() * sizeof(PairIO)));
    PSX_TRY(upload((void**)&s.ftab[ci], ft.data(), ft.size() * sizeof(PairIO)));
  }
  // the backward sweep starts at a free end: checkpoint 0 holds q = 1
  const int g = grid_1d(e->N);
  if (prec == PSX_FP64)
    hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kThreads), 0, e->stream, (double*)s.ck, e->N, 1.0);
  else
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kThreads), 0, e->stream, (float*)s.ck, e->N, 1.0f);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// Spectral factors of a plane launch for b class c: the z factors followed
// by the y factors (both include 1/n), then their (a, b) rows.
std::vector<double> plane_factors(const psx_scft* e, size_t c) {
  std::vector<double> em = axis_emul(e->axes[0], e->cell[e->axes[0].axis], e->ds, e->bval[c]);
  const std::vector<double> ey = axis_emul(e->axes[1], e->cell[e->axes[1].axis], e->ds, e->bval[c]);
  em.insert(em.end(), ey.begin(), ey.end());
  return spectral_rows(em, e->axes[0].n);
}

psx_status ensure_sweep(psx_scft* e, int prec) {
  SweepBufs& s = e->sw[prec];
  if (s.ready) return PSX_OK;
  const size_t tb = prec_bytes(prec);
  s.path.assign(e->paths.size(), PathBufs());
  for (size_t p = 0; p < e->paths.size(); ++p) {
    if (e->ckpt_K == 0) PSX_TRY(dalloc(&s.path[p].P, (size_t)(e->paths[p].len + 1) * e->N * tb));
    PSX_TRY(dalloc(&s.path[p].tmp, e->N * tb));
  }
  PSX_TRY(dalloc(&s.wf[0], e->N * tb));
  PSX_TRY(dalloc(&s.wf[1], e->N * tb));
  s.emul.assign(e->bval.size(), std::array<void*, 3>{nullptr, nullptr, nullptr});
  for (size_t i = 0; i < e->axes.size(); ++i) {
    Axis& a = e->axes[i];
    PSX_TRY(axis_tw(a, prec));
    for (size_t c = 0; c < e->bval.size(); ++c) {
      const double h = a.n == 1 ? 1.0 : e->cell[a.axis];
      std::vector<double> em = spectral_rows(axis_emul(a, h, e->ds, e->bval[c]), a.n);
      if (prec == PSX_FP64) PSX_TRY(upload(&s.emul[c][i], em.data(), em.size() * 8));
      else {
        std::vector<float> f = cast_vec<float>(em);
        PSX_TRY(upload(&s.emul[c][i], f.data(), f.size() * 4));
      }
    }
  }
  if (e->plane) {
    s.epl.assign(e->bval.size(), nullptr);
    for (size_t c = 0; c < e->bval.size(); ++c) {
      const std::vector<double> em = plane_factors(e, c);
      if (prec == PSX_FP64) PSX_TRY(upload(&s.epl[c], em.data(), em.size() * 8));
      else {
        std::vector<float> f = cast_vec<float>(em);
        PSX_TRY(upload(&s.epl[c], f.data(), f.size() * 4));
      }
    }
    if (prec == PSX_FP64) PSX_TRY(launch_plane<double>(e->axes[0], prec, nullptr, 0, 0, e->stream, true));
    else PSX_TRY(launch_plane<float>(e->axes[0], prec, nullptr, 0, 0, e->stream, true));
  }
  const int g = grid_1d(e->N);
  if (e->ckpt_K > 0) {
    PSX_TRY(ensure_ckpt(e, prec));
    s.ready = true;
    s.wfac_valid = false;
    return PSX_OK;
  }
  // field tables of every launch of every path (static: graphs capture them)
  s.rtab.assign(e->rounds.size(), nullptr);
  for (size_t r = 0; r < e->rounds.size(); ++r) PSX_TRY(upload_round_table(e, prec, (int)r));
  // free ends: q = 1 (pspolymer/PsBlock.cpp:146-147)
  for (size_t p = 0; p < e->paths.size(); ++p) {
    if (!e->paths[p].init.empty()) continue;
    if (prec == PSX_FP64)
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kThreads), 0, e->stream, (double*)s.path[p].P, e->N, 1.0);
    else
      hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kThreads), 0, e->stream, (float*)s.path[p].P, e->N, 1.0f);
  }
  PSX_HIP(hipGetLastError());
  s.ready = true;
  s.wfac_valid = false;
  return PSX_OK;
}

void free_sweep(SweepBufs& s) {
  for (auto& pb : s.path) { dfree(pb.P); dfree(pb.tmp); }
  s.path.clear();
  dfree(s.ck); dfree(s.seg); dfree(s.fch[0]); dfree(s.fch[1]); dfree(s.scr);
  dfree((void*&)s.btab);
  for (auto& t : s.gtab) dfree((void*&)t);
  for (auto& t : s.ftab) dfree((void*&)t);
  s.gtab.clear();
  s.ftab.clear();
  dfree(s.wf[0]); dfree(s.wf[1]);
  for (auto& t : s.rtab) dfree((void*&)t);
  s.rtab.clear();
  for (auto& b : s.emul) for (auto& p : b) dfree(p);
  s.emul.clear();
  for (auto& p : s.epl) dfree(p);
  s.epl.clear();
  s.ready = false;
}

psx_status compute_wfac(psx_scft* e) {
  SweepBufs& s = e->sw[e->prec];
  const int g = grid_1d(e->N);
  const double h = 0.5 * e->ds;
  if (e->prec == PSX_FP64) {
    hipLaunchKernelGGL(k_wfac<double>, dim3(g), dim3(kThreads), 0, e->stream, e->wA, e->N, h, (double*)s.wf[0]);
    hipLaunchKernelGGL(k_wfac<double>, dim3(g), dim3(kThreads), 0, e->stream, e->wB, e->N, h, (double*)s.wf[1]);
  } else {
    hipLaunchKernelGGL(k_wfac<float>, dim3(g), dim3(kThreads), 0, e->stream, e->wA, e->N, h, (float*)s.wf[0]);
    hipLaunchKernelGGL(k_wfac<float>, dim3(g), dim3(kThreads), 0, e->stream, e->wB, e->N, h, (float*)s.wf[1]);
  }
  PSX_HIP(hipGetLastError());
  s.wfac_valid = true;
  return PSX_OK;
}

// Device address of slice t of path p in precision prec.
const void* slice_ptr(const psx_scft* e, int prec, int p, int t) {
  return (const char*)e->sw[prec].path[p].P + (size_t)t * e->N * prec_bytes(prec);
}

// Junction products for the paths of round r that start at a joined end.
template <typename T>
psx_status launch_junctions(psx_scft* e, int r) {
  for (int p : e->rounds[r]) {
    const PathInfo& pi = e->paths[p];
    if (pi.init.empty()) continue;
    JunctionArgs ja;
    std::memset(&ja, 0, sizeof(ja));
    ja.out = (void*)slice_ptr(e, e->prec, p, 0);
    ja.nin = (int)pi.init.size();
    ja.N = e->N;
    for (int k = 0; k < ja.nin; ++k) ja.in[k] = slice_ptr(e, e->prec, pi.init[k].first, pi.init[k].second);
    hipLaunchKernelGGL(k_junction<T>, dim3(grid_1d(e->N)), dim3(kThreads), 0, e->stream, ja);
  }
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// One launch of a sweep list: a line pass along axis a with its variant, or
// a plane launch (a == kPlaneAxis) over the engine's x-planes.
template <typename T>
psx_status launch_sweep_op(const psx_scft* e, int a, int prec, const PairIO* fio, int nfields,
                           hipStream_t s, const int* variant) {
  if (a == kPlaneAxis) return launch_plane<T>(e->axes[0], prec, fio, nfields, e->dims[0], s);
  return launch_pair<T>(e->axes[a], prec, fio, nfields, s, variant[a]);
}

// All propagators, round by round.  Within a round every path advances on
// its own stream slot (forked from and joined back into the engine stream;
// inside a capture these become parallel graph branches), or, with
// split_sweeps off and equal path lengths, all paths share each launch.
// only_axis >= 0 keeps the launches along that axis only (timing).
template <typename T>
psx_status launch_quad_points(psx_scft* e, int mode, const psx_mixing* mix,
                              const std::vector<std::pair<int, int>>& ranges, int acc_in,
                              hipStream_t stream);

template <typename T>
psx_status sweep_range(psx_scft* e, int only_axis = -1, bool quad_chunks = false) {
  SweepBufs& s = e->sw[e->prec];
  for (size_t r = 0; r < e->rounds.size(); ++r) {
    const std::vector<int>& rp = e->rounds[r];
    const int np = (int)rp.size();
    if (only_axis < 0) PSX_TRY(launch_junctions<T>(e, (int)r));
    const PairIO* tab = s.rtab[r];
    if (!e->split_sweeps && round_uniform(e, (int)r)) {
      const PathInfo& p0 = e->paths[rp[0]];
      for (size_t j = 0; j < p0.launch_axis.size(); ++j) {
        const int a = p0.launch_axis[j];
        if (only_axis >= 0 && a != only_axis) continue;
        PSX_TRY(launch_sweep_op<T>(e, a, e->prec, tab + j * np, np, e->stream, e->variant[e->prec]));
      }
      continue;
    }
    const bool fork = e->split_sweeps && np > 1;
    if (fork) {
      PSX_HIP(hipEventRecord(e->ev_fork, e->stream));
      for (int k = 1; k < np; ++k) PSX_HIP(hipStreamWaitEvent(e->pstream[k], e->ev_fork, 0));
    }
    // launches interleaved across the slots, so that every stream has work
    // queued early when the sweep is enqueued eagerly
    size_t rows = 0;
    for (int p : rp) rows = std::max(rows, e->paths[p].launch_axis.size());
    for (size_t j = 0; j < rows; ++j)
      for (int k = 0; k < np; ++k) {
        const PathInfo& pi = e->paths[rp[k]];
        if (j >= pi.launch_axis.size()) continue;
        const int a = pi.launch_axis[j];
        if (only_axis >= 0 && a != only_axis) continue;
        hipStream_t st = (fork && k > 0) ? e->pstream[k] : e->stream;
        PSX_TRY(launch_sweep_op<T>(e, a, e->prec, tab + j * np + k, 1, st, e->variant[e->prec]));
        if (quad_chunks)
          for (size_t c = 0; c + 1 < e->qchunks.size(); ++c)
            if (e->qchunks[c].launch == (int)j) PSX_HIP(hipEventRecord(e->qchunks[c].ev[k], st));
      }
    // contour chunks whose slices now exist: raw sums on the quadrature stream
    if (quad_chunks && fork)
      for (size_t j = 0; j < rows; ++j)
        for (size_t c = 0; c + 1 < e->qchunks.size(); ++c) {
          if (e->qchunks[c].launch != (int)j) continue;
          PSX_HIP(hipStreamWaitEvent(e->qstream, e->qchunks[c].ev[0], 0));
          PSX_HIP(hipStreamWaitEvent(e->qstream, e->qchunks[c].ev[1], 0));
          PSX_TRY(launch_quad_points<T>(e, QM_ACC, nullptr, e->qchunks[c].ranges, c > 0 ? 1 : 0, e->qstream));
        }
    if (fork) {
      for (int k = 1; k < np; ++k) {
        PSX_HIP(hipEventRecord(e->ev_join[k], e->pstream[k]));
        PSX_HIP(hipStreamWaitEvent(e->stream, e->ev_join[k], 0));
      }
      if (quad_chunks) {
        PSX_HIP(hipEventRecord(e->ev_qdone, e->qstream));
        PSX_HIP(hipStreamWaitEvent(e->stream, e->ev_qdone, 0));
      }
    }
  }
  return PSX_OK;
}

// Checkpointed mode: the backward path only, writing its checkpoints.
template <typename T>
psx_status sweep_ckpt(psx_scft* e) {
  const PairIO* tab = e->sw[e->prec].btab;
  for (size_t j = 0; j < e->btab_axis.size(); ++j) {
    const int a = e->btab_axis[j];
    PSX_TRY(launch_sweep_op<T>(e, a, e->prec, tab + j, 1, e->stream, e->variant[e->prec]));
  }
  return PSX_OK;
}

psx_status sweep(psx_scft* e) {
  if (e->ckpt_K > 0) return e->prec == PSX_FP64 ? sweep_ckpt<double>(e) : sweep_ckpt<float>(e);
  if (e->prec == PSX_FP64) return sweep_range<double>(e);
  return sweep_range<float>(e);
}

// Checkpointed mode, chunk ci: recompute its backward slices (second stream)
// while the forward propagator advances over the paired points.
template <typename T>
psx_status ckpt_chunk_sweeps(psx_scft* e, size_t ci) {
  SweepBufs& s = e->sw[e->prec];
  const CkptSeg& c = e->segs[ci];
  const size_t sb = (size_t)e->N * sizeof(T);
  const int g = grid_1d(e->N);
  hipStream_t s2 = e->pstream.size() > 1 ? e->pstream[1] : e->stream;
  if (s2 != e->stream) {
    PSX_HIP(hipEventRecord(e->ev_fork, e->stream));
    PSX_HIP(hipStreamWaitEvent(s2, e->ev_fork, 0));
  }
  const int j = c.u0 / e->ckpt_K;
  hipLaunchKernelGGL(k_copy<T>, dim3(g), dim3(kThreads), 0, s2,
                     (const T*)((const char*)s.ck + (size_t)j * sb), e->N, (T*)s.seg);
  for (size_t k = 0; k < c.g_axis.size(); ++k) {
    const int a = c.g_axis[k];
    PSX_TRY(launch_sweep_op<T>(e, a, e->prec, s.gtab[ci] + k, 1, s2, e->variant[e->prec]));
  }
  T* cur = (T*)s.fch[ci % 2];
  if (ci == 0) {
    hipLaunchKernelGGL(k_fill<T>, dim3(g), dim3(kThreads), 0, e->stream, cur, e->N, (T)1);
  } else {
    const CkptSeg& pc = e->segs[ci - 1];
    const T* prev = (const T*)((const char*)s.fch[(ci - 1) % 2] + (size_t)(pc.t1 - pc.t0) * sb);
    hipLaunchKernelGGL(k_copy<T>, dim3(g), dim3(kThreads), 0, e->stream, prev, e->N, cur);
  }
  for (size_t k = 0; k < c.f_axis.size(); ++k) {
    const int a = c.f_axis[k];
    PSX_TRY(launch_sweep_op<T>(e, a, e->prec, s.ftab[ci] + k, 1, e->stream, e->variant[e->prec]));
  }
  if (s2 != e->stream) {
    PSX_HIP(hipEventRecord(e->ev_join[1], s2));
    PSX_HIP(hipStreamWaitEvent(e->stream, e->ev_join[1], 0));
  }
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

psx_status ckpt_chunk(psx_scft* e, size_t ci) {
  return e->prec == PSX_FP64 ? ckpt_chunk_sweeps<double>(e, ci) : ckpt_chunk_sweeps<float>(e, ci);
}

// Q = mean(qHeadInitial qHeadFinal) of block 0 (pspolymer/PsBlock.cpp:205-237);
// a free head has qHeadInitial = 1.
// Q of every polymer: mean(qHeadInitial qHeadFinal) of its first block
// (pspolymer/PsBlock.cpp:205-237).
psx_status launch_Q(psx_scft* e) {
  for (size_t p = 0; p < e->polys.size(); ++p) {
    const BlockInfo& b0 = e->blocks[e->polys[p].first];
    const void* fin;
    const void* ini = nullptr;
    if (e->ckpt_K > 0) {   // block 0's head is the free start of the chain
      fin = (const char*)e->sw[e->prec].ck + (size_t)(e->nck - 1) * e->N * prec_bytes(e->prec);
    } else {
      fin = slice_ptr(e, e->prec, b0.path[1], b0.off[1] + b0.steps);
      if (!b0.cnt[0].empty()) ini = slice_ptr(e, e->prec, b0.path[0], b0.off[0]);
    }
    const int g = e->nb;
    if (e->prec == PSX_FP64) {
      if (ini) hipLaunchKernelGGL(k_partial_dot<double>, dim3(g), dim3(kThreads), 0, e->stream, (const double*)ini, (const double*)fin, e->N, e->part);
      else hipLaunchKernelGGL(k_partial_sum<double>, dim3(g), dim3(kThreads), 0, e->stream, (const double*)fin, e->N, e->part);
    } else {
      if (ini) hipLaunchKernelGGL(k_partial_dot<float>, dim3(g), dim3(kThreads), 0, e->stream, (const float*)ini, (const float*)fin, e->N, e->part);
      else hipLaunchKernelGGL(k_partial_sum<float>, dim3(g), dim3(kThreads), 0, e->stream, (const float*)fin, e->N, e->part);
    }
    hipLaunchKernelGGL(k_finish_Q, dim3(1), dim3(kThreads), 0, e->stream, (const double*)e->part, g,
                       1.0 / (double)e->N, e->st, (int)p);
    PSX_HIP(hipGetLastError());
  }
  return PSX_OK;
}

QuadArgs quad_args(psx_scft* e, int mode, const psx_mixing* mix) {
  SweepBufs& s = e->sw[e->prec];
  QuadArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nblk = e->ckpt_K > 0 ? 0 : (int)e->blocks.size();   // checkpointed: per chunk
  int woff = 0;
  for (int b = 0; b < a.nblk; ++b) {
    const BlockInfo& bi = e->blocks[b];
    a.blk[b].H = slice_ptr(e, e->prec, bi.path[0], bi.off[0]);
    a.blk[b].T = slice_ptr(e, e->prec, bi.path[1], bi.off[1]);
    a.blk[b].M = bi.steps + 1;
    a.blk[b].woff = woff;
    a.blk[b].species = bi.species;
    a.blk[b].poly = bi.poly;
    a.blk[b].dens = 1.0 * e->polys[bi.poly].vf / e->polys[bi.poly].lr;
    woff += bi.steps + 1;
  }
  a.npoly = (int)e->polys.size();
  a.N = e->N;
  a.cw = e->cw;
  a.phiA = e->phiA; a.phiB = e->phiB;
  a.wA = e->wA; a.wB = e->wB; a.praw = e->praw;
  a.wfA = s.wf[0]; a.wfB = s.wf[1];
  a.ringx = e->ringx; a.ringf = e->ringf;
  a.m1 = (mix && mix->kind == 1) ? mix->depth + 1 : 1;
  a.chiN = e->chiN; a.fA = e->fA; a.fB = e->fB;
  a.lam0 = mix ? mix->lambda0 : 0.0;
  a.lam1 = mix ? mix->lambda1 : 0.0;
  a.halfds = 0.5 * e->ds;
  // one polymer: its volfrac / lengthRatio (pspolymer/PsBlock.cpp:170-173)
  // after the 1/Q; a blend scales every block by its own (blend_scale)
  a.dens = 1.0 * e->polys[0].vf / e->polys[0].lr;
  a.st = e->st;
  a.part = e->part;
  a.mode = mode;
  a.f_order = mix ? mix->f_order : 0;
  if ((mode == QM_SD || mode == QM_SD_PHI) && e->noise_rng && e->noise_strength != 0.0) a.noise = e->noise;
  return a;
}

// Draw one steepest-descent step's noise on the host, in the reference's
// order (all of w_A's cells, then w_B's: PsPolymerUpdater.cpp:282-293, each
// value (0 + strength g) * 1), and queue its copy to e->noise ahead of the
// step's kernels.  Strength 0 still draws, as the reference does, but adds
// nothing, so nothing is copied.
psx_status stage_noise(psx_scft* e) {
  if (!e->noise_rng) return PSX_OK;
  const long long n2 = 2 * e->N;
  if (e->noise_strength == 0.0) {
    for (long long i = 0; i < n2; ++i) (void)e->noise_rng->gaussian();
    return PSX_OK;
  }
  const int k = e->noise_k;
  e->noise_k ^= 1;
  PSX_HIP(hipEventSynchronize(e->noise_ev[k]));   // its previous copy has finished
  double* h = e->noise_host[k];
  const double s = e->noise_strength;
  for (long long i = 0; i < n2; ++i) h[i] = (0.0 + s * e->noise_rng->gaussian()) * 1.0;
  PSX_HIP(hipMemcpyAsync(e->noise, h, (size_t)n2 * sizeof(double), hipMemcpyHostToDevice, e->stream));
  PSX_HIP(hipEventRecord(e->noise_ev[k], e->stream));
  return PSX_OK;
}

// The quadrature kernel: FP32 slices as 16-byte vectors of 4 cells when N
// allows (every slice then starts 16-byte aligned; measured 516 -> 437 us at
// 128^3), FP64 cell by cell (double2 measured slower, 780 -> 815 us).
// Always e->nb blocks: k_finish reduces that many partials.
// Contour modes over FP32 vectors on small grids (all vectors in one round
// of 64 per block) split each cell's sum over the block's four waves
// (k_quad_split), unless psx_scft_set_quad_kernel asked for the serial sums.
template <typename T>
psx_status launch_quad_kernel(psx_scft* e, const QuadArgs& a, hipStream_t stream = nullptr) {
  constexpr int V = sizeof(T) == 4 ? 4 : 1;
  if (!stream) stream = e->stream;
  const bool contour = a.mode != QM_SD_PHI && a.mode != QM_FE;
  if (V > 1 && e->N % V == 0 && contour && e->N / V <= (long long)e->nb * 64 && e->quad_kernel == 0) {
    hipLaunchKernelGGL((k_quad_split<T, V, kThreads / 64>), dim3(e->nb), dim3(kThreads), 0, stream, a);
  } else if (V > 1 && e->N % V == 0) {
    hipLaunchKernelGGL((k_quad_update<T, V>), dim3(e->nb), dim3(kThreads), 0, stream, a);
  } else {
    hipLaunchKernelGGL((k_quad_update<T, 1>), dim3(e->nb), dim3(kThreads), 0, stream, a);
  }
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// The quadrature over forward point ranges [lo, hi] of an oriented linear
// chain (block b covers points off[0] .. off[0] + steps): point t pairs
// forward slice t (path 0) with backward slice Ns - t (path 1).
template <typename T>
psx_status launch_quad_points(psx_scft* e, int mode, const psx_mixing* mix,
                              const std::vector<std::pair<int, int>>& ranges, int acc_in,
                              hipStream_t stream) {
  QuadArgs a = quad_args(e, mode, mix);
  a.nblk = 0;
  a.acc_in = acc_in;
  const int Ns = e->paths[0].len;
  int woff = 0;
  for (const BlockInfo& bi : e->blocks) {
    const int bs = bi.off[0];
    for (const auto& r : ranges) {
      const int lo = std::max(bs, r.first), hi = std::min(bs + bi.steps, r.second);
      if (lo > hi) continue;
      if (a.nblk == kMaxBlocks) return fail(PSX_ERR_INVALID, "quadrature chunk spans too many blocks");
      QuadBlock& q = a.blk[a.nblk++];
      q.H = slice_ptr(e, e->prec, 0, lo);
      q.T = slice_ptr(e, e->prec, 1, Ns - hi);
      q.M = hi - lo + 1;
      q.woff = woff + (lo - bs);
      q.species = bi.species;
      q.poly = bi.poly;
      q.dens = 1.0 * e->polys[bi.poly].vf / e->polys[bi.poly].lr;
    }
    woff += bi.steps + 1;
  }
  return launch_quad_kernel<T>(e, a, stream);
}

// Checkpointed mode: the contour sums chunk by chunk (raw sums carried in
// phiA/phiB), the last chunk finishing in `mode`.  Every (block, point) pair
// belongs to one chunk: chunks own [t0, t1) and the last one also t1.
template <typename T>
psx_status launch_quad_ckpt(psx_scft* e, int mode, const psx_mixing* mix) {
  SweepBufs& s = e->sw[e->prec];
  const size_t sb = (size_t)e->N * sizeof(T);
  const int Ns = e->paths[1].len;
  std::vector<int> woff(e->blocks.size());
  for (size_t b = 0, w = 0; b < e->blocks.size(); ++b) { woff[b] = (int)w; w += e->blocks[b].steps + 1; }
  for (size_t ci = 0; ci < e->segs.size(); ++ci) {
    PSX_TRY(ckpt_chunk(e, ci));
    const CkptSeg& c = e->segs[ci];
    const bool last = ci + 1 == e->segs.size();
    QuadArgs a = quad_args(e, last ? mode : QM_ACC, mix);
    a.acc_in = ci > 0;
    const int tend = last ? c.t1 : c.t1 - 1;
    for (size_t b = 0; b < e->blocks.size(); ++b) {
      const BlockInfo& bi = e->blocks[b];
      const int bs = bi.off[0], lo = std::max(bs, c.t0), hi = std::min(bs + bi.steps, tend);
      if (lo > hi) continue;
      QuadBlock& q = a.blk[a.nblk++];
      q.H = (const char*)s.fch[ci % 2] + (size_t)(lo - c.t0) * sb;
      q.T = (const char*)s.seg + (size_t)(Ns - hi - c.u0) * sb;
      q.M = hi - lo + 1;
      q.woff = woff[b] + (lo - bs);
      q.species = bi.species;
    }
    PSX_TRY(launch_quad_kernel<T>(e, a));
  }
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

psx_status launch_quad(psx_scft* e, int mode, const psx_mixing* mix) {
  if (e->ckpt_K > 0 && mode != QM_SD_PHI && mode != QM_FE)
    return e->prec == PSX_FP64 ? launch_quad_ckpt<double>(e, mode, mix) : launch_quad_ckpt<float>(e, mode, mix);
  if (e->quad_partial && mode != QM_SD_PHI && mode != QM_FE) {
    // the earlier chunks ran under the sweeps; the last one finishes here
    const int acc = e->qchunks.size() > 1 ? 1 : 0;
    return e->prec == PSX_FP64
               ? launch_quad_points<double>(e, mode, mix, e->qchunks.back().ranges, acc, e->stream)
               : launch_quad_points<float>(e, mode, mix, e->qchunks.back().ranges, acc, e->stream);
  }
  QuadArgs a = quad_args(e, mode, mix);
  if (e->prec == PSX_FP64) PSX_TRY(launch_quad_kernel<double>(e, a));
  else PSX_TRY(launch_quad_kernel<float>(e, a));
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

psx_status launch_finish(psx_scft* e, int mode, double* hist) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, e->stream, (const double*)e->part, e->nb,
                     1.0 / (double)e->N, e->chiN, e->fA, e->fB, mode, e->st, hist);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

psx_status ensure_ring(psx_scft* e, int depth) {
  const int m1 = depth + 1;
  if (e->ring_m1 >= m1 && e->ringx) return PSX_OK;
  drop_graphs(e);
  dfree((void*&)e->ringx);
  dfree((void*&)e->ringf);
  const size_t bytes = (size_t)m1 * 2 * e->N * sizeof(double);
  PSX_TRY(dalloc((void**)&e->ringx, bytes));
  PSX_TRY(dalloc((void**)&e->ringf, bytes));
  PSX_HIP(hipMemset(e->ringx, 0, bytes));
  PSX_HIP(hipMemset(e->ringf, 0, bytes));
  e->ring_m1 = m1;
  const int need = e->nb * (m1 > kNPart ? m1 : kNPart);
  if (need > e->part_cap) {
    dfree((void*&)e->part);
    PSX_TRY(dalloc((void**)&e->part, need * sizeof(double)));
    e->part_cap = need;
  }
  hipLaunchKernelGGL(k_state_reset_hist, dim3(1), dim3(1), 0, e->stream, e->st);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

psx_status enqueue_post_sweep(psx_scft* e, const psx_mixing* mix);

// Enqueue one full SCFT iteration on e->stream (capturable).
psx_status enqueue_iteration(psx_scft* e, const psx_mixing* mix) {
  const bool ov = !e->qchunks.empty() && e->split_sweeps && e->ckpt_K == 0;
  if (!ov) {
    PSX_TRY(sweep(e));
    return enqueue_post_sweep(e, mix);
  }
  PSX_TRY(e->prec == PSX_FP64 ? sweep_range<double>(e, -1, true) : sweep_range<float>(e, -1, true));
  e->quad_partial = true;
  const psx_status st = enqueue_post_sweep(e, mix);
  e->quad_partial = false;
  return st;
}

// The Anderson ring holds the sweep's type R; its slots are read as 16-byte
// vectors when 2N allows (every slot then starts 16-byte aligned).
template <typename R, int V>
void launch_gram_v(psx_scft* e, int m1) {
  const R* rf = (const R*)e->ringf;
  const DevState* st = e->st;
  if (m1 <= 4)
    hipLaunchKernelGGL((k_gram<4, R, V>), dim3(e->nb), dim3(kThreads), 0, e->stream, rf, e->N, m1, st, e->part);
  else if (m1 <= 11)
    hipLaunchKernelGGL((k_gram<11, R, V>), dim3(e->nb), dim3(kThreads), 0, e->stream, rf, e->N, m1, st, e->part);
  else if (m1 <= 21)
    hipLaunchKernelGGL((k_gram<21, R, V>), dim3(e->nb), dim3(kThreads), 0, e->stream, rf, e->N, m1, st, e->part);
  else
    hipLaunchKernelGGL((k_gram<kMaxDepth + 1, R, V>), dim3(e->nb), dim3(kThreads), 0, e->stream, rf, e->N, m1,
                       st, e->part);
}

template <typename R>
void launch_gram(psx_scft* e, int m1) {
  constexpr int V = 16 / (int)sizeof(R);
  if ((2 * e->N) % V == 0) launch_gram_v<R, V>(e, m1);
  else launch_gram_v<R, 1>(e, m1);
}

// The mixing step reads 16-byte vectors of cells when N allows, with its
// history loop unrolled to the depth's bucket.
template <typename R, int V>
void launch_mix_v(psx_scft* e, const QuadArgs& a, int m) {
  const dim3 gr(e->nb), bl(kThreads);
  if (m <= 3) hipLaunchKernelGGL((k_mix<R, V, 3>), gr, bl, 0, e->stream, a);
  else if (m <= 10) hipLaunchKernelGGL((k_mix<R, V, 10>), gr, bl, 0, e->stream, a);
  else if (m <= 20) hipLaunchKernelGGL((k_mix<R, V, 20>), gr, bl, 0, e->stream, a);
  else hipLaunchKernelGGL((k_mix<R, V, kMaxDepth>), gr, bl, 0, e->stream, a);
}

template <typename R>
void launch_mix(psx_scft* e, const QuadArgs& a, int m) {
  constexpr int V = 16 / (int)sizeof(R);
  if (e->N % V == 0) launch_mix_v<R, V>(e, a, m);
  else launch_mix_v<R, 1>(e, a, m);
}

// Q, quadrature and the field update after both sweeps (capturable).
psx_status enqueue_post_sweep(psx_scft* e, const psx_mixing* mix) {
  PSX_TRY(launch_Q(e));
  if (mix->kind == 0) {
    PSX_TRY(launch_quad(e, QM_SD, mix));
    PSX_TRY(launch_finish(e, FIN_ITER_SD, e->hist));
    return PSX_OK;
  }
  const int m = mix->depth;
  const int m1 = m + 1;
  PSX_TRY(launch_quad(e, QM_RESID, mix));
  PSX_TRY(launch_finish(e, FIN_ERR, nullptr));
  if (e->prec == PSX_FP64) launch_gram<double>(e, m1);
  else launch_gram<float>(e, m1);
  hipLaunchKernelGGL(k_gram_finish, dim3(m1), dim3(kThreads), 0, e->stream, (const double*)e->part, e->nb, m1,
                     e->st);
  hipLaunchKernelGGL(k_anderson_solve, dim3(1), dim3(64), 0, e->stream, m, mix->beta, mix->lambda0,
                     mix->anderson_start, e->st);
  QuadArgs a = quad_args(e, QM_SD, mix);
  if (e->prec == PSX_FP64) launch_mix<double>(e, a, m);
  else launch_mix<float>(e, a, m);
  PSX_HIP(hipGetLastError());
  PSX_TRY(launch_finish(e, FIN_ITER_MIX, e->hist));
  return PSX_OK;
}

psx_status read_state(psx_scft* e, DevState* h) {
  PSX_HIP(hipMemcpyAsync(h, e->st, sizeof(DevState), hipMemcpyDeviceToHost, e->stream));
  PSX_HIP(hipStreamSynchronize(e->stream));
  return PSX_OK;
}

}  // namespace

extern "C" psx_status psx_scft_destroy(psx_scft* e) {
  if (!e) return PSX_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graphs(e);
  psx_fft_plan_destroy(e->fplan);
  dfree((void*&)e->noise);
  for (int k = 0; k < 2; ++k) {
    if (e->noise_host[k]) (void)hipHostFree(e->noise_host[k]);
    if (e->noise_ev[k]) (void)hipEventDestroy(e->noise_ev[k]);
  }
  for (auto& s : e->sw) free_sweep(s);
  for (auto& a : e->axes) { dfree(a.tw[0]); dfree(a.tw[1]); }
  dfree((void*&)e->wA); dfree((void*&)e->wB); dfree((void*&)e->praw);
  dfree((void*&)e->phiA); dfree((void*&)e->phiB); dfree((void*&)e->cw);
  dfree((void*&)e->part); dfree((void*&)e->st); dfree((void*&)e->hist);
  dfree((void*&)e->ringx); dfree((void*&)e->ringf); dfree(e->hstage);
  dfree(e->zbuf); dfree((void*&)e->sacc);
  dfree((void*&)e->adj); dfree((void*&)e->adj_tab);
  for (auto& p : e->kd2) dfree((void*&)p);
  for (auto& b : e->kf1) for (auto& p : b) dfree((void*&)p);
  for (auto& a : e->caxes) { dfree(a.tw[0]); dfree(a.tw[1]); }
  for (size_t k = 1; k < e->pstream.size(); ++k)
    if (e->pstream[k]) { (void)hipStreamSynchronize(e->pstream[k]); (void)hipStreamDestroy(e->pstream[k]); }
  if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
  for (auto& ev : e->ev_join) if (ev) (void)hipEventDestroy(ev);
  if (e->qstream) { (void)hipStreamSynchronize(e->qstream); (void)hipStreamDestroy(e->qstream); }
  if (e->ev_qdone) (void)hipEventDestroy(e->ev_qdone);
  for (auto& c : e->qchunks) for (auto& ev : c.ev) if (ev) (void)hipEventDestroy(ev);
  if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PSX_OK;
}

// Block steps per psbase/PsBlockBase.cpp:252-326 (forceBlockSteps = false).
static int block_steps(double len, double ds) {
  const double bs = len / ds;
  const double ip = (double)(long)bs;
  const double dx = bs - ip;
  if (dx > 0.0001 && dx < 0.9999) return -1;
  const long c = (long)std::ceil(bs), f = (long)std::floor(bs);
  long r;
  if (c == f) r = c;
  else if (std::fabs(bs - (double)c) < std::fabs(bs - (double)f)) r = c;
  else r = f;
  if (r < 4) return -2;
  return (int)r;
}

// Trapezoid weights ds [1/2, 1, ..., 1, 1/2]: the exact derivative of the
// split-step Q with respect to w(r), since each step's two w-factor
// half-multiplies touch the contour points at both of its ends.
static void trap_weights(int steps, double ds, double* w) {
  for (int j = 0; j <= steps; ++j) w[j] = ds;
  w[0] = 0.5 * ds;
  w[steps] = 0.5 * ds;
}

// Simpson (literal 0.33333) with a trapezoid tail for an even point count,
// pspolymer/PsBlock.cpp:256-380, as per-point weights over M = steps+1.
static void quad_weights(int steps, double ds, double* w) {
  const int M = steps + 1;
  const bool even = (M % 2) == 0;
  const int last = even ? M - 2 : M - 1;   // Simpson covers points 0..last
  for (int j = 0; j < M; ++j) w[j] = 0.0;
  for (int n = 0; n + 2 <= last; n += 2) {
    w[n] += 1.0; w[n + 1] += 4.0; w[n + 2] += 1.0;
  }
  for (int j = 0; j <= last; ++j) w[j] *= 0.33333 * ds;
  if (even) { w[M - 2] += 0.5 * ds; w[M - 1] += 0.5 * ds; }
}

// Per-block quadrature weights, blocks in order (mode 0 reference Simpson,
// 1 trapezoid).
static std::vector<double> all_weights(const psx_scft* e, int mode) {
  std::vector<double> w;
  for (const BlockInfo& b : e->blocks) {
    const size_t o = w.size();
    w.resize(o + b.steps + 1);
    if (mode == 0) quad_weights(b.steps, e->ds, w.data() + o);
    else trap_weights(b.steps, e->ds, w.data() + o);
  }
  return w;
}

// Blocks, junctions and propagation paths of the graph d[0..nb).
// Launch lists of every path for the engine's sweep mode.
static void plan_launches(psx_scft* e) {
  const int na = (int)e->axes.size();
  for (PathInfo& pi : e->paths) {
    pi.launch_axis.clear();
    pi.launch_nops.clear();
    for (const auto& l : sweep_launches(na, pi.len, e->plane)) {
      pi.launch_axis.push_back(l.first.axis);
      pi.launch_nops.push_back(l.second);
    }
  }
}

// Plane launches need a 3D grid whose y and z lengths are equal and one of
// the plane kernel's lengths.
static bool plane_available(const psx_scft* e) {
  return e->axes.size() == 3 && e->axes[0].n == e->axes[1].n && plane_length(e->axes[0].n);
}

// The block graph of all polymers.  e->polys is set: polymer p holds blocks
// first .. first + n - 1, and the head/tail indices of `d` count within the
// block's polymer.
static psx_status build_graph(psx_scft* e, int nb, const psx_block* d) {
  if (nb < 1 || nb > PSX_MAX_BLOCKS)
    return fail(PSX_ERR_INVALID, "block count must be in [1, " + std::to_string(PSX_MAX_BLOCKS) + "]");
  e->blocks.assign(nb, BlockInfo());
  for (size_t p = 0; p < e->polys.size(); ++p)
    for (int i = e->polys[p].first; i < e->polys[p].first + e->polys[p].n; ++i) e->blocks[i].poly = (int)p;
  e->bval.clear();
  for (int i = 0; i < nb; ++i) {
    const psx_block& b = d[i];
    if (b.species != 0 && b.species != 1) return fail(PSX_ERR_INVALID, "block species must be 0 (A) or 1 (B)");
    if (!(b.lengthfrac > 0) || !(b.bSegRatio > 0)) return fail(PSX_ERR_INVALID, "need lengthfrac > 0 and bSegRatio > 0");
    if (b.nhead < 0 || b.ntail < 0 || b.nhead > PSX_MAX_JOIN || b.ntail > PSX_MAX_JOIN)
      return fail(PSX_ERR_INVALID, "joined-block counts must be in [0, " + std::to_string(PSX_MAX_JOIN) + "]");
    BlockInfo& bi = e->blocks[i];
    bi.species = b.species;
    bi.frac = b.lengthfrac;
    bi.b = b.bSegRatio;
    // blockSteps = lengthRatio * lengthFrac / ds (psbase/PsBlockBase.cpp:192-197)
    const double lr = e->polys[bi.poly].lr;
    bi.steps = block_steps(lr * b.lengthfrac, e->ds);
    if (bi.steps < 0)
      return fail(PSX_ERR_STEPS, "block " + std::to_string(i) + ": non-integral or too few block steps "
                                 "(lengthRatio*lengthfrac/ds = " + std::to_string(lr * b.lengthfrac / e->ds) + ")");
    bi.bclass = -1;
    for (size_t c = 0; c < e->bval.size(); ++c)
      if (e->bval[c] == b.bSegRatio) bi.bclass = (int)c;
    if (bi.bclass < 0) { bi.bclass = (int)e->bval.size(); e->bval.push_back(b.bSegRatio); }
  }
  // junctions (PsBlockCopolymer::buildCntBlockType): block j named at end
  // E of block i is joined at j's tail when j's tail list names i
  for (int i = 0; i < nb; ++i)
    for (int E = 0; E < 2; ++E) {
      const int n = E == 0 ? d[i].nhead : d[i].ntail;
      const int* lst = E == 0 ? d[i].head : d[i].tail;
      const auto& pi = e->polys[e->blocks[i].poly];
      for (int k = 0; k < n; ++k) {
        if (lst[k] < 0 || lst[k] >= pi.n)
          return fail(PSX_ERR_INVALID, "block " + std::to_string(i) + ": joined block index out of range");
        const int j = pi.first + lst[k];
        if (j == i)
          return fail(PSX_ERR_INVALID, "block " + std::to_string(i) + ": joined to itself");
        for (const auto& c : e->blocks[i].cnt[E])
          if (c.first == j) return fail(PSX_ERR_INVALID, "block " + std::to_string(i) + " joins block " + std::to_string(j) + " twice");
        bool at_tail = false;
        for (int m = 0; m < d[j].ntail; ++m)
          if (pi.first + d[j].tail[m] == i) at_tail = true;
        e->blocks[i].cnt[E].push_back({j, at_tail ? 1 : 0});
      }
      // junction products run in ascending block index (std::map order)
      std::sort(e->blocks[i].cnt[E].begin(), e->blocks[i].cnt[E].end());
    }
  // Propagation tasks: task 2b + dir is block b propagated from its head
  // (dir 0, ends at the tail) or from its tail (dir 1, ends at the head).
  // Its inputs are the tasks that end at the blocks joined to its start end.
  const int nt = 2 * nb;
  auto inputs = [&](int t) {
    std::vector<int> in;
    for (const auto& c : e->blocks[t / 2].cnt[t % 2]) in.push_back(2 * c.first + (c.second == 1 ? 0 : 1));
    return in;
  };
  std::vector<int> tpath(nt, -1), toff(nt, 0), plast;
  e->paths.clear();
  int left = nt;
  while (left > 0) {
    bool progress = false;
    for (int t = 0; t < nt; ++t) {
      if (tpath[t] >= 0) continue;
      const std::vector<int> in = inputs(t);
      bool ready = true;
      for (int c : in) ready = ready && tpath[c] >= 0;
      if (!ready) continue;
      const int b = t / 2, steps = e->blocks[b].steps;
      int p;
      if (in.size() == 1 && plast[tpath[in[0]]] == in[0]) {
        p = tpath[in[0]];               // continues the path of its single input
      } else {
        p = (int)e->paths.size();
        PathInfo pi;
        for (int c : in) {
          pi.init.push_back({tpath[c], toff[c] + e->blocks[c / 2].steps});
          pi.round = std::max(pi.round, e->paths[tpath[c]].round + 1);
        }
        e->paths.push_back(pi);
        plast.push_back(-1);
      }
      PathInfo& pi = e->paths[p];
      tpath[t] = p;
      toff[t] = pi.len;
      for (int k = 0; k < steps; ++k) pi.step_block.push_back(b);
      pi.len += steps;
      plast[p] = t;
      e->blocks[b].path[t % 2] = p;
      e->blocks[b].off[t % 2] = toff[t];
      --left;
      progress = true;
    }
    if (!progress)
      return fail(PSX_ERR_INVALID, "block graph has a cycle or a junction no free end reaches");
  }
  int nr = 0;
  for (const PathInfo& pi : e->paths) nr = std::max(nr, pi.round + 1);
  e->rounds.assign(nr, {});
  const int na = (int)e->axes.size();
  for (size_t p = 0; p < e->paths.size(); ++p) {
    PathInfo& pi = e->paths[p];
    pi.slot = (int)e->rounds[pi.round].size();
    e->rounds[pi.round].push_back((int)p);
  }
  (void)na;
  plan_launches(e);
  // species totals (lengthWeight * lengthFrac * volfrac, psbase/PsBlockBase.cpp:
  // 185-189), contour steps
  e->fA = e->fB = 0.0;
  e->Ns = 0;
  for (const BlockInfo& b : e->blocks) {
    (b.species == 0 ? e->fA : e->fB) += 1.0 * b.frac * e->polys[b.poly].vf;
    e->Ns += b.steps;
  }
  e->diblock = nb == 2 && e->polys.size() == 1 && e->blocks[0].species == 0 && e->blocks[1].species == 1 &&
               d[0].nhead == 0 && d[0].ntail == 1 && d[0].tail[0] == 1 &&
               d[1].nhead == 1 && d[1].head[0] == 0 && d[1].ntail == 0;
  return PSX_OK;
}

// Chunks of contour points for the quadrature under the sweeps (see
// psx_scft::qchunks): oriented linear chains, two paths in one round.
static void plan_quad_overlap(psx_scft* e) {
  // a re-plan (new launch lists) keeps the chunks' events
  std::vector<std::array<hipEvent_t, 2>> evs;
  for (const auto& c : e->qchunks) evs.push_back({c.ev[0], c.ev[1]});
  struct Restore {
    psx_scft* e;
    std::vector<std::array<hipEvent_t, 2>>& evs;
    ~Restore() {
      for (size_t c = 0; c < evs.size() && c < e->qchunks.size(); ++c) {
        e->qchunks[c].ev[0] = evs[c][0];
        e->qchunks[c].ev[1] = evs[c][1];
      }
    }
  } restore{e, evs};
  e->qchunks.clear();
  if (!e->overlap_quad || e->rounds.size() != 1 || e->paths.size() != 2) return;
  if (!e->blocks[0].cnt[0].empty() || 2 * e->blocks.size() > (size_t)kMaxBlocks) return;
  for (const BlockInfo& b : e->blocks)
    if (b.path[0] != 0 || b.path[1] != 1) return;
  const int Ns = e->paths[0].len;
  if (Ns < 8) return;
  // highest slice each launch of a path has written
  const int na = (int)e->axes.size();
  std::vector<int> written;
  int w = 0;
  for (const auto& l : sweep_launches(na, Ns, e->plane)) {
    const SweepOp& o = l.first;
    if (o.axis == kPlaneAxis) {
      if (l.second == 4 || o.last) w = std::max(w, o.step + 1);
    } else if (l.second == 2) {
      w = std::max(w, o.step + 1 + (na == 1 ? 1 : 0));
    } else if (o.last) {
      w = std::max(w, o.step + 1);
    }
    written.push_back(w);
  }
  // point t is ready at step r(t) = max(t, Ns - t) in [rmin, Ns]
  const int nch = 4, rmin = (Ns + 1) / 2;
  std::vector<int> S(nch);
  for (int c = 0; c < nch; ++c) S[c] = rmin + ((Ns - rmin) * (c + 1) + nch - 1) / nch;
  S[nch - 1] = Ns;
  e->qchunks.assign(nch, psx_scft::QChunk());
  for (int t = 0; t <= Ns; ++t) {
    const int r = std::max(t, Ns - t);
    int c = 0;
    while (r > S[c]) ++c;
    auto& rg = e->qchunks[c].ranges;
    if (!rg.empty() && rg.back().second == t - 1) rg.back().second = t;
    else rg.push_back({t, t});
  }
  for (int c = 0; c < nch; ++c) {
    int j = 0;
    while (j + 1 < (int)written.size() && written[j] < S[c]) ++j;
    e->qchunks[c].launch = j;
  }
}

static psx_status create_engine(const psx_grid* g, int nb, const psx_block* blocks, double ds,
                                double chiN, psx_prec prec, void* stream, psx_scft** out,
                                int npoly = 0, const psx_polymer* polymers = nullptr) {
  if (!out || !g || !blocks) return fail(PSX_ERR_INVALID, "null argument");
  *out = nullptr;
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(PSX_ERR_SHAPE, "grid dims must be >= 1");
  if (!(g->dx > 0 && g->dy > 0 && g->dz > 0)) return fail(PSX_ERR_INVALID, "cell sizes must be > 0");
  if ((long long)g->nx * g->ny * g->nz > (1LL << 30)) return fail(PSX_ERR_SHAPE, "grid too large");
  if (prec != PSX_FP64 && prec != PSX_FP32) return fail(PSX_ERR_INVALID, "bad precision");
  if (!(ds > 0)) return fail(PSX_ERR_INVALID, "need ds > 0");
  psx_scft* e = new psx_scft();
  e->dims[0] = g->nx; e->dims[1] = g->ny; e->dims[2] = g->nz;
  e->N = (long long)g->nx * g->ny * g->nz;
  e->cell[0] = g->dx; e->cell[1] = g->dy; e->cell[2] = g->dz;
  e->ds = ds;
  e->chiN = chiN;
  e->prec = prec;
  psx_status st = make_axes(e->dims, &e->axes);
  if (npoly <= 0) {
    e->polys.assign(1, psx_scft::PolyInfo());
    e->polys[0].n = nb;
  } else {
    // lengthRatio = length / the first polymer's length (pspolymer/PsPolymerHldr.cpp:69-83)
    int first = 0;
    for (int p = 0; p < npoly && st == PSX_OK; ++p) {
      const psx_polymer& q = polymers[p];
      if (!(q.volfrac > 0) || !(q.length > 0) || q.nblocks < 1)
        st = fail(PSX_ERR_INVALID, "polymer " + std::to_string(p) + ": need volfrac > 0, length > 0 and a block");
      psx_scft::PolyInfo pi;
      pi.vf = q.volfrac;
      pi.lr = q.length / polymers[0].length;
      pi.first = first;
      pi.n = q.nblocks;
      first += q.nblocks;
      e->polys.push_back(pi);
    }
    if (st == PSX_OK && first != nb) st = fail(PSX_ERR_INVALID, "the polymers' block counts do not add up to nblocks");
  }
  e->plane = plane_available(e);
  if (const char* v = getenv("PSX_PLANE")) e->plane = e->plane && v[0] != '0';
  if (st == PSX_OK) st = build_graph(e, nb, blocks);
  if (st != PSX_OK) { delete e; return st; }
#define E_TRY(x) do { if (st == PSX_OK) st = (x); } while (0)
  E_TRY((hipGetDevice(&e->dev) == hipSuccess) ? PSX_OK : fail(PSX_ERR_HIP, "hipGetDevice failed"));
  E_TRY(make_stream(stream, &e->stream, &e->own_stream));
  size_t slots = 1;
  for (const auto& r : e->rounds) slots = std::max(slots, r.size());
  e->pstream.assign(slots, nullptr);
  e->ev_join.assign(slots, nullptr);
  e->pstream[0] = e->stream;
  for (size_t k = 1; k < slots; ++k)
    E_TRY(hipStreamCreateWithFlags(&e->pstream[k], hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&e->ev_join[k], hipEventDisableTiming) == hipSuccess
              ? PSX_OK : fail(PSX_ERR_HIP, "stream creation failed"));
  E_TRY(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) == hipSuccess
            ? PSX_OK : fail(PSX_ERR_HIP, "event creation failed"));
  if (const char* v = getenv("PSX_SPLIT_SWEEPS")) e->split_sweeps = v[0] != '0';
  if (const char* v = getenv("PSX_OVERLAP_QUAD")) e->overlap_quad = v[0] == '1';
  if (st == PSX_OK) {
    plan_quad_overlap(e);
    if (!e->qchunks.empty()) {
      E_TRY(hipStreamCreateWithFlags(&e->qstream, hipStreamNonBlocking) == hipSuccess &&
                    hipEventCreateWithFlags(&e->ev_qdone, hipEventDisableTiming) == hipSuccess
                ? PSX_OK : fail(PSX_ERR_HIP, "stream creation failed"));
      for (auto& c : e->qchunks)
        for (auto& ev : c.ev)
          E_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess
                    ? PSX_OK : fail(PSX_ERR_HIP, "event creation failed"));
    }
  }
  const size_t fb = e->N * sizeof(double);
  E_TRY(dalloc((void**)&e->wA, fb));
  E_TRY(dalloc((void**)&e->wB, fb));
  E_TRY(dalloc((void**)&e->praw, fb));
  E_TRY(dalloc((void**)&e->phiA, fb));
  E_TRY(dalloc((void**)&e->phiB, fb));
  e->nb = grid_1d(e->N);
  e->part_cap = e->nb * kNPart;
  E_TRY(dalloc((void**)&e->part, e->part_cap * sizeof(double)));
  E_TRY(dalloc((void**)&e->st, sizeof(DevState)));
  if (st == PSX_OK) {
    std::vector<double> w = all_weights(e, 0);
    st = upload((void**)&e->cw, w.data(), w.size() * sizeof(double));
  }
  if (st == PSX_OK) {
    DevState h;
    std::memset(&h, 0, sizeof(h));
    h.Q = 1.0;
    h.npoly = (int)e->polys.size();
    for (int p = 0; p < kMaxPoly; ++p) {
      h.Qp[p] = 1.0;
      h.pvf[p] = p < h.npoly ? e->polys[p].vf : 0.0;
      h.plr[p] = p < h.npoly ? e->polys[p].lr : 1.0;
    }
    st = (hipMemcpy(e->st, &h, sizeof(h), hipMemcpyHostToDevice) == hipSuccess)
             ? PSX_OK : fail(PSX_ERR_HIP, "state upload failed");
  }
  if (st == PSX_OK) {
    const bool ok = hipMemset(e->wA, 0, fb) == hipSuccess && hipMemset(e->wB, 0, fb) == hipSuccess &&
                    hipMemset(e->praw, 0, fb) == hipSuccess && hipMemset(e->phiA, 0, fb) == hipSuccess &&
                    hipMemset(e->phiB, 0, fb) == hipSuccess;
    if (!ok) st = fail(PSX_ERR_HIP, "field init failed");
  }
  E_TRY(ensure_sweep(e, prec));
#undef E_TRY
  if (st != PSX_OK) { psx_scft_destroy(e); return st; }
  *out = e;
  return PSX_OK;
}

extern "C" psx_status psx_scft_create(const psx_grid* g, const psx_diblock* c,
                                      psx_prec prec, void* stream, psx_scft** out) {
  if (!out || !g || !c) return fail(PSX_ERR_INVALID, "null argument");
  *out = nullptr;
  if (!(c->ds > 0) || !(c->fA > 0 && c->fA < 1)) return fail(PSX_ERR_INVALID, "need ds > 0 and 0 < fA < 1");
  if (block_steps(c->fA, c->ds) < 0 || block_steps(1.0 - c->fA, c->ds) < 0)
    return fail(PSX_ERR_STEPS, "non-integral or too few block steps (fA/ds = " +
                                   std::to_string(c->fA / c->ds) + ", (1-fA)/ds = " +
                                   std::to_string((1.0 - c->fA) / c->ds) + ")");
  // A (head free) joined at its tail to the head of B (tail free)
  psx_block b[2];
  std::memset(b, 0, sizeof(b));
  b[0].species = 0; b[0].lengthfrac = c->fA; b[0].bSegRatio = c->bA; b[0].ntail = 1; b[0].tail[0] = 1;
  b[1].species = 1; b[1].lengthfrac = 1.0 - c->fA; b[1].bSegRatio = c->bB; b[1].nhead = 1; b[1].head[0] = 0;
  return create_engine(g, 2, b, c->ds, c->chiN, prec, stream, out);
}

extern "C" psx_status psx_scft_create_blocks(const psx_grid* g, int nblocks, const psx_block* blocks,
                                             double ds, double chiN, psx_prec prec, void* stream,
                                             psx_scft** out) {
  return create_engine(g, nblocks, blocks, ds, chiN, prec, stream, out);
}

extern "C" psx_status psx_scft_create_blend(const psx_grid* g, int npolymers, const psx_polymer* polymers,
                                            const psx_block* blocks, double ds, double chiN, psx_prec prec,
                                            void* stream, psx_scft** out) {
  if (!out || !polymers) return fail(PSX_ERR_INVALID, "null argument");
  *out = nullptr;
  if (npolymers < 1 || npolymers > PSX_MAX_POLYMERS)
    return fail(PSX_ERR_INVALID, "polymer count must be in [1, " + std::to_string(PSX_MAX_POLYMERS) + "]");
  int nb = 0;
  for (int p = 0; p < npolymers; ++p) nb += polymers[p].nblocks > 0 ? polymers[p].nblocks : 0;
  return create_engine(g, nb, blocks, ds, chiN, prec, stream, out, npolymers, polymers);
}

extern "C" psx_status psx_scft_graph_info(const psx_scft* e, int* nblocks, int* npaths, int* nrounds) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (nblocks) *nblocks = (int)e->blocks.size();
  if (npaths) *npaths = (int)e->paths.size();
  if (nrounds) *nrounds = (int)e->rounds.size();
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_checkpointing(psx_scft* e, int every) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (every < 0) return fail(PSX_ERR_INVALID, "checkpoint interval must be >= 0");
  if (every > 0) {
    if (e->in_batch) return fail(PSX_ERR_STATE, "engine belongs to a batch");
    bool chain = e->rounds.size() == 1 && e->paths.size() == 2 && e->paths[0].init.empty() &&
                 e->blocks[0].cnt[0].empty();
    for (const BlockInfo& b : e->blocks) chain = chain && b.path[0] == 0 && b.path[1] == 1;
    if (!chain)
      return fail(PSX_ERR_INVALID, "checkpointing needs a linear chain with every block oriented "
                                   "head to tail from block 0");
  }
  if (every == e->ckpt_K) return PSX_OK;
  PSX_HIP(hipStreamSynchronize(e->stream));
  drop_graphs(e);
  for (auto& sw : e->sw) free_sweep(sw);
  e->ckpt_K = every;
  if (every > 0) ckpt_plan(e);
  else { e->segs.clear(); e->btab_axis.clear(); e->nck = 0; }
  PSX_TRY(ensure_sweep(e, e->prec));
  e->swept = false;
  e->dens_valid = false;
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_quadrature(psx_scft* e, int mode) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (mode != 0 && mode != 1) return fail(PSX_ERR_INVALID, "quadrature mode must be 0 or 1");
  std::vector<double> w = all_weights(e, mode);
  PSX_HIP(hipStreamSynchronize(e->stream));
  PSX_HIP(hipMemcpy(e->cw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  e->quad_mode = mode;
  e->dens_valid = false;
  return PSX_OK;
}

// FP32 ring slots (x and f) to FP64 in place, highest slot first: the FP64
// copy of slot s lies above every FP32 slot below s, and slot 0 goes
// through a temporary.
static psx_status ring_to_f64(psx_scft* e) {
  const long long n = 2 * e->N;
  double* tmp = nullptr;
  PSX_TRY(dalloc((void**)&tmp, (size_t)n * sizeof(double)));
  psx_status st = PSX_OK;
  for (void* ring : {e->ringx, e->ringf})
    for (int s = e->ring_m1 - 1; s >= 0 && st == PSX_OK; --s) {
      hipLaunchKernelGGL(k_slot_f32_to_f64, dim3(grid_1d(n)), dim3(kThreads), 0, e->stream,
                         (const float*)ring + (size_t)s * n, n, tmp);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync((double*)ring + (size_t)s * n, tmp, (size_t)n * sizeof(double),
                         hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
        st = fail(PSX_ERR_HIP, "ring conversion failed");
    }
  if (hipStreamSynchronize(e->stream) != hipSuccess && st == PSX_OK) st = fail(PSX_ERR_HIP, "sync failed");
  dfree((void*&)tmp);
  return st;
}

static psx_status set_precision_impl(psx_scft* e, psx_prec prec) {
  if (e->prec == (int)prec) return PSX_OK;
  PSX_TRY(ensure_sweep(e, prec));
  const int old = e->prec;
  e->prec = prec;
  // The Anderson ring is kept in the sweep's type.  FP32 -> FP64 converts
  // its entries exactly, so the history (and the Gram matrix built from
  // them) carries on; FP64 -> FP32 would round them, so the history starts
  // again.
  if (e->ringx) {
    if (old == PSX_FP32 && prec == PSX_FP64) {
      PSX_TRY(ring_to_f64(e));
    } else {
      hipLaunchKernelGGL(k_state_reset_hist, dim3(1), dim3(1), 0, e->stream, e->st);
      PSX_HIP(hipGetLastError());
    }
  }
  e->sw[prec].wfac_valid = false;
  e->swept = false;
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_precision(psx_scft* e, psx_prec prec) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (prec != PSX_FP64 && prec != PSX_FP32) return fail(PSX_ERR_INVALID, "bad precision");
  if (e->prec == (int)prec) return PSX_OK;
  // a batch's class launches write the slices of the batch's precision, and
  // the instance's own quadrature reads those of e->prec: they switch together
  if (e->in_batch) return fail(PSX_ERR_STATE, "engine belongs to a batch: use psx_batch_set_precision");
  return set_precision_impl(e, prec);
}

extern "C" psx_status psx_scft_set_quad_kernel(psx_scft* e, int kind) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (kind != 0 && kind != 1) return fail(PSX_ERR_INVALID, "quadrature kernel must be 0 or 1");
  if (kind == e->quad_kernel) return PSX_OK;
  PSX_HIP(hipStreamSynchronize(e->stream));
  drop_graphs(e);
  e->quad_kernel = kind;
  return PSX_OK;
}

extern "C" psx_status psx_scft_block_steps(const psx_scft* e, int* sA, int* sB) {
  if (!e || !sA || !sB) return fail(PSX_ERR_INVALID, "null argument");
  *sA = 0; *sB = 0;
  for (const BlockInfo& b : e->blocks) (b.species == 0 ? *sA : *sB) += b.steps;
  return PSX_OK;
}

static void fields_changed(psx_scft* e) {
  e->sw[0].wfac_valid = false;
  e->sw[1].wfac_valid = false;
  e->swept = false;
  e->dens_valid = false;
}

// p is stored as p_raw with pbar = 0.
static psx_status set_pbar_zero(psx_scft* e) {
  DevState h;
  PSX_TRY(read_state(e, &h));
  h.pbar = 0.0;
  PSX_HIP(hipMemcpyAsync(e->st, &h, sizeof(h), hipMemcpyHostToDevice, e->stream));
  PSX_HIP(hipStreamSynchronize(e->stream));
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_fields(psx_scft* e, const double* wA, const double* wB,
                                          const double* p) {
  if (!e || !wA || !wB) return fail(PSX_ERR_INVALID, "null argument");
  const size_t fb = e->N * sizeof(double);
  PSX_HIP(hipMemcpyAsync(e->wA, wA, fb, hipMemcpyDeviceToDevice, e->stream));
  PSX_HIP(hipMemcpyAsync(e->wB, wB, fb, hipMemcpyDeviceToDevice, e->stream));
  if (p) PSX_HIP(hipMemcpyAsync(e->praw, p, fb, hipMemcpyDeviceToDevice, e->stream));
  else {   // a fill kernel, not a memset node, so a caller may capture the call
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_1d(e->N)), dim3(kThreads), 0, e->stream, e->praw, e->N, 0.0);
    PSX_HIP(hipGetLastError());
  }
  PSX_TRY(set_pbar_zero(e));
  fields_changed(e);
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_fields_host(psx_scft* e, const double* wA, const double* wB,
                                               const double* p) {
  if (!e || !wA || !wB) return fail(PSX_ERR_INVALID, "null argument");
  const size_t fb = e->N * sizeof(double);
  PSX_HIP(hipMemcpyAsync(e->wA, wA, fb, hipMemcpyHostToDevice, e->stream));
  PSX_HIP(hipMemcpyAsync(e->wB, wB, fb, hipMemcpyHostToDevice, e->stream));
  if (p) PSX_HIP(hipMemcpyAsync(e->praw, p, fb, hipMemcpyHostToDevice, e->stream));
  else {   // a fill kernel, not a memset node, so a caller may capture the call
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_1d(e->N)), dim3(kThreads), 0, e->stream, e->praw, e->N, 0.0);
    PSX_HIP(hipGetLastError());
  }
  PSX_TRY(set_pbar_zero(e));
  fields_changed(e);
  return PSX_OK;
}

// Spectral filter updaters on the engine's conjugate fields, in place:
// PsSimpleSpecFilter / PsMultiSpecFilter::update for updatefields naming the
// A (bit 0) and B (bit 1) monomer densities (pseffhamil/PsSimpleSpecFilter.cpp:
// 74-103).  The next iteration sees the filtered fields; densities and F of
// the last iteration stay as they were.
extern "C" psx_status psx_scft_spec_filter(psx_scft* e, int fields, double cutoff_factor,
                                           double filter_strength, const int* spec_cells) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (fields < 0 || fields > 3) return fail(PSX_ERR_INVALID, "fields is a mask of 1 (w_A) and 2 (w_B)");
  if (e->in_batch) return fail(PSX_ERR_STATE, "engine belongs to a batch");
  if (!e->fplan) PSX_TRY(psx_fft_plan_create(e->dims[0], e->dims[1], e->dims[2], PSX_FP64, e->stream, &e->fplan));
  if (fields & 1) PSX_TRY(psx_fft_spec_filter(e->fplan, e->wA, cutoff_factor, filter_strength, spec_cells, 0.0));
  if (fields & 2) PSX_TRY(psx_fft_spec_filter(e->fplan, e->wB, cutoff_factor, filter_strength, spec_cells, 0.0));
  e->sw[0].wfac_valid = false;
  e->sw[1].wfac_valid = false;
  e->swept = false;
  return PSX_OK;
}

extern "C" psx_status psx_rng_create(unsigned long long seed, psx_rng** out) {
  if (!out) return fail(PSX_ERR_INVALID, "null argument");
  psx_rng* r = new psx_rng();
  r->seed((uint32_t)(seed & 0xffffffffull));
  *out = r;
  return PSX_OK;
}

extern "C" psx_status psx_rng_destroy(psx_rng* r) {
  delete r;
  return PSX_OK;
}

extern "C" psx_status psx_rng_uniform(psx_rng* r, long long n, double* h_out) {
  if (!r || (n > 0 && !h_out) || n < 0) return fail(PSX_ERR_INVALID, "bad argument");
  for (long long i = 0; i < n; ++i) h_out[i] = r->uniform();
  return PSX_OK;
}

extern "C" psx_status psx_rng_gaussian(psx_rng* r, long long n, double* h_out) {
  if (!r || (n > 0 && !h_out) || n < 0) return fail(PSX_ERR_INVALID, "bad argument");
  for (long long i = 0; i < n; ++i) h_out[i] = r->gaussian();
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_noise(psx_scft* e, double strength, psx_rng* rng) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (e->in_batch) return fail(PSX_ERR_STATE, "engine belongs to a batch");
  if (!std::isfinite(strength)) return fail(PSX_ERR_INVALID, "noise strength must be finite");
  const bool was = e->noise_rng && e->noise_strength != 0.0;
  const bool now = rng && strength != 0.0;
  if (now && !e->noise) {
    PSX_TRY(dalloc((void**)&e->noise, 2 * (size_t)e->N * sizeof(double)));
    for (int k = 0; k < 2; ++k) {
      PSX_HIP(hipHostMalloc((void**)&e->noise_host[k], 2 * (size_t)e->N * sizeof(double), hipHostMallocDefault));
      PSX_HIP(hipEventCreateWithFlags(&e->noise_ev[k], hipEventDisableTiming));
      PSX_HIP(hipEventRecord(e->noise_ev[k], e->stream));
    }
  }
  if (was != now) drop_graphs(e);
  e->noise_rng = rng;
  e->noise_strength = rng ? strength : 0.0;
  return PSX_OK;
}

static psx_status get_fields_impl(psx_scft* e, double* wA, double* wB, double* p, hipMemcpyKind k) {
  const size_t fb = e->N * sizeof(double);
  if (wA) PSX_HIP(hipMemcpyAsync(wA, e->wA, fb, k, e->stream));
  if (wB) PSX_HIP(hipMemcpyAsync(wB, e->wB, fb, k, e->stream));
  if (p) {
    if (!e->hstage) PSX_TRY(dalloc(&e->hstage, fb));
    hipLaunchKernelGGL(k_pressure, dim3(grid_1d(e->N)), dim3(kThreads), 0, e->stream,
                       (const double*)e->praw, (const DevState*)e->st, e->N, (double*)e->hstage);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipMemcpyAsync(p, e->hstage, fb, k, e->stream));
  }
  PSX_HIP(hipStreamSynchronize(e->stream));
  return PSX_OK;
}

extern "C" psx_status psx_scft_get_fields(psx_scft* e, double* wA, double* wB, double* p) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  return get_fields_impl(e, wA, wB, p, hipMemcpyDeviceToDevice);
}

extern "C" psx_status psx_scft_get_fields_host(psx_scft* e, double* wA, double* wB, double* p) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  return get_fields_impl(e, wA, wB, p, hipMemcpyDeviceToHost);
}

extern "C" psx_status psx_scft_propagate(psx_scft* e) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  PSX_TRY(compute_wfac(e));
  PSX_TRY(sweep(e));
  e->swept = true;
  e->dens_valid = false;
  return PSX_OK;
}

extern "C" psx_status psx_scft_Q(psx_scft* e, double* Q) {
  if (!e || !Q) return fail(PSX_ERR_INVALID, "null argument");
  if (!e->swept) return fail(PSX_ERR_STATE, "propagate before Q");
  PSX_TRY(launch_Q(e));
  DevState h;
  PSX_TRY(read_state(e, &h));
  *Q = h.Q;
  if (h.nanflag || !(h.Q == h.Q) || std::isinf(h.Q)) {
    int z = 0;
    (void)hipMemcpy(&e->st->nanflag, &z, sizeof(int), hipMemcpyHostToDevice);
    return fail(PSX_ERR_NAN, "Q is NaN/Inf (PsBlock::calcBigQ check)");
  }
  return PSX_OK;
}

extern "C" psx_status psx_scft_polymer_Q(psx_scft* e, double* Qs, int* n) {
  if (!e || !Qs) return fail(PSX_ERR_INVALID, "null argument");
  double Q0;
  PSX_TRY(psx_scft_Q(e, &Q0));
  DevState h;
  PSX_TRY(read_state(e, &h));
  for (int p = 0; p < (int)e->polys.size(); ++p) Qs[p] = h.Qp[p];
  if (n) *n = (int)e->polys.size();
  return PSX_OK;
}

extern "C" psx_status psx_scft_densities(psx_scft* e, double* phiA, double* phiB) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (!e->swept) return fail(PSX_ERR_STATE, "propagate before densities");
  double Q;
  PSX_TRY(psx_scft_Q(e, &Q));
  PSX_TRY(launch_quad(e, QM_DENS, nullptr));
  const size_t fb = e->N * sizeof(double);
  if (phiA) PSX_HIP(hipMemcpyAsync(phiA, e->phiA, fb, hipMemcpyDeviceToDevice, e->stream));
  if (phiB) PSX_HIP(hipMemcpyAsync(phiB, e->phiB, fb, hipMemcpyDeviceToDevice, e->stream));
  PSX_HIP(hipStreamSynchronize(e->stream));
  e->dens_valid = true;
  return PSX_OK;
}

extern "C" psx_status psx_scft_get_densities_host(psx_scft* e, double* phiA, double* phiB) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (!e->dens_valid) return fail(PSX_ERR_STATE, "no densities computed yet");
  const size_t fb = e->N * sizeof(double);
  if (phiA) PSX_HIP(hipMemcpyAsync(phiA, e->phiA, fb, hipMemcpyDeviceToHost, e->stream));
  if (phiB) PSX_HIP(hipMemcpyAsync(phiB, e->phiB, fb, hipMemcpyDeviceToHost, e->stream));
  PSX_HIP(hipStreamSynchronize(e->stream));
  return PSX_OK;
}

static psx_status copy_slice_host(psx_scft* e, const void* src, double* out) {
  const size_t fb = e->N * sizeof(double);
  if (!e->hstage) PSX_TRY(dalloc(&e->hstage, fb));
  if (e->prec == PSX_FP64)
    hipLaunchKernelGGL(k_to_f64<double>, dim3(grid_1d(e->N)), dim3(kThreads), 0, e->stream,
                       (const double*)src, e->N, (double*)e->hstage);
  else
    hipLaunchKernelGGL(k_to_f64<float>, dim3(grid_1d(e->N)), dim3(kThreads), 0, e->stream,
                       (const float*)src, e->N, (double*)e->hstage);
  PSX_HIP(hipGetLastError());
  PSX_HIP(hipMemcpyAsync(out, e->hstage, fb, hipMemcpyDeviceToHost, e->stream));
  PSX_HIP(hipStreamSynchronize(e->stream));
  return PSX_OK;
}

extern "C" psx_status psx_scft_get_slice_host(psx_scft* e, int path, int t, double* out) {
  if (!e || !out) return fail(PSX_ERR_INVALID, "null argument");
  if (e->ckpt_K > 0) return fail(PSX_ERR_STATE, "slices are not stored in checkpointed mode");
  if (path < 0 || path >= (int)e->paths.size()) return fail(PSX_ERR_INVALID, "path index out of range");
  if (t < 0 || t > e->paths[path].len) return fail(PSX_ERR_INVALID, "slice index out of range");
  return copy_slice_host(e, slice_ptr(e, e->prec, path, t), out);
}

extern "C" psx_status psx_scft_get_block_slice_host(psx_scft* e, int block, int from_tail, int t,
                                                    double* out) {
  if (!e || !out) return fail(PSX_ERR_INVALID, "null argument");
  if (block < 0 || block >= (int)e->blocks.size()) return fail(PSX_ERR_INVALID, "block index out of range");
  if (from_tail != 0 && from_tail != 1) return fail(PSX_ERR_INVALID, "from_tail must be 0 or 1");
  if (e->ckpt_K > 0) return fail(PSX_ERR_STATE, "slices are not stored in checkpointed mode");
  const BlockInfo& b = e->blocks[block];
  if (t < 0 || t > b.steps) return fail(PSX_ERR_INVALID, "slice index out of range");
  return copy_slice_host(e, slice_ptr(e, e->prec, b.path[from_tail], b.off[from_tail] + t), out);
}

extern "C" psx_status psx_scft_update_steepest(psx_scft* e, double lambda0, double lambda1,
                                               double* F) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (!e->dens_valid) return fail(PSX_ERR_STATE, "densities before the update");
  psx_mixing mix{};
  mix.kind = 0; mix.lambda0 = lambda0; mix.lambda1 = lambda1; mix.f_order = 0;
  hipLaunchKernelGGL(k_state_set_call, dim3(1), dim3(1), 0, e->stream, e->st);
  PSX_TRY(stage_noise(e));
  PSX_TRY(launch_quad(e, QM_SD_PHI, &mix));
  PSX_TRY(launch_finish(e, FIN_ITER_SD, nullptr));
  DevState h;
  PSX_TRY(read_state(e, &h));
  if (F) *F = h.F;
  e->sw[e->prec].wfac_valid = true;
  e->sw[e->prec ^ 1].wfac_valid = false;
  e->swept = false;
  e->dens_valid = false;
  return PSX_OK;
}

extern "C" psx_status psx_scft_free_energy(psx_scft* e, double* F) {
  if (!e || !F) return fail(PSX_ERR_INVALID, "null argument");
  if (!e->dens_valid) return fail(PSX_ERR_STATE, "densities before the free energy");
  PSX_TRY(launch_quad(e, QM_FE, nullptr));
  PSX_TRY(launch_finish(e, FIN_FE, nullptr));
  DevState h;
  PSX_TRY(read_state(e, &h));
  *F = h.F;
  return PSX_OK;
}

extern "C" psx_status psx_scft_reset_history(psx_scft* e) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  hipLaunchKernelGGL(k_state_reset_hist, dim3(1), dim3(1), 0, e->stream, e->st);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

static psx_status prepare_iterate(psx_scft* e, const psx_mixing* mix, int n_iters) {
  if (mix->kind != 0 && mix->kind != 1) return fail(PSX_ERR_INVALID, "mixing kind must be 0 or 1");
  if (mix->kind == 1 && (mix->depth < 0 || mix->depth > kMaxDepth))
    return fail(PSX_ERR_INVALID, "Anderson depth must be in [0, 32]");
  if (mix->kind == 1) PSX_TRY(ensure_ring(e, mix->depth));
  if (n_iters > e->hist_cap) {   // grows rarely: captured graphs hold the pointer
    const int cap = std::max(n_iters, std::max(2 * e->hist_cap, 1024));
    drop_graphs(e);
    dfree((void*&)e->hist);
    PSX_TRY(dalloc((void**)&e->hist, (size_t)cap * 3 * sizeof(double)));
    e->hist_cap = cap;
  }
  if (!e->sw[e->prec].wfac_valid) PSX_TRY(compute_wfac(e));
  hipLaunchKernelGGL(k_state_set_call, dim3(1), dim3(1), 0, e->stream, e->st);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

static GraphKey graph_key(int prec, const psx_mixing* mix) {
  GraphKey key;
  std::memset(&key, 0, sizeof(key));
  key.prec = prec; key.kind = mix->kind; key.depth = mix->kind ? mix->depth : 0;
  key.f_order = mix->f_order; key.lam0 = mix->lambda0; key.lam1 = mix->lambda1;
  key.beta = mix->kind ? mix->beta : 0.0; key.start = mix->kind ? mix->anderson_start : 0.0;
  return key;
}

// Copy the first `it` history rows to the host and update the validity flags.
static psx_status finish_iterate(psx_scft* e, int it, double* hF, double* hQ, double* herr) {
  if (hF || hQ || herr) {
    std::vector<double> hh((size_t)it * 3);
    if (it > 0)
      PSX_HIP(hipMemcpy(hh.data(), e->hist, hh.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < it; ++i) {
      if (hF) hF[i] = hh[3 * i];
      if (hQ) hQ[i] = hh[3 * i + 1];
      if (herr) herr[i] = hh[3 * i + 2];
    }
  }
  e->sw[e->prec].wfac_valid = true;
  e->sw[e->prec ^ 1].wfac_valid = false;
  e->swept = false;
  e->dens_valid = true;   // phi of the last iteration's sweep
  return PSX_OK;
}

extern "C" psx_status psx_scft_iterate(psx_scft* e, const psx_mixing* mix, int n_iters, double tol,
                                       int use_graph, double* hF, double* hQ, double* herr,
                                       int* done) {
  if (!e || !mix) return fail(PSX_ERR_INVALID, "null argument");
  if (done) *done = 0;
  if (n_iters <= 0) return PSX_OK;
  if (mix->kind != 0 && e->noise_rng && e->noise_strength != 0.0)
    return fail(PSX_ERR_INVALID, "noise belongs to the steepest-descent update (mixing kind 0)");
  PSX_TRY(prepare_iterate(e, mix, n_iters));

  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    GraphKey key = graph_key(e->prec, mix);
    auto it = e->graphs.find(key);
    if (it != e->graphs.end()) exec = it->second;
    else {
      hipGraph_t graph;
      PSX_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      psx_status cs = enqueue_iteration(e, mix);
      hipError_t ce = hipStreamEndCapture(e->stream, &graph);
      if (cs != PSX_OK) { if (ce == hipSuccess) (void)hipGraphDestroy(graph); return cs; }
      if (ce != hipSuccess) return fail(PSX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) return fail(PSX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      e->graphs[key] = exec;
    }
  }
  int it = 0;
  DevState h;
  psx_status result = PSX_OK;
  for (; it < n_iters; ++it) {
    if (mix->kind == 0) PSX_TRY(stage_noise(e));
    if (exec) PSX_HIP(hipGraphLaunch(exec, e->stream));
    else PSX_TRY(enqueue_iteration(e, mix));
    if (tol > 0.0) {
      PSX_TRY(read_state(e, &h));
      if (h.nanflag) { result = fail(PSX_ERR_NAN, "Q is NaN/Inf (PsBlock::calcBigQ check)"); ++it; break; }
      if (h.err < tol) { ++it; break; }
    }
  }
  PSX_TRY(read_state(e, &h));
  if (result == PSX_OK && h.nanflag) result = fail(PSX_ERR_NAN, "Q is NaN/Inf (PsBlock::calcBigQ check)");
  if (h.nanflag) {
    int z = 0;
    (void)hipMemcpy(&e->st->nanflag, &z, sizeof(int), hipMemcpyHostToDevice);
  }
  if (done) *done = it;
  PSX_TRY(finish_iterate(e, it, hF, hQ, herr));
  return result;
}

// (Re)write the spectral tables of precision `prec` for the current cell.
static psx_status write_emul(psx_scft* e, int prec) {
  SweepBufs& s = e->sw[prec];
  for (size_t i = 0; i < e->axes.size(); ++i) {
    const Axis& a = e->axes[i];
    for (size_t c = 0; c < e->bval.size(); ++c) {
      const double h = a.n == 1 ? 1.0 : e->cell[a.axis];
      std::vector<double> em = spectral_rows(axis_emul(a, h, e->ds, e->bval[c]), a.n);
      if (prec == PSX_FP64) {
        PSX_HIP(hipMemcpy(s.emul[c][i], em.data(), em.size() * 8, hipMemcpyHostToDevice));
      } else {
        std::vector<float> f = cast_vec<float>(em);
        PSX_HIP(hipMemcpy(s.emul[c][i], f.data(), f.size() * 4, hipMemcpyHostToDevice));
      }
    }
  }
  for (size_t c = 0; c < s.epl.size(); ++c) {   // the plane launches' copies
    const std::vector<double> em = plane_factors(e, c);
    if (prec == PSX_FP64) {
      PSX_HIP(hipMemcpy(s.epl[c], em.data(), em.size() * 8, hipMemcpyHostToDevice));
    } else {
      std::vector<float> f = cast_vec<float>(em);
      PSX_HIP(hipMemcpy(s.epl[c], f.data(), f.size() * 4, hipMemcpyHostToDevice));
    }
  }
  return PSX_OK;
}

static psx_status write_stress_tables(psx_scft* e, bool alloc) {
  const double twopi = 2.0 * std::acos(-1.0);
  if (alloc) e->kf1.assign(e->bval.size(), std::array<double*, 3>{nullptr, nullptr, nullptr});
  for (int d = 0; d < 3; ++d) {
    const int n = e->dims[d];
    std::vector<double> k2(n);
    const double n2 = (double)n / 2.0;
    for (int i = 0; i < n; ++i) {
      const double nk = n2 - std::fabs((double)i - n2);
      const double kv = twopi * nk / (double)n / e->cell[d];
      k2[i] = kv * kv;
    }
    if (alloc) PSX_TRY(upload((void**)&e->kd2[d], k2.data(), n * 8));
    else PSX_HIP(hipMemcpy(e->kd2[d], k2.data(), n * 8, hipMemcpyHostToDevice));
    for (size_t c = 0; c < e->bval.size(); ++c) {
      std::vector<double> kf(n);
      for (int i = 0; i < n; ++i) kf[i] = std::exp(-1.0 * e->ds * k2[i] * e->bval[c] * e->bval[c]);
      if (alloc) PSX_TRY(upload((void**)&e->kf1[c][d], kf.data(), n * 8));
      else PSX_HIP(hipMemcpy(e->kf1[c][d], kf.data(), n * 8, hipMemcpyHostToDevice));
    }
  }
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_cell(psx_scft* e, double dx, double dy, double dz) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (!(dx > 0 && dy > 0 && dz > 0)) return fail(PSX_ERR_INVALID, "cell sizes must be > 0");
  PSX_HIP(hipStreamSynchronize(e->stream));
  e->cell[0] = dx; e->cell[1] = dy; e->cell[2] = dz;
  for (int p = 0; p < 2; ++p)
    if (e->sw[p].ready) PSX_TRY(write_emul(e, p));
  if (e->zbuf) PSX_TRY(write_stress_tables(e, false));
  e->swept = false;
  e->dens_valid = false;
  return PSX_OK;
}

static psx_status ensure_stress(psx_scft* e) {
  if (e->zbuf) return PSX_OK;
  PSX_TRY(dalloc(&e->zbuf, (size_t)e->N * 16));
  PSX_TRY(dalloc((void**)&e->sacc, 4 * sizeof(double)));
  PSX_TRY(make_axes(e->dims, &e->caxes));
  for (auto& a : e->caxes) PSX_TRY(axis_tw(a, PSX_FP64));
  PSX_TRY(write_stress_tables(e, true));
  const int need = e->nb * 3;
  if (need > e->part_cap) {
    drop_graphs(e);   // captured iterations hold the partials' pointer
    dfree((void*&)e->part);
    PSX_TRY(dalloc((void**)&e->part, need * sizeof(double)));
    e->part_cap = need;
  }
  return PSX_OK;
}

// One contour step's contribution to the stress: Z = (W f, W g) packed,
// transformed, and its weighted spectral sum accumulated.
template <typename T>
static psx_status stress_step(psx_scft* e, const T* f, const T* gg, const BlockInfo& blk) {
  SweepBufs& s = e->sw[e->prec];
  const int g = grid_1d(e->N);
  const T* w = (const T*)s.wf[blk.species];
  hipLaunchKernelGGL(k_stress_pack<T>, dim3(g), dim3(kThreads), 0, e->stream, f, gg, w, e->N,
                     (double2*)e->zbuf);
  for (auto& a : e->caxes) {
    CplxIO io{};
    io.in = e->zbuf; io.out = e->zbuf; io.kmul = nullptr;
    io.in_mode = 0; io.out_mode = 0; io.do_fwd = 1; io.do_inv = 0;
    PSX_TRY(launch_cplx<double>(a, PSX_FP64, io, e->stream));
  }
  StressArgs sa;
  sa.z = (const double2*)e->zbuf;
  sa.nx = e->dims[0]; sa.ny = e->dims[1]; sa.nz = e->dims[2];
  for (int d = 0; d < 3; ++d) { sa.kf[d] = e->kf1[blk.bclass][d]; sa.kd2[d] = e->kd2[d]; }
  sa.part = e->part;
  hipLaunchKernelGGL(k_stress_reduce, dim3(e->nb), dim3(kThreads), 0, e->stream, sa);
  hipLaunchKernelGGL(k_stress_accum, dim3(1), dim3(kThreads), 0, e->stream, (const double*)e->part,
                     e->nb, e->ds * blk.b * blk.b, e->sacc);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// The stress accumulators cleared by a fill kernel (no memset node in a
// caller's captured graph).
static psx_status clear_sacc(psx_scft* e) {
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(kThreads), 0, e->stream, e->sacc, 4, 0.0);
  PSX_HIP(hipGetLastError());
  return PSX_OK;
}

// For every contour step t of every block: q_t from the head and the
// propagator from the tail after steps-1-t steps meet across the step
// operator; sum their spectral product weighted by dK/dL.  Checkpointed:
// the same pairs, chunk by chunk (step t of the chain pairs forward slice t
// with backward slice Ns-1-t).
template <typename T>
static psx_status stress_sweep(psx_scft* e, int poly) {
  PSX_TRY(clear_sacc(e));
  if (e->ckpt_K > 0) {
    SweepBufs& s = e->sw[e->prec];
    const size_t sb = (size_t)e->N * sizeof(T);
    const int Ns = e->paths[1].len;
    const PathInfo& fwd = e->paths[0];
    for (size_t ci = 0; ci < e->segs.size(); ++ci) {
      PSX_TRY(ckpt_chunk(e, ci));
      const CkptSeg& c = e->segs[ci];
      for (int t = c.t0; t < c.t1; ++t) {
        const T* f = (const T*)((const char*)s.fch[ci % 2] + (size_t)(t - c.t0) * sb);
        const T* gg = (const T*)((const char*)s.seg + (size_t)(Ns - 1 - t - c.u0) * sb);
        PSX_TRY(stress_step<T>(e, f, gg, e->blocks[fwd.step_block[t]]));
      }
    }
    return PSX_OK;
  }
  for (const BlockInfo& blk : e->blocks) {
    if (blk.poly != poly) continue;
    for (int t = 0; t < blk.steps; ++t) {
      const T* f = (const T*)slice_ptr(e, e->prec, blk.path[0], blk.off[0] + t);
      const T* gg = (const T*)slice_ptr(e, e->prec, blk.path[1], blk.off[1] + blk.steps - 1 - t);
      PSX_TRY(stress_step<T>(e, f, gg, blk));
    }
  }
  return PSX_OK;
}

extern "C" psx_status psx_scft_box_stress(psx_scft* e, double* dFdL) {
  if (!e || !dFdL) return fail(PSX_ERR_INVALID, "null argument");
  if (!e->swept) return fail(PSX_ERR_STATE, "propagate before the box stress");
  PSX_TRY(ensure_stress(e));
  double Qs[PSX_MAX_POLYMERS];
  int np = 0;
  PSX_TRY(psx_scft_polymer_Q(e, Qs, &np));
  // F = -sum_p (volfrac_p / lengthRatio_p) ln Q_p + terms without L at fixed fields
  double sum[3] = {0.0, 0.0, 0.0};
  const double N = (double)e->N;
  for (int p = 0; p < np; ++p) {
    if (e->prec == PSX_FP64) PSX_TRY(stress_sweep<double>(e, p));
    else PSX_TRY(stress_sweep<float>(e, p));
    double acc[3];
    PSX_HIP(hipMemcpyAsync(acc, e->sacc, 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    PSX_HIP(hipStreamSynchronize(e->stream));
    const double c = e->polys[p].vf / e->polys[p].lr;
    for (int d = 0; d < 3; ++d) {
      const double L = e->dims[d] * e->cell[d];
      const double dQ = acc[d] / L / N / N;
      sum[d] += c * (dQ / Qs[p]);
    }
  }
  for (int d = 0; d < 3; ++d) dFdL[d] = e->dims[d] > 1 ? -sum[d] : 0.0;
  return PSX_OK;
}

// ---------------------------------------------------------------------------
// Reverse mode of the forward op (w_A, w_B, box) -> (phi_A, phi_B, Q)
// ---------------------------------------------------------------------------
static bool oriented_chain(const psx_scft* e) {
  bool chain = e->rounds.size() == 1 && e->paths.size() == 2 && e->paths[0].init.empty() &&
               e->blocks[0].cnt[0].empty() && e->polys.size() == 1;
  for (const BlockInfo& b : e->blocks) chain = chain && b.path[0] == 0 && b.path[1] == 1;
  return chain;
}

// Quadrature weights per contour point t by species: point t of block b at
// index j = t - off_b gets the block's weight j (the junction point gets one
// from each block).
static void point_weights(const psx_scft* e, std::vector<double>* cA, std::vector<double>* cB) {
  const int Ns = e->paths[0].len;
  cA->assign(Ns + 1, 0.0);
  cB->assign(Ns + 1, 0.0);
  const std::vector<double> w = all_weights(e, e->quad_mode);
  size_t woff = 0;
  for (const BlockInfo& b : e->blocks) {
    for (int j = 0; j <= b.steps; ++j) (b.species == 0 ? *cA : *cB)[b.off[0] + j] += w[woff + j];
    woff += b.steps + 1;
  }
}

static double host_dot(psx_scft* e, const double* x, const double* y) {
  hipLaunchKernelGGL(k_partial_dot<double>, dim3(e->nb), dim3(kThreads), 0, e->stream, x, y, e->N, e->part);
  std::vector<double> h(e->nb);
  if (hipMemcpyAsync(h.data(), e->part, e->nb * sizeof(double), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess)
    return std::nan("");
  double v = 0.0;
  for (double x2 : h) v += x2;
  return v;
}

extern "C" psx_status psx_scft_backward(psx_scft* e, const double* gA, const double* gB, double gQ,
                                        double* gwA, double* gwB, double* gL) {
  if (!e || !gwA || !gwB) return fail(PSX_ERR_INVALID, "null argument");
  if (!e->swept) return fail(PSX_ERR_STATE, "propagate before the backward pass");
  if (e->prec != PSX_FP64) return fail(PSX_ERR_STATE, "the backward pass runs on FP64 sweeps");
  if (e->ckpt_K > 0) return fail(PSX_ERR_STATE, "the backward pass reads every stored slice (checkpointing on)");
  if (!oriented_chain(e))
    return fail(PSX_ERR_INVALID, "the backward pass needs one linear chain oriented head to tail from block 0");
  const long long N = e->N;
  const int Ns = e->paths[0].len;
  const size_t fb = (size_t)N * sizeof(double);
  const size_t rows = (size_t)2 * Ns * e->axes.size();
  if (!e->adj) PSX_TRY(dalloc((void**)&e->adj, 7 * fb));
  if (e->adj_tab_rows < rows) {
    dfree((void*&)e->adj_tab);
    PSX_TRY(dalloc((void**)&e->adj_tab, rows * sizeof(PairIO)));
    e->adj_tab_rows = rows;
  }
  double* lam[2] = {e->adj, e->adj + N};
  double* u = e->adj + 2 * N;
  double* grad[2] = {e->adj + 3 * N, e->adj + 4 * N};
  double* qg[2] = {e->adj + 5 * N, e->adj + 6 * N};
  SweepBufs& s = e->sw[PSX_FP64];
  auto F = [&](int t) { return (const double*)slice_ptr(e, PSX_FP64, 0, t); };
  auto G = [&](int t) { return (const double*)slice_ptr(e, PSX_FP64, 1, t); };
  double Q;
  PSX_TRY(psx_scft_Q(e, &Q));
  if (!e->dens_valid) PSX_TRY(psx_scft_densities(e, nullptr, nullptr));
  std::vector<double> cA, cB;
  point_weights(e, &cA, &cB);
  // line passes of both adjoint sweeps: step k of sweep d applies the step
  // operator of contour step s (d = 0: s = Ns-1-k, from the tail; d = 1:
  // s = k, in the backward sweep's order), the first pass times W, the last
  // pass times W
  const int na = (int)e->axes.size();
  std::vector<PairIO> tab(rows);
  for (int d = 0; d < 2; ++d)
    for (int k = 0; k < Ns; ++k) {
      const int sstep = d == 0 ? Ns - 1 - k : k;
      const BlockInfo& blk = e->blocks[e->paths[0].step_block[sstep]];
      for (int a = 0; a < na; ++a) {
        PairIO& r = tab[((size_t)d * Ns + k) * na + a];
        std::memset(&r, 0, sizeof(r));
        r.in = a == 0 ? (const void*)lam[k % 2] : (const void*)u;
        r.out = u;
        r.pre = a == 0 ? s.wf[blk.species] : nullptr;
        r.post = a == na - 1 ? s.wf[blk.species] : nullptr;
        r.emul = s.emul[blk.bclass][a];
      }
    }
  PSX_HIP(hipMemcpyAsync(e->adj_tab, tab.data(), rows * sizeof(PairIO), hipMemcpyHostToDevice, e->stream));
  const int g = grid_1d(N);
  const double halfds = 0.5 * e->ds;
  const bool box = gL != nullptr;
  double accJ[3] = {0.0, 0.0, 0.0};
  if (box) PSX_TRY(ensure_stress(e));
  hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kThreads), 0, e->stream, e->adj + 3 * N, 4 * N, 0.0);
  for (int d = 0; d < 2; ++d) {
    // d = 0: adjoint of the forward sweep, from lam_Ns = h_Ns G_0 down to t = 0;
    // d = 1: adjoint of the backward sweep, from mu_Ns = h_0 F_0
    if (box) PSX_TRY(clear_sacc(e));
    const int t0 = d == 0 ? Ns : 0;
    hipLaunchKernelGGL(k_adj_init, dim3(g), dim3(kThreads), 0, e->stream, d == 0 ? G(0) : F(0), gA, gB,
                       cA[t0], cB[t0], N, lam[0]);
    for (int k = 0; k < Ns; ++k) {
      const int sstep = d == 0 ? Ns - 1 - k : k;
      const BlockInfo& blk = e->blocks[e->paths[0].step_block[sstep]];
      for (int a = 0; a < na; ++a)
        PSX_TRY(launch_pair<double>(e->axes[a], PSX_FP64, e->adj_tab + ((size_t)d * Ns + k) * na + a, 1,
                                    e->stream, 0));
      const int sp = blk.species;
      if (d == 0) {
        // forward step s = Ns-1-k: lam_s = T lam_{s+1} + h_s G_{Ns-s}
        const int st = Ns - 1 - k;
        hipLaunchKernelGGL(k_adj_step, dim3(g), dim3(kThreads), 0, e->stream, u, lam[k % 2], F(st), F(st + 1),
                           G(Ns - st), gA, gB, cA[st], cB[st], halfds, N, lam[(k + 1) % 2], grad[sp],
                           G(Ns - 1 - st), G(Ns - st), qg[sp]);
        if (box) PSX_TRY(stress_step<double>(e, F(st), lam[k % 2], blk));
      } else {
        // backward step u = k (the operator of forward step Ns-1-k):
        // mu_u = T mu_{u+1} + h_{Ns-u} F_{Ns-u}
        const int uu = Ns - 1 - k;
        hipLaunchKernelGGL(k_adj_step, dim3(g), dim3(kThreads), 0, e->stream, u, lam[k % 2], G(uu), G(uu + 1),
                           F(Ns - uu), gA, gB, cA[Ns - uu], cB[Ns - uu], halfds, N, lam[(k + 1) % 2],
                           grad[sp], nullptr, nullptr, nullptr);
        if (box) PSX_TRY(stress_step<double>(e, G(uu), lam[k % 2], blk));
      }
      PSX_HIP(hipGetLastError());
    }
    if (box) {
      double acc[3];
      PSX_HIP(hipMemcpyAsync(acc, e->sacc, 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
      PSX_HIP(hipStreamSynchronize(e->stream));
      for (int i = 0; i < 3; ++i) accJ[i] += acc[i];
    }
  }
  // L = J / Q + gQ Q with J = Q <g, phi>
  double gphi = 0.0;
  if (gA) gphi += host_dot(e, gA, e->phiA);
  if (gB) gphi += host_dot(e, gB, e->phiB);
  if (!(gphi == gphi)) return fail(PSX_ERR_HIP, "reduction failed");
  // dL/dw = (1/Q) dJ/dw + (gQ - <g, phi> / Q) dQ/dw, dQ/dw = -(ds / 2N) qgrad
  const double coefQ = (gQ - gphi / Q) * (-halfds / (double)N);
  hipLaunchKernelGGL(k_adj_combine, dim3(g), dim3(kThreads), 0, e->stream, grad[0], qg[0], 1.0 / Q, coefQ, N, gwA);
  hipLaunchKernelGGL(k_adj_combine, dim3(g), dim3(kThreads), 0, e->stream, grad[1], qg[1], 1.0 / Q, coefQ, N, gwB);
  PSX_HIP(hipGetLastError());
  if (box) {
    // dQ/dL from the pairs (F_t, G_{Ns-1-t}), as the box stress
    PSX_TRY(stress_sweep<double>(e, 0));
    double accQ[3];
    PSX_HIP(hipMemcpyAsync(accQ, e->sacc, 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    PSX_HIP(hipStreamSynchronize(e->stream));
    for (int dd = 0; dd < 3; ++dd) {
      const double L = e->dims[dd] * e->cell[dd];
      const double dJ = accJ[dd] / L / (double)N;
      const double dQ = accQ[dd] / L / (double)N / (double)N;
      gL[dd] = e->dims[dd] > 1 ? dJ / Q + (gQ - gphi / Q) * dQ : 0.0;
    }
  }
  PSX_HIP(hipStreamSynchronize(e->stream));
  return PSX_OK;
}

extern "C" psx_status psx_scft_backward_host(psx_scft* e, const double* h_gA, const double* h_gB, double gQ,
                                             double* h_gwA, double* h_gwB, double* h_gL) {
  if (!e || !h_gwA || !h_gwB) return fail(PSX_ERR_INVALID, "null argument");
  const size_t fb = (size_t)e->N * sizeof(double);
  double* buf = nullptr;
  PSX_TRY(dalloc((void**)&buf, 4 * fb));
  struct Free { double*& p; ~Free() { dfree((void*&)p); } } fr{buf};
  const long long N = e->N;
  if (h_gA) PSX_HIP(hipMemcpy(buf, h_gA, fb, hipMemcpyHostToDevice));
  if (h_gB) PSX_HIP(hipMemcpy(buf + N, h_gB, fb, hipMemcpyHostToDevice));
  PSX_TRY(psx_scft_backward(e, h_gA ? buf : nullptr, h_gB ? buf + N : nullptr, gQ, buf + 2 * N, buf + 3 * N,
                            h_gL));
  PSX_HIP(hipMemcpy(h_gwA, buf + 2 * N, fb, hipMemcpyDeviceToHost));
  PSX_HIP(hipMemcpy(h_gwB, buf + 3 * N, fb, hipMemcpyDeviceToHost));
  return PSX_OK;
}

extern "C" psx_status psx_scft_param_grad(psx_scft* e, double* dFdchiN, double* dFdfA) {
  if (!e) return fail(PSX_ERR_INVALID, "null argument");
  if (!e->dens_valid) return fail(PSX_ERR_STATE, "densities before the parameter gradient");
  if (!e->swept) return fail(PSX_ERR_STATE, "propagate before the parameter gradient");
  if (dFdfA && !e->diblock)
    return fail(PSX_ERR_INVALID, "dF/dfA is defined for the AB diblock only");
  if (e->ckpt_K > 0) return fail(PSX_ERR_STATE, "the junction slices are not stored in checkpointed mode");
  double Q;
  PSX_TRY(psx_scft_Q(e, &Q));
  // junction of the diblock: q of B's head and q+ arriving there
  const BlockInfo& bb = e->blocks.back();
  const void* fj = slice_ptr(e, e->prec, bb.path[0], bb.off[0]);
  const void* gj = slice_ptr(e, e->prec, bb.path[1], bb.off[1] + bb.steps);
  if (e->prec == PSX_FP64)
    hipLaunchKernelGGL(k_param_grad<double>, dim3(e->nb), dim3(kThreads), 0, e->stream,
                       (const double*)e->phiA, (const double*)e->phiB, (const double*)fj,
                       (const double*)gj, (const double*)e->wA, (const double*)e->wB, e->N, e->part);
  else
    hipLaunchKernelGGL(k_param_grad<float>, dim3(e->nb), dim3(kThreads), 0, e->stream,
                       (const double*)e->phiA, (const double*)e->phiB, (const float*)fj,
                       (const float*)gj, (const double*)e->wA, (const double*)e->wB, e->N, e->part);
  PSX_HIP(hipGetLastError());
  std::vector<double> h((size_t)e->nb * 2);
  PSX_HIP(hipMemcpyAsync(h.data(), e->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  PSX_HIP(hipStreamSynchronize(e->stream));
  double pp = 0.0, jt = 0.0;
  for (int b = 0; b < e->nb; ++b) { pp += h[2 * b]; jt += h[2 * b + 1]; }
  const double N = (double)e->N;
  if (dFdchiN) *dFdchiN = pp / N - e->fA * e->fB;
  // Along the fixed points of the field update, the explicit
  // -chiN (fB - fA) of the disorder term cancels against the density terms
  // chiN fB d<phiA>/dfA + chiN fA d<phiB>/dfA (the update shifts densities
  // by fA, fB), leaving the junction term of -ln Q.
  if (dFdfA) *dFdfA = -(jt / N) / Q;
  return PSX_OK;
}

// The launch of path pi that stands for axis a in tuning: a fused one if the
// path has one.
static size_t representative_launch(const PathInfo& pi, int a) {
  size_t best = 0;
  bool found = false;
  for (size_t j = 0; j < pi.launch_axis.size(); ++j) {
    if (pi.launch_axis[j] != a) continue;
    if (!found || (pi.launch_nops[j] == 2 && pi.launch_nops[best] != 2)) { best = j; found = true; }
  }
  return best;
}

extern "C" psx_status psx_scft_time_step(psx_scft* e, int reps, double* ms_step, double* ms_pass) {
  if (!e || reps < 1) return fail(PSX_ERR_INVALID, "bad argument");
  if (e->ckpt_K > 0) return fail(PSX_ERR_STATE, "timing runs on the stored-slice sweep (checkpointing off)");
  if (!e->sw[e->prec].wfac_valid) PSX_TRY(compute_wfac(e));
  hipEvent_t a, b;
  PSX_HIP(hipEventCreate(&a));
  PSX_HIP(hipEventCreate(&b));
  // reps whole sweeps; per-step figures divide by reps * Ns
  const int na = (int)e->axes.size();
  auto run = [&](int axis) -> psx_status {
    for (int r = 0; r < reps; ++r) {
      if (e->prec == PSX_FP64) PSX_TRY(sweep_range<double>(e, axis));
      else PSX_TRY(sweep_range<float>(e, axis));
    }
    return PSX_OK;
  };
  PSX_TRY(run(-1));  // warm
  PSX_HIP(hipEventRecord(a, e->stream));
  PSX_TRY(run(-1));
  PSX_HIP(hipEventRecord(b, e->stream));
  PSX_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  PSX_HIP(hipEventElapsedTime(&ms, a, b));
  const double per = 1.0 / ((double)reps * e->Ns);
  if (ms_step) *ms_step = ms * per;
  for (int i = 0; i < 3; ++i) {
    double v = 0.0;
    // plane mode: pass 0 times the plane launches (z and y), pass 1 is 0
    if (i < na && !(e->plane && i == 1)) {
      PSX_HIP(hipEventRecord(a, e->stream));
      PSX_TRY(run(e->plane && i == 0 ? kPlaneAxis : i));
      PSX_HIP(hipEventRecord(b, e->stream));
      PSX_HIP(hipEventSynchronize(b));
      PSX_HIP(hipEventElapsedTime(&ms, a, b));
      v = ms * per;
    }
    if (ms_pass) ms_pass[i] = v;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  e->swept = false;
  return PSX_OK;
}

// ===========================================================================
// Line-kernel autotuner and batched parameter sweeps
// ===========================================================================
namespace {

// Candidate line-kernel variants for length n: -1 default (n = 128: lane
// pairs 8x16; n = 96: lane pairs 12x8; FP32 strided passes at n = 64, 96,
// 128: 128-byte tiles), 1 LDS Stockham, 2 alternative register split
// (n = 64: 16x4, n = 96: 24x4, n = 128: 32x4),
// 3 second alternative (n = 32: lane pairs 4x8, n = 128: 16x8), 4 tiles on
// strided passes (n = 64, 96, 128), 5 lane pairs 8x16 (n = 128).
std::vector<int> variant_candidates(int n) {
  std::vector<int> v{-1, 1};
  if (n == 64 || n == 96 || n == 128) v.push_back(2);
  if (n == 32 || n == 128) v.push_back(3);
  if (n == 64 || n == 96 || n == 128) v.push_back(4);
  if (n == 128) v.push_back(5);
  return v;
}

// Time `reps` launches of pass `i` over the field table `fio` (nfields) for
// each candidate; return the fastest variant.
template <typename T>
psx_status tune_pass(const Axis& a, int prec, const PairIO* fio, int nfields, hipStream_t s,
                     int reps, int* best) {
  hipEvent_t e0, e1;
  PSX_HIP(hipEventCreate(&e0));
  PSX_HIP(hipEventCreate(&e1));
  float best_ms = 1e30f;
  *best = -1;
  for (int v : variant_candidates(a.n)) {
    PSX_TRY(launch_pair<T>(a, prec, fio, nfields, s, v));   // warm
    PSX_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) PSX_TRY(launch_pair<T>(a, prec, fio, nfields, s, v));
    PSX_HIP(hipEventRecord(e1, s));
    PSX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PSX_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) { best_ms = ms; *best = v; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return PSX_OK;
}

struct BatchClass {
  int dims[3];
  std::vector<int> members;         // engine indices
  PairIO* fio[2] = {nullptr, nullptr};   // [t][pass][2 * members] per precision
  int variant[2][3] = {{-1, -1, -1}, {-1, -1, -1}};
};

}  // namespace

extern "C" psx_status psx_scft_set_variant(psx_scft* e, int pass, int variant) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (pass < 0 || pass >= (int)e->axes.size()) return fail(PSX_ERR_INVALID, "pass index out of range");
  if (variant != -1 && (variant < 1 || variant > 5))
    return fail(PSX_ERR_INVALID, "line-kernel variant must be -1 or 1..5");
  e->variant[e->prec][pass] = variant;
  drop_graphs(e);
  e->swept = false;
  return PSX_OK;
}

// Switch between line launches and plane launches: the launch lists and
// field tables are rebuilt (sweep buffers reallocated, captured graphs
// dropped); fields are kept, the propagators must be recomputed.
static psx_status apply_plane(psx_scft* e, bool on) {
  if (on == e->plane) return PSX_OK;
  if (on && !plane_available(e))
    return fail(PSX_ERR_INVALID, "plane launches need a 3D grid with ny == nz in {32, 48, 64, 96, 128}");
  PSX_HIP(hipStreamSynchronize(e->stream));
  drop_graphs(e);
  for (auto& sw : e->sw) free_sweep(sw);
  e->plane = on;
  plan_launches(e);
  if (e->ckpt_K > 0) ckpt_plan(e);
  plan_quad_overlap(e);
  PSX_TRY(ensure_sweep(e, e->prec));
  e->swept = false;
  e->dens_valid = false;
  return PSX_OK;
}

extern "C" psx_status psx_scft_set_sweep_mode(psx_scft* e, int mode) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (mode != 0 && mode != 1) return fail(PSX_ERR_INVALID, "sweep mode must be 0 (lines) or 1 (planes)");
  if (e->in_batch && (mode == 1) != e->plane)
    return fail(PSX_ERR_STATE, "engine belongs to a batch (use psx_batch_set_sweep_mode)");
  return apply_plane(e, mode == 1);
}

extern "C" psx_status psx_scft_sweep_mode(const psx_scft* e, int* mode, int* available) {
  if (!e) return fail(PSX_ERR_INVALID, "null engine");
  if (mode) *mode = e->plane ? 1 : 0;
  if (available) *available = plane_available(e) ? 1 : 0;
  return PSX_OK;
}

// Per pass, time whole sweeps (both sweeps on their streams, as an iteration
// runs them) for each candidate line kernel and keep the fastest; passes are
// tuned one after another, the others holding their current choice.  Where
// plane launches are available, whole sweeps with them and with the tuned
// line launches are then timed and the faster mode is kept.
extern "C" psx_status psx_scft_autotune(psx_scft* e, int reps, int* h_choice) {
  if (!e || reps < 1) return fail(PSX_ERR_INVALID, "bad argument");
  if (e->ckpt_K > 0) return fail(PSX_ERR_STATE, "autotune runs on the stored-slice sweep (checkpointing off)");
  const bool modes = plane_available(e) && !e->in_batch;
  if (modes) PSX_TRY(apply_plane(e, false));     // line variants are tuned with line launches
  if (!e->sw[e->prec].wfac_valid) PSX_TRY(compute_wfac(e));
  const int na = (int)e->axes.size();
  hipEvent_t a, b;
  PSX_HIP(hipEventCreate(&a));
  PSX_HIP(hipEventCreate(&b));
  psx_status st = PSX_OK;
  auto time_sweeps = [&](float* ms) -> psx_status {
    PSX_TRY(sweep(e));                               // warm
    PSX_HIP(hipEventRecord(a, e->stream));
    for (int r = 0; r < reps; ++r) PSX_TRY(sweep(e));
    PSX_HIP(hipEventRecord(b, e->stream));
    PSX_HIP(hipEventSynchronize(b));
    PSX_HIP(hipEventElapsedTime(ms, a, b));
    return PSX_OK;
  };
  for (int i = 0; i < na && st == PSX_OK; ++i) {
    if (e->plane && i <= 1) continue;                // inside the plane launches
    int best = e->variant[e->prec][i];
    float best_ms = 1e30f;
    for (int v : variant_candidates(e->axes[i].n)) {
      e->variant[e->prec][i] = v;
      float ms = 0.f;
      st = time_sweeps(&ms);
      if (st != PSX_OK) break;
      if (ms < best_ms) { best_ms = ms; best = v; }
    }
    e->variant[e->prec][i] = best;
  }
  if (st == PSX_OK && modes) {
    float ms_lines = 0.f, ms_plane = 0.f;
    st = time_sweeps(&ms_lines);
    if (st == PSX_OK) st = apply_plane(e, true);
    if (st == PSX_OK) st = compute_wfac(e);
    if (st == PSX_OK) st = time_sweeps(&ms_plane);
    if (st == PSX_OK && ms_lines < ms_plane) st = apply_plane(e, false);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  PSX_TRY(st);
  if (h_choice)
    for (int i = 0; i < 3; ++i) h_choice[i] = i < na ? e->variant[e->prec][i] : -1;
  drop_graphs(e);
  e->swept = false;
  return PSX_OK;
}

// A batch runs on kBatchStreams streams (the caller's and three of its
// own): each size class sweeps on one of them and each instance's
// post-sweep work (Q, quadrature, Anderson, mixing) on one of them after its
// class's sweep, so the small classes' latency-bound launches overlap the
// large classes' HBM-bound ones.  The placement is a list schedule over
// estimated costs, fixed at creation; results do not depend on it.
constexpr int kBatchStreams = 4;

// Size classes per group: a batch of more grid sizes runs its classes as
// consecutive groups of at most this many, each group forked from and
// joined back to the batch stream and captured as a graph of its own.  Every
// captured graph then holds at most four class chains, the form the batch
// runs at its 64-instance workload.
constexpr int kGroupClasses = 4;

struct BatchGroup {
  std::vector<int> cls;                 // class indices
  std::vector<int> post;                // the classes' instances in post-sweep enqueue order
};

struct psx_batch {
  std::vector<psx_scft*> eng;
  std::vector<BatchClass> cls;
  std::vector<int> inst_cls;            // per instance: its class
  std::vector<BatchGroup> groups;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int prec = PSX_FP64;
  int Ns = 0;
  std::map<std::pair<GraphKey, int>, hipGraphExec_t> graphs;   // per (mixing, group)
  hipStream_t ss[kBatchStreams] = {};   // ss[0] = stream
  std::vector<long long> graph_gen;     // engines' graph_gen when the graphs were captured
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kBatchStreams] = {};
  std::vector<hipEvent_t> ev_cls;       // per class: its sweep is done
  std::vector<int> cls_stream;          // per class
  std::vector<int> inst_stream;         // per instance
};

// Estimated device time (us) of a class's sweep and of one instance's
// post-sweep work: HBM bytes at 4 TB/s plus 4 us per launch.
static double batch_sweep_cost(const psx_batch* b, const BatchClass& c) {
  const psx_scft* e0 = b->eng[c.members[0]];
  const double cells = (double)e0->N * c.members.size();
  const double pb = prec_bytes(b->prec);
  return b->Ns * (12.0 * pb * cells / 4.0e6) + 4.0 * e0->paths[0].launch_axis.size();
}

static double batch_post_cost(const psx_batch* b, const psx_scft* e) {
  const double pb = prec_bytes(b->prec);
  return (double)e->N * (2.0 * (b->Ns + 1) * pb + 200.0) / 4.0e6 + 9 * 4.0;
}

// The list schedule of each group on its own (the groups run one after
// another).
static void batch_schedule(psx_batch* b) {
  const int nc = (int)b->cls.size();
  std::vector<double> done(nc, 0.0);
  b->cls_stream.assign(nc, 0);
  b->inst_stream.assign(b->eng.size(), 0);
  for (BatchGroup& grp : b->groups) {
    double ready[kBatchStreams] = {};
    std::vector<int> order = grp.cls;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return batch_sweep_cost(b, b->cls[x]) > batch_sweep_cost(b, b->cls[y]);
    });
    for (int c : order) {
      int k = 0;
      for (int j = 1; j < kBatchStreams; ++j) if (ready[j] < ready[k]) k = j;
      b->cls_stream[c] = k;
      ready[k] += batch_sweep_cost(b, b->cls[c]);
      done[c] = ready[k];
    }
    grp.post.clear();
    for (size_t i = 0; i < b->eng.size(); ++i)
      if (std::find(grp.cls.begin(), grp.cls.end(), b->inst_cls[i]) != grp.cls.end()) grp.post.push_back((int)i);
    std::stable_sort(grp.post.begin(), grp.post.end(), [&](int x, int y) {
      return done[b->inst_cls[x]] < done[b->inst_cls[y]];
    });
    for (int i : grp.post) {
      int k = 0;
      double best = 1e300;
      for (int j = 0; j < kBatchStreams; ++j) {
        const double start = std::max(ready[j], done[b->inst_cls[i]]);
        if (start < best) { best = start; k = j; }
      }
      b->inst_stream[i] = k;
      ready[k] = best + batch_post_cost(b, b->eng[i]);
    }
  }
}

static void batch_drop_graphs(psx_batch* b) {
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}

static psx_status batch_tables(psx_batch* b, int prec) {
  for (auto& c : b->cls) {
    if (c.fio[prec]) continue;
    psx_scft* e0 = b->eng[c.members[0]];
    const int na = (int)e0->axes.size();
    const int K = (int)c.members.size();
    // members are diblocks: paths 0 and 1 (forward, backward) of equal length
    std::vector<std::vector<PairIO>> per;
    for (int m : c.members)
      for (int h = 0; h < 2; ++h) per.push_back(path_fio(b->eng[m], prec, h));
    (void)na;
    const size_t nl = per[0].size();
    std::vector<PairIO> tab(nl * 2 * K);
    for (size_t j = 0; j < nl; ++j)
      for (int k = 0; k < K; ++k)
        for (int h = 0; h < 2; ++h) tab[(j * K + k) * 2 + h] = per[2 * k + h][j];
    PSX_TRY(upload((void**)&c.fio[prec], tab.data(), tab.size() * sizeof(PairIO)));
  }
  return PSX_OK;
}

// Plane launches by default where they measured faster than the line
// launches: lengths below 96, or few fields (one engine's two sweeps).  At
// 96^3 with 38 fields (19 instances) the line launches move 5.2 TB/s and a
// plane launch, three blocks of six waves per CU, 2.0 TB/s.
static bool plane_default(int n, int nfields) { return n < 96 || nfields <= 8; }

// Line or plane launches for every member of class ci; the class tables
// are rebuilt for the batch's precision (the other one on its next switch).
static psx_status batch_class_mode(psx_batch* b, size_t ci, bool on) {
  BatchClass& c = b->cls[ci];
  for (auto s : b->ss) if (s) PSX_HIP(hipStreamSynchronize(s));
  for (int m : c.members) PSX_TRY(apply_plane(b->eng[m], on));
  dfree((void*&)c.fio[0]);
  dfree((void*&)c.fio[1]);
  batch_drop_graphs(b);
  PSX_TRY(batch_tables(b, b->prec));
  return PSX_OK;
}

extern "C" psx_status psx_batch_destroy(psx_batch* b) {
  if (!b) return PSX_OK;
  for (auto s : b->ss) if (s) (void)hipStreamSynchronize(s);
  batch_drop_graphs(b);
  for (int k = 1; k < kBatchStreams; ++k) if (b->ss[k]) (void)hipStreamDestroy(b->ss[k]);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  for (auto ev : b->ev_join) if (ev) (void)hipEventDestroy(ev);
  for (auto ev : b->ev_cls) if (ev) (void)hipEventDestroy(ev);
  for (auto& c : b->cls) { dfree((void*&)c.fio[0]); dfree((void*&)c.fio[1]); }
  for (auto* e : b->eng) psx_scft_destroy(e);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PSX_OK;
}

extern "C" psx_status psx_batch_create(int n, const psx_grid* grids, const psx_diblock* chains,
                                       psx_prec prec, void* stream, psx_batch** out) {
  if (!out || !grids || !chains || n < 1) return fail(PSX_ERR_INVALID, "bad argument");
  *out = nullptr;
  psx_batch* b = new psx_batch();
  psx_status st = make_stream(stream, &b->stream, &b->own_stream);
  b->prec = prec;
  for (int i = 0; i < n && st == PSX_OK; ++i) {
    psx_scft* e = nullptr;
    st = psx_scft_create(&grids[i], &chains[i], prec, b->stream, &e);
    if (st != PSX_OK) break;
    e->in_batch = true;
    b->eng.push_back(e);
    if (i == 0) b->Ns = e->Ns;
    else if (e->Ns != b->Ns)
      st = fail(PSX_ERR_INVALID, "all batch instances need the same number of contour steps");
  }
  if (st == PSX_OK) {
    for (int i = 0; i < n; ++i) {
      const psx_scft* e = b->eng[i];
      BatchClass* found = nullptr;
      for (auto& c : b->cls)
        if (c.dims[0] == e->dims[0] && c.dims[1] == e->dims[1] && c.dims[2] == e->dims[2]) found = &c;
      if (!found) {
        b->cls.emplace_back();
        found = &b->cls.back();
        for (int d = 0; d < 3; ++d) found->dims[d] = e->dims[d];
      }
      found->members.push_back(i);
    }
    b->inst_cls.assign(n, 0);
    for (size_t c = 0; c < b->cls.size(); ++c) {
      for (int m : b->cls[c].members) b->inst_cls[m] = (int)c;
      if (c % kGroupClasses == 0) b->groups.emplace_back();
      b->groups.back().cls.push_back((int)c);
    }
    for (auto& c : b->cls) {
      psx_scft* e0 = b->eng[c.members[0]];
      const bool on = plane_available(e0) && plane_default(e0->axes[0].n, 2 * (int)c.members.size());
      for (int m : c.members)
        if (st == PSX_OK) st = apply_plane(b->eng[m], on);
    }
    if (st == PSX_OK) st = batch_tables(b, prec);
  }
  if (st == PSX_OK) {
    b->ss[0] = b->stream;
    for (int k = 1; k < kBatchStreams && st == PSX_OK; ++k)
      if (hipStreamCreateWithFlags(&b->ss[k], hipStreamNonBlocking) != hipSuccess)
        st = fail(PSX_ERR_HIP, "hipStreamCreate");
    auto mk = [&](hipEvent_t* ev) {
      if (st == PSX_OK && hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess)
        st = fail(PSX_ERR_HIP, "hipEventCreate");
    };
    mk(&b->ev_fork);
    for (auto& ev : b->ev_join) mk(&ev);
    b->ev_cls.assign(b->cls.size(), nullptr);
    for (auto& ev : b->ev_cls) mk(&ev);
    if (st == PSX_OK) batch_schedule(b);
  }
  if (st != PSX_OK) { psx_batch_destroy(b); return st; }
  *out = b;
  return PSX_OK;
}

extern "C" psx_status psx_batch_size(const psx_batch* b, int* n, int* nclasses) {
  if (!b) return fail(PSX_ERR_INVALID, "null batch");
  if (n) *n = (int)b->eng.size();
  if (nclasses) *nclasses = (int)b->cls.size();
  return PSX_OK;
}

extern "C" psx_status psx_batch_instance(psx_batch* b, int i, psx_scft** out) {
  if (!b || !out || i < 0 || i >= (int)b->eng.size()) return fail(PSX_ERR_INVALID, "bad instance");
  *out = b->eng[i];
  return PSX_OK;
}

extern "C" psx_status psx_batch_set_precision(psx_batch* b, psx_prec prec) {
  if (!b) return fail(PSX_ERR_INVALID, "null batch");
  if (prec != PSX_FP64 && prec != PSX_FP32) return fail(PSX_ERR_INVALID, "bad precision");
  for (auto* e : b->eng) PSX_TRY(set_precision_impl(e, prec));
  PSX_TRY(batch_tables(b, prec));
  b->prec = prec;
  batch_schedule(b);
  return PSX_OK;
}

extern "C" psx_status psx_batch_set_sweep_mode(psx_batch* b, int cls, int mode) {
  if (!b || cls < 0 || cls >= (int)b->cls.size()) return fail(PSX_ERR_INVALID, "bad batch or class");
  if (mode != 0 && mode != 1) return fail(PSX_ERR_INVALID, "sweep mode must be 0 (lines) or 1 (planes)");
  psx_scft* e0 = b->eng[b->cls[cls].members[0]];
  if ((mode == 1) == e0->plane) return PSX_OK;
  if (mode == 1 && !plane_available(e0))
    return fail(PSX_ERR_INVALID, "plane launches need a 3D grid with ny == nz in {32, 48, 64, 96, 128}");
  PSX_TRY(batch_class_mode(b, (size_t)cls, mode == 1));
  batch_schedule(b);
  return PSX_OK;
}

extern "C" psx_status psx_batch_sweep_mode(const psx_batch* b, int cls, int* mode) {
  if (!b || cls < 0 || cls >= (int)b->cls.size() || !mode) return fail(PSX_ERR_INVALID, "bad argument");
  *mode = b->eng[b->cls[cls].members[0]]->plane ? 1 : 0;
  return PSX_OK;
}

static psx_status batch_time_class(psx_batch* b, size_t ci, int reps, float* ms);

extern "C" psx_status psx_batch_autotune(psx_batch* b, int reps, int* h_choice) {
  if (!b || reps < 1) return fail(PSX_ERR_INVALID, "bad argument");
  for (size_t ci = 0; ci < b->cls.size(); ++ci) {
    BatchClass& c = b->cls[ci];
    psx_scft* e0 = b->eng[c.members[0]];
    const bool modes = plane_available(e0);
    if (modes && e0->plane) PSX_TRY(batch_class_mode(b, ci, false));
    for (int m : c.members)
      if (!b->eng[m]->sw[b->prec].wfac_valid) PSX_TRY(compute_wfac(b->eng[m]));
    const int na = (int)e0->axes.size();
    const int K = (int)c.members.size();
    for (int i = 0; i < na; ++i) {
      const PairIO* fio = c.fio[b->prec] + representative_launch(e0->paths[0], i) * 2 * K;
      int* v = &c.variant[b->prec][i];
      if (b->prec == PSX_FP64) PSX_TRY(tune_pass<double>(e0->axes[i], b->prec, fio, 2 * K, b->stream, reps, v));
      else PSX_TRY(tune_pass<float>(e0->axes[i], b->prec, fio, 2 * K, b->stream, reps, v));
      if (h_choice) h_choice[ci * 3 + i] = *v;
    }
    if (modes) {
      // whole class sweeps, line launches against plane launches
      float ms[2] = {0.f, 0.f};
      for (int on = 0; on < 2; ++on) {
        if (on) {
          PSX_TRY(batch_class_mode(b, ci, true));
          for (int m : c.members) PSX_TRY(compute_wfac(b->eng[m]));
        }
        PSX_TRY(batch_time_class(b, ci, reps, &ms[on]));
      }
      if (ms[0] < ms[1]) PSX_TRY(batch_class_mode(b, ci, false));
    }
  }
  batch_schedule(b);
  batch_drop_graphs(b);
  for (auto* e : b->eng) e->swept = false;
  return PSX_OK;
}

template <typename T>
static psx_status batch_sweep(psx_batch* b, size_t ci) {
  BatchClass& c = b->cls[ci];
  psx_scft* e0 = b->eng[c.members[0]];
  const PathInfo& pi = e0->paths[0];
  const int K = (int)c.members.size();
  hipStream_t s = b->ss[b->cls_stream[ci]];
  for (size_t j = 0; j < pi.launch_axis.size(); ++j) {
    const int a = pi.launch_axis[j];
    PSX_TRY(launch_sweep_op<T>(e0, a, b->prec, c.fio[b->prec] + j * 2 * K, 2 * K, s, c.variant[b->prec]));
  }
  PSX_HIP(hipEventRecord(b->ev_cls[ci], s));
  return PSX_OK;
}

// Device time (ms) of reps sweeps of class ci on its stream.
static psx_status batch_time_class(psx_batch* b, size_t ci, int reps, float* ms) {
  hipStream_t s = b->ss[b->cls_stream[ci]];
  hipEvent_t e0, e1;
  PSX_HIP(hipEventCreate(&e0));
  PSX_HIP(hipEventCreate(&e1));
  psx_status st = PSX_OK;
  for (int r = 0; r <= reps && st == PSX_OK; ++r) {   // r = 0 warms
    if (r == 1 && hipEventRecord(e0, s) != hipSuccess) st = fail(PSX_ERR_HIP, "event record");
    if (st == PSX_OK) st = b->prec == PSX_FP64 ? batch_sweep<double>(b, ci) : batch_sweep<float>(b, ci);
  }
  if (st == PSX_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(ms, e0, e1) != hipSuccess))
    st = fail(PSX_ERR_HIP, "event timing");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

// One batch iteration of one group's instances on the batch streams, forked
// from and joined back to b->stream (capturable).  An instance's post-sweep
// work is enqueued on its batch stream by lending that stream to the engine
// for the call.
static psx_status batch_enqueue(psx_batch* b, const psx_mixing* mix, const BatchGroup& grp) {
  PSX_HIP(hipEventRecord(b->ev_fork, b->stream));
  for (int k = 1; k < kBatchStreams; ++k) PSX_HIP(hipStreamWaitEvent(b->ss[k], b->ev_fork, 0));
  for (int ci : grp.cls) {
    if (b->prec == PSX_FP64) PSX_TRY(batch_sweep<double>(b, (size_t)ci));
    else PSX_TRY(batch_sweep<float>(b, (size_t)ci));
  }
  for (int i : grp.post) {
    psx_scft* e = b->eng[i];
    hipStream_t s = b->ss[b->inst_stream[i]];
    const int c = b->inst_cls[i];
    if (b->cls_stream[c] != b->inst_stream[i]) PSX_HIP(hipStreamWaitEvent(s, b->ev_cls[c], 0));
    hipStream_t own = e->stream;
    e->stream = s;
    const psx_status st = enqueue_post_sweep(e, mix);
    e->stream = own;
    PSX_TRY(st);
  }
  for (int k = 1; k < kBatchStreams; ++k) {
    PSX_HIP(hipEventRecord(b->ev_join[k], b->ss[k]));
    PSX_HIP(hipStreamWaitEvent(b->stream, b->ev_join[k], 0));
  }
  return PSX_OK;
}

extern "C" psx_status psx_batch_iterate(psx_batch* b, const psx_mixing* mix, int n_iters, int use_graph,
                                        double* hF, double* herr, int* done) {
  if (!b || !mix) return fail(PSX_ERR_INVALID, "null argument");
  if (done) *done = 0;
  if (n_iters <= 0) return PSX_OK;
  for (auto* e : b->eng) PSX_TRY(prepare_iterate(e, mix, n_iters));
  // prepare_iterate may have reallocated an engine's history, ring or
  // partials; graphs captured with the old pointers must not replay
  std::vector<long long> gen;
  for (auto* e : b->eng) gen.push_back(e->graph_gen);
  if (gen != b->graph_gen) {
    batch_drop_graphs(b);
    b->graph_gen = gen;
  }
  const size_t ng = b->groups.size();
  std::vector<hipGraphExec_t> exec(ng, nullptr);
  if (use_graph) {
    const GraphKey key = graph_key(b->prec, mix);
    for (size_t gi = 0; gi < ng; ++gi) {
      auto it = b->graphs.find({key, (int)gi});
      if (it != b->graphs.end()) { exec[gi] = it->second; continue; }
      hipGraph_t graph;
      PSX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      psx_status cs = batch_enqueue(b, mix, b->groups[gi]);
      hipError_t ce = hipStreamEndCapture(b->stream, &graph);
      if (cs != PSX_OK) { if (ce == hipSuccess) (void)hipGraphDestroy(graph); return cs; }
      if (ce != hipSuccess) return fail(PSX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&exec[gi], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) return fail(PSX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
      b->graphs[{key, (int)gi}] = exec[gi];
    }
  }
  for (int it = 0; it < n_iters; ++it) {
    for (size_t gi = 0; gi < ng; ++gi) {
      if (exec[gi]) PSX_HIP(hipGraphLaunch(exec[gi], b->stream));
      else PSX_TRY(batch_enqueue(b, mix, b->groups[gi]));
    }
  }
  PSX_HIP(hipStreamSynchronize(b->stream));
  psx_status result = PSX_OK;
  const size_t n = b->eng.size();
  for (size_t i = 0; i < n; ++i) {
    psx_scft* e = b->eng[i];
    DevState h;
    PSX_TRY(read_state(e, &h));
    if (h.nanflag && result == PSX_OK)
      result = fail(PSX_ERR_NAN, "Q is NaN/Inf in batch instance " + std::to_string(i));
    PSX_TRY(finish_iterate(e, n_iters, hF ? hF + i * n_iters : nullptr, nullptr,
                           herr ? herr + i * n_iters : nullptr));
  }
  if (done) *done = n_iters;
  return result;
}
