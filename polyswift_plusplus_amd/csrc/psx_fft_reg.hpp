// psx_fft_reg.hpp -- register-resident round-trip line FFT (one wave per tile).
//
// A line of n = N1 * N2 complex points is held by N2 lanes, N1 points each:
// lane t owns x[t + N2 r], r < N1.  The forward transform is the two-level
// split  X[k1 + N1 k2] = sum_t W_N2^{t k2} W_n^{t k1} sum_r x[t + N2 r] W_N1^{r k1}:
// an in-register DFT_N1, a twiddle, one LDS transpose inside the wave, then
// N1/N2 in-register DFT_N2 per lane.  The spectral multiply happens in
// registers and the inverse retraces the same steps, so the result lands in
// the lane layout of the input: one round trip costs two wave-local LDS
// transposes.  A wave holds 64/N2 complex lines = 2*64/N2 real lines (two
// real lines share one complex line; the operator is real and even, so
// their results come back in the real and imaginary parts).
#pragma once
#include <hip/hip_runtime.h>
#include "psx_fft_lds.hpp"

namespace psx {

// Transpose buffers.  A wave's 64/N2 complex lines sit TBuf::line complex
// words apart.  For one register index the N2 lanes of a line touch N2
// consecutive words of one row (XOR-swizzled inside it), so with lines n
// words apart (a multiple of 32 words) every line of the wave fell on the
// same LDS banks: 2-4 times the conflict-free access time in FP32.  N2 more
// words per line shift consecutive lines onto disjoint banks.
template <int N1, int N2>
struct TBuf {
  static constexpr int line = N1 * N2 + N2;       // complex words per line
  static constexpr int wave = (64 / N2) * line;   // per wave: 64 (N1 + 1)
};

// 16-byte vectors for staged contiguous rows.
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double2 type; };
template <> struct Vec16<float> { typedef float4 type; };
__device__ __forceinline__ void vmul(double2& a, const double2& b) { a.x *= b.x; a.y *= b.y; }
__device__ __forceinline__ void vmul(float4& a, const float4& b) {
  a.x *= b.x; a.y *= b.y; a.z *= b.z; a.w *= b.w;
}
// LDS rows are padded, so staged 16-byte chunks are only element-aligned.
__device__ __forceinline__ void vstore_lds(double* p, const double2& v) { p[0] = v.x; p[1] = v.y; }
__device__ __forceinline__ void vstore_lds(float* p, const float4& v) {
  p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w;
}
template <typename T> __device__ __forceinline__ typename Vec16<T>::type vload_lds(const T* p);
template <> __device__ __forceinline__ double2 vload_lds<double>(const double* p) {
  double2 v; v.x = p[0]; v.y = p[1]; return v;
}
template <> __device__ __forceinline__ float4 vload_lds<float>(const float* p) {
  float4 v; v.x = p[0]; v.y = p[1]; v.z = p[2]; v.w = p[3]; return v;
}

template <int N>
constexpr int first_radix() {
  return (N % 8 == 0 && N > 8) ? 8 : (N % 4 == 0 ? 4 : (N % 2 == 0 ? 2 : (N % 3 == 0 ? 3 : (N % 5 == 0 ? 5 : N))));
}

template <int SIGN, typename C>
__device__ __forceinline__ C twiddle(const C* __restrict__ tw, int idx) {
  C w = tw[idx];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// u * W_D^e for SIGN = -1 (W_D = exp(-2 pi i / D)), u * conj(W_D^e) for
// SIGN = +1; tw[idx] is the same twiddle in the size-n table.  e and D are
// compile-time once the callers' loops are unrolled, so the branches fold:
// powers of -i take no multiply, odd powers of W_8 a scaled sum and
// difference (1/sqrt(2)), everything else the table entry.
template <int SIGN, typename T, typename C>
__device__ __forceinline__ C tw_mul(const C& u, int e, int D, const C* __restrict__ tw, int idx) {
  if ((4 * e) % D == 0) {
    int k = ((4 * e) / D) & 3;          // W_4^k = (-i)^k
    if (SIGN > 0) k = (4 - k) & 3;
    C r;
    if (k == 0) r = u;
    else if (k == 1) { r.x = u.y; r.y = -u.x; }
    else if (k == 2) { r.x = -u.x; r.y = -u.y; }
    else { r.x = -u.y; r.y = u.x; }
    return r;
  }
  if ((8 * e) % D == 0) {
    int k = ((8 * e) / D) & 7;          // odd: W_8^k = ((+-1) + (+-1) i) / sqrt(2)
    if (SIGN > 0) k = (8 - k) & 7;
    const T h = (T)0.70710678118654752440;
    C r;
    if (k == 1) { r.x = h * (u.x + u.y); r.y = h * (u.y - u.x); }
    else if (k == 3) { r.x = h * (u.y - u.x); r.y = -h * (u.x + u.y); }
    else if (k == 5) { r.x = -h * (u.x + u.y); r.y = h * (u.x - u.y); }
    else { r.x = h * (u.x - u.y); r.y = h * (u.x + u.y); }
    return r;
  }
  return cmul(u, twiddle<SIGN>(tw, idx));
}

// In-register natural-order DFT of size N (all indices compile-time).
// tw is the size-n table exp(-2 pi i j / n); twst = n / N.
template <int N, int SIGN, typename T, typename C>
__device__ __forceinline__ void dft_reg(C* v, const C* __restrict__ tw, int twst) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 8) {
    bfly<N, SIGN, T>(v);
  } else {
    constexpr int R = first_radix<N>();
    constexpr int M = N / R;
    static_assert(R * M == N && R > 1 && M > 1, "unsupported in-register DFT size");
    C U[N];
#pragma unroll
    for (int b = 0; b < M; ++b) {
      C u[R];
#pragma unroll
      for (int a = 0; a < R; ++a) u[a] = v[b + M * a];
      bfly<R, SIGN, T>(u);
#pragma unroll
      for (int c = 0; c < R; ++c) {
        U[c * M + b] = tw_mul<SIGN, T>(u[c], b * c, N, tw, b * c * twst);
      }
    }
#pragma unroll
    for (int c = 0; c < R; ++c) {
      C m[M];
#pragma unroll
      for (int b = 0; b < M; ++b) m[b] = U[c * M + b];
      dft_reg<M, SIGN, T>(m, tw, twst * R);
#pragma unroll
      for (int k2 = 0; k2 < M; ++k2) v[c + R * k2] = m[k2];
    }
  }
}

// Hand-over of LDS data between the lanes of the transposes: a block barrier,
// or (WAVE) wave-scope ordering only -- a wave's LDS operations complete in
// issue order, so fences that keep the compiler from moving them across the
// hand-over are enough when no other wave touches the buffer.
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One round trip of a wave's lines held in registers (lane t of line p owns
// points t + N2 r):  v <- IFFT(emul * FFT(v)), result in the same layout.
// L is the wave's LDS transpose buffer (p-th line at L + p n).
// tw: global twiddle table (uniform indices, scalar loads); twl, emul: LDS
// copies of the twiddles and the spectral multiplier (per-lane indices).
template <typename T, int N1, int N2, bool WS, typename C>
__device__ __forceinline__ void reg_roundtrip(C* v, C* L, const C* __restrict__ tw,
                                              const C* twl, const T* emul, int t) {
  constexpr int U = N1 / N2;
  // forward: DFT_N1 in registers, twiddle W_n^{t k1}
  dft_reg<N1, -1, T>(v, tw, N2);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twiddle<-1>(twl, t * k1));
  // transpose: element (k1, t) at k1*N2 + (t ^ (k1 % N2)) (XOR swizzle)
  lds_sync<WS>();
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) L[k1 * N2 + (t ^ (k1 % N2))] = v[k1];
  lds_sync<WS>();
  // lane t now takes rows k1 = t + N2 u, u < U; y[u][s] = element (k1, s)
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int s = 0; s < N2; ++s) v[u * N2 + s] = L[(t + N2 * u) * N2 + (s ^ t)];
  // DFT_N2, spectral multiply at k = k1 + N1 k2, inverse DFT_N2
#pragma unroll
  for (int u = 0; u < U; ++u) {
    C* y = v + u * N2;
    dft_reg<N2, -1, T>(y, tw, N1);
    const int k1 = t + N2 * u;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const T m = emul[k1 + N1 * k2];
      y[k2].x *= m; y[k2].y *= m;
    }
    dft_reg<N2, +1, T>(y, tw, N1);
  }
  lds_sync<WS>();
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int s = 0; s < N2; ++s) L[(t + N2 * u) * N2 + (s ^ t)] = v[u * N2 + s];
  lds_sync<WS>();
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) v[k1] = L[k1 * N2 + (t ^ (k1 % N2))];
  // inverse: conjugate twiddle, inverse DFT_N1; lanes hold x[t + N2 r]
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twiddle<+1>(twl, t * k1));
  dft_reg<N1, +1, T>(v, tw, N2);
}

// Lane-pair splits (N1 % N2 != 0) end their forward transform with a radix 2
// across a lane pair and start the inverse with one.  Between the two the
// spectral multiply is diagonal, so the three steps are one 2x2 operator per
// pair of spectral points k = k1 + N1 j and k + n/2 (W = W_N2^j, K the row of
// spectral factors, a = K(k) + K(k + n/2), b = K(k) - K(k + n/2)):
//   lane h = 0 (own E_0, partner E_1):  E_0 <- a E_0 + b W E_1
//   lane h = 1 (own E_1, partner E_0):  E_1 <- a E_1 + b conj(W) E_0
// Entry k < n/2 of a row's (a, b) row holds a and b; the host stores it
// behind every row of spectral factors (spectral_rows), and W comes from the
// twiddle table.  Per point: one exchange and 4-9 multiply-adds instead of
// two exchanges, two twiddle products, the selects of the lane roles and the
// multiply.  Measured at 128^3 (static VALU instructions, dual contour step):
// the FP32 plane kernel 2266 -> 1731 and 18.2 -> 17.7 us; the FP64 x launch
// 1527 -> 1120 at 126 instead of 129 VGPRs and 32.5 -> 31.7 us; at 96^3 the
// FP32 step 18.3 -> 16.5 us, the FP64 step 25.0 -> 22.0 us.
template <typename T>
struct __align__(2 * sizeof(T)) PairAB { T a, b; };

template <int N1, int N2>
constexpr bool pair_split() { return N1 % N2 != 0; }

// LDS bytes of a kernel's table copies: the twiddles (n complex) and `rows`
// rows of n spectral values (the factors, or their (a, b) rows).
template <typename T, int N1, int N2>
constexpr int table_bytes(int rows) {
  return N1 * N2 * 2 * (int)sizeof(T) + rows * N1 * N2 * (int)sizeof(T);
}

// The row a round trip reads for the spectral row K (n values, its (a, b)
// row the n values behind it in line launches): the (a, b) row for lane-pair
// splits, K itself otherwise.
template <int N1, int N2, typename T>
__device__ __forceinline__ const T* pair_row(const T* K) {
  return pair_split<N1, N2>() ? K + N1 * N2 : K;
}

// Exchange with the partner lane (lane ^ 1) through a DPP quad permutation
// [1, 0, 3, 2] (VALU moves, no LDS).
__device__ __forceinline__ float pair_swap(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pair_swap(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// The fused radix 2 / multiply / inverse radix 2 of one point from its (a, b)
// entry: y <- a y + b W_h o with o the partner's value, W_h = W_N2^j for lane
// h = 0 and conj(W_N2^j) for h = 1 (sg = +1 / -1), W_N2^j = tw[tj].
// j is compile-time: j = 0 takes no twiddle, j = N2/4 (W = -i) a sign.
template <int N2, typename T, typename C>
__device__ __forceinline__ void pair_apply_ab(C& y, const PairAB<T>& q, int j, T sg,
                                              const C* __restrict__ tw, int tj) {
  C o;
  o.x = pair_swap(y.x);
  o.y = pair_swap(y.y);
  if (j == 0) {
    y.x = q.a * y.x + q.b * o.x;
    y.y = q.a * y.y + q.b * o.y;
  } else if ((4 * j) % N2 == 0) {   // W_h o = sg (o.y, -o.x)
    const T bs = sg * q.b;
    y.x = q.a * y.x + bs * o.y;
    y.y = q.a * y.y - bs * o.x;
  } else {
    const C w = tw[tj];
    const T s = sg * w.y;
    C u;
    u.x = w.x * o.x - s * o.y;
    u.y = w.x * o.y + s * o.x;
    y.x = q.a * y.x + q.b * u.x;
    y.y = q.a * y.y + q.b * u.y;
  }
}

// Lane-pair round trip for n = 2 N1^2 on N2 = 2 N1 lanes, N1 points per lane
// (lane t owns x[t + N2 r]): half the registers per lane and twice the lines
// in flight of the N1 = N2^2 / n split.  Forward, with k = k1 + N1 k2g:
//   in-register DFT_N1 over r, twiddle W_n^{t k1};
//   LDS transpose: lane l = 2 k1 + h takes the points t = h + 2m of column k1
//   (XOR swizzle 2 k1 on the column index keeps writes and reads on distinct banks);
//   in-register DFT_N1 over m: E_h[k2] = sum_m W_N1^{m k2} z[h + 2m];
//   radix 2 across the pair: lane h keeps X[k1 + N1 (k2 + N1 h)] =
//   E_0[k2] + (-1)^h W_N2^{k2} E_1[k2], k2 < N1.
// That radix 2, the spectral multiply and the inverse radix 2 are one
// operator (pair_apply_ab) from the (a, b) row ab (entry k1 + N1 k2); then
// the forward steps inverted.
template <typename T, int N1, bool WS, typename C>
__device__ __forceinline__ void pair_roundtrip(C* v, C* L, const C* __restrict__ tw,
                                               const C* twl, const PairAB<T>* ab, int t) {
  constexpr int N2 = 2 * N1;
  dft_reg<N1, -1, T>(v, tw, N2);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twiddle<-1>(twl, t * k1));
  lds_sync<WS>();
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) L[k1 * N2 + (t ^ ((2 * k1) % N2))] = v[k1];
  lds_sync<WS>();
  const int k1 = t >> 1, h = t & 1;
  const int sw = (2 * k1) % N2;
#pragma unroll
  for (int m = 0; m < N1; ++m) v[m] = L[k1 * N2 + ((h + 2 * m) ^ sw)];
  dft_reg<N1, -1, T>(v, tw, N2);
  const T sg = h ? (T)-1 : (T)1;
#pragma unroll
  for (int k2 = 0; k2 < N1; ++k2) pair_apply_ab<N2>(v[k2], ab[k1 + N1 * k2], k2, sg, tw, k2 * N1);
  dft_reg<N1, +1, T>(v, tw, N2);
  lds_sync<WS>();
#pragma unroll
  for (int m = 0; m < N1; ++m) L[k1 * N2 + ((h + 2 * m) ^ sw)] = v[m];
  lds_sync<WS>();
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = L[j * N2 + (t ^ ((2 * j) % N2))];
#pragma unroll
  for (int j = 1; j < N1; ++j) v[j] = cmul(v[j], twiddle<+1>(twl, t * j));
  dft_reg<N1, +1, T>(v, tw, N2);
}

// Lane-pair round trip for the general split 2 N1 % N2 == 0 (N2 a power of
// two): the pair k = t >> 1 takes the P = 2 N1 / N2 columns k1 = k + H u
// (H = N2 / 2, u < P), lane h of the pair the points t' = h + 2 m (m < H) of
// each; per column an in-register DFT_H over m and the fused radix 2 /
// multiply / inverse radix 2 across the pair ((a, b) entry k1 + N1 j), as in
// pair_roundtrip (its P = 1 case).  Registers: column u holds
// v[u H .. u H + H - 1].  With N1 = 12, N2 = 8 a line of 96 takes 8 lanes
// of 12 points: twice the lines per wave of the 24 x 4 split.
template <typename T, int N1, int N2, bool WS, typename C>
__device__ __forceinline__ void pairs_roundtrip(C* v, C* L, const C* __restrict__ tw,
                                                const C* twl, const PairAB<T>* ab, int t) {
  constexpr int n = N1 * N2;
  constexpr int H = N2 / 2;
  constexpr int P = N1 / H;
  static_assert(P * H == N1 && (N2 & (N2 - 1)) == 0, "bad lane-pair split");
  dft_reg<N1, -1, T>(v, tw, N2);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twiddle<-1>(twl, t * k1));
  lds_sync<WS>();
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) L[k1 * N2 + (t ^ ((2 * k1) % N2))] = v[k1];
  lds_sync<WS>();
  const int k = t >> 1, h = t & 1;
#pragma unroll
  for (int u = 0; u < P; ++u) {
    const int k1 = k + H * u, sw = (2 * k1) % N2;
#pragma unroll
    for (int m = 0; m < H; ++m) v[u * H + m] = L[k1 * N2 + ((h + 2 * m) ^ sw)];
  }
#pragma unroll
  for (int u = 0; u < P; ++u) dft_reg<H, -1, T>(v + u * H, tw, n / H);
  const T sg = h ? (T)-1 : (T)1;
#pragma unroll
  for (int u = 0; u < P; ++u) {
    const int k1 = k + H * u;
#pragma unroll
    for (int j = 0; j < H; ++j) pair_apply_ab<N2>(v[u * H + j], ab[k1 + N1 * j], j, sg, tw, j * N1);
  }
#pragma unroll
  for (int u = 0; u < P; ++u) dft_reg<H, +1, T>(v + u * H, tw, n / H);
  lds_sync<WS>();
#pragma unroll
  for (int u = 0; u < P; ++u) {
    const int k1 = k + H * u, sw = (2 * k1) % N2;
#pragma unroll
    for (int m = 0; m < H; ++m) L[k1 * N2 + ((h + 2 * m) ^ sw)] = v[u * H + m];
  }
  lds_sync<WS>();
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = L[j * N2 + (t ^ ((2 * j) % N2))];
#pragma unroll
  for (int j = 1; j < N1; ++j) v[j] = cmul(v[j], twiddle<+1>(twl, t * j));
  dft_reg<N1, +1, T>(v, tw, N2);
}

// Round trip on the lanes of the split (N1, N2): lane-pair form when
// N2 = 2 N1, the N1 % N2 == 0 form, else the general lane-pair form.
// WS: the transpose buffer L is private to the wave, so its write/read
// hand-overs need only wave-level ordering (multi-wave blocks whose callers
// separate the shared uses of that memory with block barriers).
// Lane-pair splits read the (a, b) row of the spectral factors (emul =
// pair_row of the factor row); the other splits the factors themselves.
template <typename T, int N1, int N2, bool WS = false, typename C>
__device__ __forceinline__ void line_roundtrip(C* v, C* L, const C* __restrict__ tw,
                                               const C* twl, const T* emul, int t) {
  if constexpr (N2 == 2 * N1) pair_roundtrip<T, N1, WS>(v, L, tw, twl, (const PairAB<T>*)emul, t);
  else if constexpr (N1 % N2 == 0) reg_roundtrip<T, N1, N2, WS>(v, L, tw, twl, emul, t);
  else pairs_roundtrip<T, N1, N2, WS>(v, L, tw, twl, (const PairAB<T>*)emul, t);
}

// Round trip(s) along one axis for 2*64/N2 real lines per wave:
//   first:  u = post * IFFT(emul * FFT(pre * in))
//   fused:  mid = u;  out = post2 * IFFT(emul2 * FFT(pre2 * u))   (emul2 != null)
//   plain:  out = u                                               (emul2 == null)
// The fused form joins the last pass of one contour step and the first pass
// of the next when both run along the same axis: the step's result is stored
// (mid) and carried on in registers.  VEC: the lines of a wave are
// consecutive in memory (inner % lines-per-wave == 0), so a real pair is one
// aligned complex load.
template <typename T, int N1, int N2, bool CONTIG, bool VEC>
__global__ void __launch_bounds__(64)
k_line_reg(LineGeom g, const PairIO* __restrict__ fio) {
  typedef typename Cx<T>::type C;
  typedef typename Vec16<T>::type V;
  constexpr int n = N1 * N2;
  constexpr int CL = 64 / N2;     // complex lines per wave
  constexpr int RL = 2 * CL;      // real lines per wave
  static_assert((N1 % N2 == 0 || (2 * N1) % N2 == 0) && 64 % N2 == 0, "bad split");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = (C*)smem_raw;
  int tile, field;
  xcd_tile(g.ntiles, g.nfields, &tile, &field);

  const PairIO fd = fio[field];
  const C* __restrict__ tw = (const C*)g.tw;
  const bool fused = fd.emul2 != nullptr;

  const int lane = threadIdx.x;
  const int p = lane / N2, t = lane % N2;
  const int la = tile * RL + 2 * p;   // real lines la, la + 1
  const bool va = la < g.nlines, vb = la + 1 < g.nlines;
  const int inner = g.inner;
  // element k of line l: base(l) + k * stride
  int ba, bb, stride;
  if (CONTIG) { ba = la * n; bb = ba + n; stride = 1; }
  else {
    const int oa = la / inner, ob = (la + 1) / inner;
    ba = oa * n * inner + (la - oa * inner);
    bb = ob * n * inner + (la + 1 - ob * inner);
    stride = inner;
  }
  C* L = lds + p * TBuf<N1, N2>::line;
  // FP32: LDS copies of the per-lane tables after the transpose/staging
  // buffer (shorter latency than L2 in the middle of the round trip).  FP64
  // keeps them in global memory: the extra 4 KB would drop the wave below
  // 8 waves per CU of LDS, and a 128^3 pass needs exactly 8.
  constexpr bool TLDS = sizeof(T) == 4;
  constexpr int TBB = TBuf<N1, N2>::wave * 2 * (int)sizeof(T);
  constexpr int BUF = (TBB > 2 * 64 / N2 * (n + N2 / 2) * (int)sizeof(T))
                          ? TBB
                          : 2 * 64 / N2 * (n + N2 / 2) * (int)sizeof(T);
  const C* twl = tw;
  const T* e1 = pair_row<N1, N2>((const T*)fd.emul);
  const T* e2 = fused ? pair_row<N1, N2>((const T*)fd.emul2) : nullptr;
  if (TLDS) {
    C* tws = (C*)(smem_raw + BUF);
    T* e1s = (T*)(tws + n);
    T* e2s = e1s + n;
    for (int i = lane; i < n; i += 64) {
      tws[i] = tw[i];
      e1s[i] = e1[i];
      if (fused) e2s[i] = e2[i];
    }
    __syncthreads();
    twl = tws; e1 = e1s; e2 = e2s;
  }

  // Contiguous lines: the wave's RL rows are one contiguous block of RL*n
  // elements, moved through LDS with 16-byte accesses; rows are padded by
  // N2/2 elements so that the FFT-layout accesses hit distinct banks.
  constexpr int PAD = N2 / 2;
  constexpr int RS = n + PAD;
  constexpr int VW = 16 / sizeof(T);
  T* S = (T*)smem_raw;
  const int row0 = tile * RL;
  const int nrows = min(RL, g.nlines - row0);
  const int nchunk = nrows * n / VW;

  C v[N1];
  // ---- load (times pre)
  if (CONTIG) {
    const T* src = (const T*)fd.in + (size_t)row0 * n;
    const T* psrc = fd.pre ? (const T*)fd.pre + (size_t)row0 * n : nullptr;
#pragma unroll 4
    for (int c = lane; c < nchunk; c += 64) {
      const int e = c * VW, row = e / n, col = e - row * n;
      V x = *(const V*)(src + e);
      if (psrc) { const V w = *(const V*)(psrc + e); vmul(x, w); }
      vstore_lds(S + row * RS + col, x);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int k = t + N2 * r;
      C x;
      x.x = (2 * p < nrows) ? S[(2 * p) * RS + k] : (T)0;
      x.y = (2 * p + 1 < nrows) ? S[(2 * p + 1) * RS + k] : (T)0;
      v[r] = x;
    }
  } else {
    const T* __restrict__ in = (const T*)fd.in;
    const T* __restrict__ pre = (const T*)fd.pre;
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int k = t + N2 * r;
      C x;
      if (VEC) {
        if (va) {
          x = *(const C*)(in + ba + k * stride);
          if (pre) { const C w = *(const C*)(pre + ba + k * stride); x.x *= w.x; x.y *= w.y; }
        } else { x.x = (T)0; x.y = (T)0; }
      } else {
        x.x = va ? in[ba + k * stride] : (T)0;
        x.y = vb ? in[bb + k * stride] : (T)0;
        if (pre) {
          if (va) x.x *= pre[ba + k * stride];
          if (vb) x.y *= pre[bb + k * stride];
        }
      }
      v[r] = x;
    }
  }

  line_roundtrip<T, N1, N2>(v, L, tw, twl, e1, t);

  // ---- store: the step result (post) to mid and on through the second
  // round trip, or straight to out
  for (int stage = 0; stage < 2; ++stage) {
    const bool last = stage == 1 || !fused;
    T* __restrict__ dst = (T*)(last ? fd.out : fd.mid);
    const T* __restrict__ post = (const T*)(stage == 0 ? fd.post : fd.post2);
    const T* __restrict__ pre2 = last ? nullptr : (const T*)fd.pre2;
    if (CONTIG) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        const int k = t + N2 * r;
        if (2 * p < nrows) S[(2 * p) * RS + k] = v[r].x;
        if (2 * p + 1 < nrows) S[(2 * p + 1) * RS + k] = v[r].y;
      }
      __syncthreads();
      T* d = dst + (size_t)row0 * n;
      const T* pp = post ? post + (size_t)row0 * n : nullptr;
      const T* qq = pre2 ? pre2 + (size_t)row0 * n : nullptr;
#pragma unroll 4
      for (int c = lane; c < nchunk; c += 64) {
        const int e = c * VW, row = e / n, col = e - row * n;
        V x = vload_lds<T>(S + row * RS + col);
        if (pp) { const V w = *(const V*)(pp + e); vmul(x, w); }
        *(V*)(d + e) = x;
        if (!last) {
          if (qq) { const V w = (qq == pp) ? *(const V*)(pp + e) : *(const V*)(qq + e); vmul(x, w); }
          vstore_lds(S + row * RS + col, x);
        }
      }
      if (!last) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < N1; ++r) {
          const int k = t + N2 * r;
          v[r].x = (2 * p < nrows) ? S[(2 * p) * RS + k] : (T)0;
          v[r].y = (2 * p + 1 < nrows) ? S[(2 * p + 1) * RS + k] : (T)0;
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        const int k = t + N2 * r;
        C x = v[r];
        if (VEC) {
          if (va) {
            C w;
            if (post) { w = *(const C*)(post + ba + k * stride); x.x *= w.x; x.y *= w.y; }
            *(C*)(dst + ba + k * stride) = x;
            if (!last && pre2) {
              if (pre2 != post) w = *(const C*)(pre2 + ba + k * stride);
              x.x *= w.x; x.y *= w.y;
            }
          }
        } else {
          if (post) {
            if (va) x.x *= post[ba + k * stride];
            if (vb) x.y *= post[bb + k * stride];
          }
          if (va) dst[ba + k * stride] = x.x;
          if (vb) dst[bb + k * stride] = x.y;
          if (!last && pre2) {
            if (va) x.x *= pre2[ba + k * stride];
            if (vb) x.y *= pre2[bb + k * stride];
          }
        }
        v[r] = x;
      }
    }
    if (last) break;
    line_roundtrip<T, N1, N2>(v, L, tw, twl, e2, t);
  }
}

// Strided axes through 128-byte rows: a block of W waves takes LT = W * RL
// lines that sit next to each other in memory (one 128-byte segment per
// point), loads the tile [n points][LT lines] with 16-byte chunks into LDS,
// and every wave reads its RL lines in the lane layout of line_roundtrip.
// The staging area doubles as the waves' transpose buffers (barriers keep the
// two uses apart).  Stores go back through the same tile.  Same stages as
// k_line_reg (pre, fused mid / pre2 / second round trip, post).
template <typename T, int N1, int N2, int W>
__global__ void __launch_bounds__(64 * W)
k_line_tile(LineGeom g, const PairIO* __restrict__ fio) {
  typedef typename Cx<T>::type C;
  typedef typename Vec16<T>::type V;
  constexpr int n = N1 * N2;
  constexpr int RL = 2 * 64 / N2;                 // real lines per wave
  constexpr int LT = W * RL;                      // real lines per tile
  constexpr int LS = LT + 2;                      // LDS row stride (elements)
  constexpr int VW = 16 / (int)sizeof(T);
  constexpr int CPR = LT / VW;                    // 16-byte chunks per row
  constexpr int NCH = n * CPR;
  constexpr int SB = n * LS * (int)sizeof(T);
  constexpr int TB = W * TBuf<N1, N2>::wave * 2 * (int)sizeof(T);
  constexpr int BUF = ((SB > TB ? SB : TB) + 15) / 16 * 16;
  static_assert(LT % VW == 0, "tile rows must be whole 16-byte chunks");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* S = (T*)smem_raw;
  int tile, field;
  xcd_tile(g.ntiles, g.nfields, &tile, &field);
  const PairIO fd = fio[field];
  const C* __restrict__ tw = (const C*)g.tw;
  const bool fused = fd.emul2 != nullptr;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int p = lane / N2, t = lane % N2;
  const int inner = g.inner;
  const int l0 = tile * LT;
  const long long base = (long long)(l0 / inner) * n * inner + (l0 - (l0 / inner) * inner);
  C* L = (C*)smem_raw + wave * TBuf<N1, N2>::wave + p * TBuf<N1, N2>::line;
  const int j0 = wave * RL + 2 * p;               // this lane's real-line pair in the tile

  constexpr bool TLDS = sizeof(T) == 4;
  const C* twl = tw;
  const T* e1 = pair_row<N1, N2>((const T*)fd.emul);
  const T* e2 = fused ? pair_row<N1, N2>((const T*)fd.emul2) : nullptr;
  if (TLDS) {
    C* tws = (C*)(smem_raw + BUF);
    T* e1s = (T*)(tws + n);
    T* e2s = e1s + n;
    for (int i = tid; i < n; i += 64 * W) {
      tws[i] = tw[i];
      e1s[i] = e1[i];
      if (fused) e2s[i] = e2[i];
    }
    twl = tws; e1 = e1s; e2 = e2s;
  }

  // ---- load the tile (times pre), then the lane layout
  {
    const T* src = (const T*)fd.in;
    const T* pre = (const T*)fd.pre;
#pragma unroll 4
    for (int c = tid; c < NCH; c += 64 * W) {
      const int row = c / CPR, col = (c - row * CPR) * VW;
      const long long ga = base + (long long)row * inner + col;
      V x = *(const V*)(src + ga);
      if (pre) { const V w = *(const V*)(pre + ga); vmul(x, w); }
      vstore_lds(S + row * LS + col, x);
    }
  }
  __syncthreads();
  C v[N1];
#pragma unroll
  for (int r = 0; r < N1; ++r) {
    const int k = t + N2 * r;
    v[r].x = S[k * LS + j0];
    v[r].y = S[k * LS + j0 + 1];
  }
  __syncthreads();
  line_roundtrip<T, N1, N2, true>(v, L, tw, twl, e1, t);

  for (int stage = 0; stage < 2; ++stage) {
    const bool last = stage == 1 || !fused;
    T* __restrict__ dst = (T*)(last ? fd.out : fd.mid);
    const T* __restrict__ post = (const T*)(stage == 0 ? fd.post : fd.post2);
    const T* __restrict__ pre2 = last ? nullptr : (const T*)fd.pre2;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int k = t + N2 * r;
      S[k * LS + j0] = v[r].x;
      S[k * LS + j0 + 1] = v[r].y;
    }
    __syncthreads();
#pragma unroll 4
    for (int c = tid; c < NCH; c += 64 * W) {
      const int row = c / CPR, col = (c - row * CPR) * VW;
      const long long ga = base + (long long)row * inner + col;
      V x = vload_lds<T>(S + row * LS + col);
      if (post) { const V w = *(const V*)(post + ga); vmul(x, w); }
      *(V*)(dst + ga) = x;
      if (!last) {
        if (pre2) { const V w = (pre2 == post) ? *(const V*)(post + ga) : *(const V*)(pre2 + ga); vmul(x, w); }
        vstore_lds(S + row * LS + col, x);
      }
    }
    if (last) break;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int k = t + N2 * r;
      v[r].x = S[k * LS + j0];
      v[r].y = S[k * LS + j0 + 1];
    }
    __syncthreads();
    line_roundtrip<T, N1, N2, true>(v, L, tw, twl, e2, t);
  }
}

// Whole line in one lane (n <= 8): no transposes.  Same stages as k_line_reg.
template <typename T, int NN, bool CONTIG>
__global__ void __launch_bounds__(64)
k_line_lane(LineGeom g, const PairIO* __restrict__ fio) {
  typedef typename Cx<T>::type C;
  int tile, field;
  xcd_tile(g.ntiles, g.nfields, &tile, &field);
  const PairIO fd = fio[field];
  const C* __restrict__ tw = (const C*)g.tw;
  const int la = (tile * 64 + threadIdx.x) * 2;
  const bool va = la < g.nlines, vb = la + 1 < g.nlines;
  const int inner = g.inner;
  int ba, bb, stride;
  if (CONTIG) { ba = la * NN; bb = ba + NN; stride = 1; }
  else {
    const int oa = la / inner, ob = (la + 1) / inner;
    ba = oa * NN * inner + (la - oa * inner);
    bb = ob * NN * inner + (la + 1 - ob * inner);
    stride = inner;
  }
  const T* __restrict__ in = (const T*)fd.in;
  const T* __restrict__ pre = (const T*)fd.pre;
  C v[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    C x;
    x.x = va ? in[ba + k * stride] : (T)0;
    x.y = vb ? in[bb + k * stride] : (T)0;
    if (pre) {
      if (va) x.x *= pre[ba + k * stride];
      if (vb) x.y *= pre[bb + k * stride];
    }
    v[k] = x;
  }
  const bool fused = fd.emul2 != nullptr;
  for (int stage = 0; stage < 2; ++stage) {
    const T* __restrict__ emul = (const T*)(stage == 0 ? fd.emul : fd.emul2);
    dft_reg<NN, -1, T>(v, tw, 1);
#pragma unroll
    for (int k = 0; k < NN; ++k) { const T m = emul[k]; v[k].x *= m; v[k].y *= m; }
    dft_reg<NN, +1, T>(v, tw, 1);
    const bool last = stage == 1 || !fused;
    T* __restrict__ dst = (T*)(last ? fd.out : fd.mid);
    const T* __restrict__ post = (const T*)(stage == 0 ? fd.post : fd.post2);
    const T* __restrict__ pre2 = last ? nullptr : (const T*)fd.pre2;
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      if (post) {
        if (va) v[k].x *= post[ba + k * stride];
        if (vb) v[k].y *= post[bb + k * stride];
      }
      if (va) dst[ba + k * stride] = v[k].x;
      if (vb) dst[bb + k * stride] = v[k].y;
      if (pre2) {
        if (va) v[k].x *= pre2[ba + k * stride];
        if (vb) v[k].y *= pre2[bb + k * stride];
      }
    }
    if (last) break;
  }
}

}  // namespace psx
