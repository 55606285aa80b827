// psx_fft_cline.hpp -- register-resident complex line passes for the general
// FFT plugin (forward, inverse, or forward * kmul * inverse along one axis).
//
// The lane layout and the two-level split are those of the real-line round
// trips in psx_fft_reg.hpp: a line of n = N1 * N2 complex points sits on N2
// lanes, lane t holding x[t + N2 r].  Here the forward and inverse halves are
// separate, so a pass can stop in the frequency domain: after the forward
// half, register i of lane t holds X[spec_k(t, i)], and the pass stores it at
// its natural position (the next pass, along another axis, reads natural
// order).  The inverse half starts from that layout.  A wave holds 64 / N2
// complex lines; lanes past the last line join the transposes but load and
// store nothing.
#pragma once
#include "psx_fft_reg.hpp"

namespace psx {

// Split forms: N1 % N2 == 0 (the k1 columns of a lane follow t + N2 u), or
// lane pairs for 2 N1 % N2 == 0 (columns k + H u of the pair k = t >> 1,
// H = N2 / 2; N2 = 2 N1 is its one-column case).
template <int N1, int N2>
constexpr bool cline_pairs() { return !(N1 % N2 == 0) || N2 == 2 * N1; }

// Natural frequency index of register i of lane t after the forward half.
template <int N1, int N2>
__device__ __forceinline__ int spec_k(int t, int i) {
  if constexpr (!cline_pairs<N1, N2>()) {
    const int u = i / N2, k2 = i - u * N2;
    return (t + N2 * u) + N1 * k2;
  } else {
    constexpr int H = N2 / 2;
    const int u = i / H, j = i - u * H;
    return ((t >> 1) + H * u) + N1 * (j + H * (t & 1));
  }
}

template <typename T, int N1, int N2, typename C>
__device__ __forceinline__ void fwd_half(C* v, C* L, const C* __restrict__ tw, int t) {
  constexpr int n = N1 * N2;
  dft_reg<N1, -1, T>(v, tw, N2);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twiddle<-1>(tw, t * k1));
  if constexpr (!cline_pairs<N1, N2>()) {
    constexpr int U = N1 / N2;
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) L[k1 * N2 + (t ^ (k1 % N2))] = v[k1];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int s = 0; s < N2; ++s) v[u * N2 + s] = L[(t + N2 * u) * N2 + (s ^ t)];
#pragma unroll
    for (int u = 0; u < U; ++u) dft_reg<N2, -1, T>(v + u * N2, tw, N1);
  } else {
    constexpr int H = N2 / 2, P = N1 / H;
    static_assert(P * H == N1 && (N2 & (N2 - 1)) == 0, "bad lane-pair split");
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) L[k1 * N2 + (t ^ ((2 * k1) % N2))] = v[k1];
    __syncthreads();
    const int k = t >> 1, h = t & 1;
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const int k1 = k + H * u, sw = (2 * k1) % N2;
#pragma unroll
      for (int m = 0; m < H; ++m) v[u * H + m] = L[k1 * N2 + ((h + 2 * m) ^ sw)];
    }
#pragma unroll
    for (int u = 0; u < P; ++u) dft_reg<H, -1, T>(v + u * H, tw, n / H);
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const int j = i % H;
      C& y = v[i];
      C o;
      o.x = pair_swap(y.x);
      o.y = pair_swap(y.y);
      const C e0 = h ? o : y;
      const C e1 = h ? y : o;
      const C w1 = tw_mul<-1, T>(e1, j, N2, tw, j * N1);
      if (h) { y.x = e0.x - w1.x; y.y = e0.y - w1.y; }
      else { y.x = e0.x + w1.x; y.y = e0.y + w1.y; }
    }
  }
}

template <typename T, int N1, int N2, typename C>
__device__ __forceinline__ void inv_half(C* v, C* L, const C* __restrict__ tw, int t) {
  constexpr int n = N1 * N2;
  if constexpr (!cline_pairs<N1, N2>()) {
    constexpr int U = N1 / N2;
#pragma unroll
    for (int u = 0; u < U; ++u) dft_reg<N2, +1, T>(v + u * N2, tw, N1);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int s = 0; s < N2; ++s) L[(t + N2 * u) * N2 + (s ^ t)] = v[u * N2 + s];
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) v[k1] = L[k1 * N2 + (t ^ (k1 % N2))];
  } else {
    constexpr int H = N2 / 2, P = N1 / H;
    const int k = t >> 1, h = t & 1;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const int j = i % H;
      C& y = v[i];
      C o;
      o.x = pair_swap(y.x);
      o.y = pair_swap(y.y);
      const C lo = h ? o : y;
      const C hi = h ? y : o;
      if (h) {
        C d; d.x = lo.x - hi.x; d.y = lo.y - hi.y;
        y = tw_mul<+1, T>(d, j, N2, tw, j * N1);
      } else {
        y.x = lo.x + hi.x; y.y = lo.y + hi.y;
      }
    }
#pragma unroll
    for (int u = 0; u < P; ++u) dft_reg<H, +1, T>(v + u * H, tw, n / H);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const int k1 = k + H * u, sw = (2 * k1) % N2;
#pragma unroll
      for (int m = 0; m < H; ++m) L[k1 * N2 + ((h + 2 * m) ^ sw)] = v[u * H + m];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N1; ++j) v[j] = L[j * N2 + (t ^ ((2 * j) % N2))];
  }
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twiddle<+1>(tw, t * k1));
  dft_reg<N1, +1, T>(v, tw, N2);
}

// One pass along one axis of the general pair, CL = 64 / N2 complex lines
// per wave (one-wave blocks).  MODE 0: forward (natural in, natural
// spectrum out); 1: inverse; 2: forward, times the real full-grid kmul at
// each (line, frequency), inverse.  in_mode / out_mode as CplxIO.
template <typename T, int N1, int N2, int MODE>
__global__ void __launch_bounds__(64)
k_cline(LineGeom g, CplxIO f) {
  typedef typename Cx<T>::type C;
  constexpr int n = N1 * N2;
  constexpr int CL = 64 / N2;
  static_assert((N1 % N2 == 0 || (2 * N1) % N2 == 0) && 64 % N2 == 0, "bad split");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const C* __restrict__ tw = (const C*)g.tw;
  const int lane = threadIdx.x;
  const int p = lane / N2, t = lane % N2;
  const int l = blockIdx.x * CL + p;
  const bool valid = l < g.nlines;
  C* L = (C*)smem_raw + p * TBuf<N1, N2>::line;
  const int inner = g.inner;
  const long long lb = valid ? (long long)(l / inner) * n * inner + (l - (l / inner) * inner) : 0;
  C v[N1];
  if (MODE != 1) {
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const long long a = lb + (long long)(t + N2 * r) * inner;
      C x; x.x = (T)0; x.y = (T)0;
      if (valid) {
        if (f.in_mode == 1) x.x = ((const T*)f.in)[a];
        else if (f.in_mode == 2) x.y = ((const T*)f.in)[a];
        else x = ((const C*)f.in)[a];
      }
      v[r] = x;
    }
    fwd_half<T, N1, N2>(v, L, tw, t);
  } else {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const long long a = lb + (long long)spec_k<N1, N2>(t, i) * inner;
      C x; x.x = (T)0; x.y = (T)0;
      if (valid) {
        if (f.in_mode == 1) x.x = ((const T*)f.in)[a];
        else if (f.in_mode == 2) x.y = ((const T*)f.in)[a];
        else x = ((const C*)f.in)[a];
      }
      v[i] = x;
    }
  }
  if (MODE == 2 && f.kmul) {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const long long a = lb + (long long)spec_k<N1, N2>(t, i) * inner;
      const T m = valid ? ((const T*)f.kmul)[a] : (T)0;
      v[i].x *= m; v[i].y *= m;
    }
  }
  if (MODE == 0) {
    if (!valid) return;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const long long a = lb + (long long)spec_k<N1, N2>(t, i) * inner;
      const C x = v[i];
      if (f.out_mode == 1) ((T*)f.out)[a] = x.x;
      else if (f.out_mode == 2) ((T*)f.out)[a] = x.x * x.x + x.y * x.y;
      else ((C*)f.out)[a] = x;
    }
    return;
  }
  inv_half<T, N1, N2>(v, L, tw, t);
  if (!valid) return;
#pragma unroll
  for (int r = 0; r < N1; ++r) {
    const long long a = lb + (long long)(t + N2 * r) * inner;
    const C x = v[r];
    if (f.out_mode == 1) ((T*)f.out)[a] = x.x;
    else if (f.out_mode == 2) ((T*)f.out)[a] = x.x * x.x + x.y * x.y;
    else ((C*)f.out)[a] = x;
  }
}

}  // namespace psx
