// psx_fft_lds.hpp -- batched complex FFTs of short lines held in LDS.
//
// Mixed-radix Stockham autosort (radix 2, 3, 4, 5, 8 and a generic odd
// radix up to 15).  Each stage reads one LDS buffer and writes the other, so
// a transform of P lines of length n needs 2 * P * ls complex words of LDS
// (ls = line stride, n + 1 to break power-of-two bank strides).  Stage s with
// radix R and sub-transform length Ns does, for butterfly j < n/R:
//   v_r = x[j + r n/R] * W_{Ns R}^{r (j mod Ns)}, v = DFT_R(v),
//   y[(j - j mod Ns) R + j mod Ns + r Ns] = v_r.
// SIGN = -1 is the forward transform exp(-2 pi i jk/n), +1 the inverse; both
// unnormalised.  The twiddle table tw[j] = exp(-2 pi i j / n) is built on the
// host in long double and rounded once to T.
#pragma once
#include <hip/hip_runtime.h>

namespace psx {

template <typename T> struct Cx;
template <> struct Cx<double> { typedef double2 type; };
template <> struct Cx<float> { typedef float2 type; };

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename C, typename T>
__device__ __forceinline__ C cscale(C a, T s) { C r; r.x = a.x * s; r.y = a.y * s; return r; }
// a * (SIGN i)
template <int SIGN, typename C>
__device__ __forceinline__ C mul_si(C a) {
  C r;
  if (SIGN < 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
  return r;
}

template <int SIGN, typename T, typename C>
__device__ __forceinline__ void bfly2(C* v) {
  C a = v[0], b = v[1];
  v[0] = cadd(a, b); v[1] = csub(a, b);
}

template <int SIGN, typename T, typename C>
__device__ __forceinline__ void bfly4(C* v) {
  C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  C t2 = cadd(v[1], v[3]), t3 = mul_si<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
}

template <int SIGN, typename T, typename C>
__device__ __forceinline__ void bfly8(C* v) {
  C e[4] = {v[0], v[2], v[4], v[6]};
  C o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<SIGN, T>(e);
  bfly4<SIGN, T>(o);
  const T h = (T)0.70710678118654752440;
  C o1, o3;
  if (SIGN < 0) {  // W8 = (1 - i)/sqrt2, W8^3 = (-1 - i)/sqrt2
    o1.x = (o[1].x + o[1].y) * h; o1.y = (o[1].y - o[1].x) * h;
    o3.x = (o[3].y - o[3].x) * h; o3.y = -(o[3].x + o[3].y) * h;
  } else {         // W8 = (1 + i)/sqrt2, W8^3 = (-1 + i)/sqrt2
    o1.x = (o[1].x - o[1].y) * h; o1.y = (o[1].x + o[1].y) * h;
    o3.x = -(o[3].x + o[3].y) * h; o3.y = (o[3].x - o[3].y) * h;
  }
  C o2 = mul_si<SIGN>(o[2]);
  v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
}

template <int SIGN, typename T, typename C>
__device__ __forceinline__ void bfly3(C* v) {
  const T s = (T)0.86602540378443864676;  // sin(2 pi / 3)
  C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
  C y0 = cadd(v[0], t1);
  C m = csub(v[0], cscale(t1, (T)0.5));
  C n = cscale(mul_si<SIGN>(t2), s);
  v[0] = y0; v[1] = cadd(m, n); v[2] = csub(m, n);
}

template <int SIGN, typename T, typename C>
__device__ __forceinline__ void bfly5(C* v) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  C y0 = cadd(v[0], cadd(t1, t2));
  C m1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
  C m2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
  C n1 = mul_si<SIGN>(cadd(cscale(t3, s1), cscale(t4, s2)));
  C n2 = mul_si<SIGN>(csub(cscale(t3, s2), cscale(t4, s1)));
  v[0] = y0;
  v[1] = cadd(m1, n1); v[4] = csub(m1, n1);
  v[2] = cadd(m2, n2); v[3] = csub(m2, n2);
}

template <int R, int SIGN, typename T, typename C>
__device__ __forceinline__ void bfly(C* v) {
  if (R == 2) bfly2<SIGN, T>(v);
  else if (R == 3) bfly3<SIGN, T>(v);
  else if (R == 4) bfly4<SIGN, T>(v);
  else if (R == 5) bfly5<SIGN, T>(v);
  else if (R == 8) bfly8<SIGN, T>(v);
}

// One Stockham stage of fixed radix R over P lines (line stride ls).
template <int R, int SIGN, typename T, typename C>
__device__ __forceinline__ void stage_fixed(const C* __restrict__ src, C* __restrict__ dst,
                                            int n, int Ns, int P, int ls,
                                            const C* __restrict__ tw, int tid, int nt) {
  const int nb = n / R;
  const int total = P * nb;
  const int twstep = n / (Ns * R);
  for (int b = tid; b < total; b += nt) {
    const int line = b / nb;
    const int j = b - line * nb;
    const int k = j % Ns;
    const C* s = src + line * ls;
    C* d = dst + line * ls;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * nb];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = tw[r * k * twstep];
        if (SIGN > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    bfly<R, SIGN, T>(v);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[base + r * Ns] = v[r];
  }
}

// Generic odd radix R <= 15 (direct DFT_R from the size-n twiddle table).
template <int SIGN, typename T, typename C>
__device__ void stage_generic(const C* __restrict__ src, C* __restrict__ dst, int R,
                              int n, int Ns, int P, int ls, const C* __restrict__ tw,
                              int tid, int nt) {
  const int nb = n / R;
  const int total = P * nb;
  const int twstep = n / (Ns * R);
  const int rstep = n / R;
  for (int b = tid; b < total; b += nt) {
    const int line = b / nb;
    const int j = b - line * nb;
    const int k = j % Ns;
    const C* s = src + line * ls;
    C* d = dst + line * ls;
    C v[15];
    for (int r = 0; r < R; ++r) {
      C x = s[j + r * nb];
      if (Ns > 1 && r > 0) {
        C w = tw[r * k * twstep];
        if (SIGN > 0) w.y = -w.y;
        x = cmul(x, w);
      }
      v[r] = x;
    }
    const int base = (j - k) * R + k;
    for (int q = 0; q < R; ++q) {
      C acc = v[0];
      for (int r = 1; r < R; ++r) {
        C w = tw[((r * q) % R) * rstep];
        if (SIGN > 0) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      d[base + q * Ns] = acc;
    }
  }
}

// Up to 8 stages; radices in rad[0..nst).
struct FftPlanDev {
  int n;
  int nst;
  int rad[8];
};

// Transform P lines held in `a` (stride ls); `b` is scratch of equal size.
// Returns the buffer that holds the result.  Ends with a barrier.
template <int SIGN, typename T, typename C>
__device__ C* fft_lines(C* a, C* b, const FftPlanDev& pl, int P, int ls,
                        const C* __restrict__ tw, int tid, int nt) {
  int Ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.rad[s];
    switch (R) {
      case 2: stage_fixed<2, SIGN, T>(a, b, pl.n, Ns, P, ls, tw, tid, nt); break;
      case 3: stage_fixed<3, SIGN, T>(a, b, pl.n, Ns, P, ls, tw, tid, nt); break;
      case 4: stage_fixed<4, SIGN, T>(a, b, pl.n, Ns, P, ls, tw, tid, nt); break;
      case 5: stage_fixed<5, SIGN, T>(a, b, pl.n, Ns, P, ls, tw, tid, nt); break;
      case 8: stage_fixed<8, SIGN, T>(a, b, pl.n, Ns, P, ls, tw, tid, nt); break;
      default: stage_generic<SIGN, T>(a, b, R, pl.n, Ns, P, ls, tw, tid, nt); break;
    }
    __syncthreads();
    C* t = a; a = b; b = t;
    Ns *= R;
  }
  return a;
}

}  // namespace psx
