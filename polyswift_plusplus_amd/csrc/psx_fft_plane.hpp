// psx_fft_plane.hpp -- both in-plane axes of a contour step in one launch.
//
// A 3D grid (n_x, n, n) stores each x-plane as one contiguous n x n block
// (z fastest).  The diffusion operator of a step is the product of three
// commuting 1D round trips (z, y, x), so the z and y round trips of one
// plane need no data from any other plane: a block stages its plane in LDS,
// runs the z round trips (two rows per complex line) and the y round trips
// (two columns per complex line) with the register line FFT of
// psx_fft_reg.hpp, and writes the plane back.  Like the fused line launches,
// one launch may join two contour steps: the plane's result of the first
// step is multiplied by post, stored (mid), multiplied by pre2 and carried on
// in LDS through the second step's y and z round trips.
//
// Against the line launches this moves 4 fields per contour step instead of
// 6: the step's x round trip stays a (fused) strided line launch, and the y
// and z round trips of a step pair become one plane launch.
#pragma once
#include "psx_fft_reg.hpp"

namespace psx {

// Spectral factors of a plane launch: emul (and emul2) point at 4n values,
// the z factors and the y factors (each includes 1/n), then their (a, b)
// rows for lane-pair splits (PairAB).
//
// NW waves of 64 / N2 complex lines each cover the plane's n / 2 complex
// lines; where they do not divide evenly (n = 48: 2 waves of 16 for 24
// lines) the last wave's extra lanes join its transposes but load and store
// nothing.  REUSE: the waves' transpose buffers alias the
// plane (block barriers around each pass) so the block needs only the plane's
// LDS; otherwise they sit behind it and a pass needs no barrier of its own
// (every wave reads and writes only its own lines of the plane).
template <typename T, int N1, int N2, int NW, bool REUSE>
struct PlaneShape {
  static constexpr int n = N1 * N2;
  static constexpr int CL = 64 / N2;                       // complex lines per wave
  // LDS row stride (elements): 4 more than n.  FP64 measured fewer bank
  // conflicts than with 2 (128^3 plane launch 19.3 -> 19.1 us); FP32 at
  // n = 128 measured twice the conflicts with 8.
  static constexpr int RS = n + 4;
  static constexpr int PB = n * RS * (int)sizeof(T);       // plane bytes
  static constexpr int TB = NW * TBuf<N1, N2>::wave * 2 * (int)sizeof(T);   // transpose buffers
  static constexpr int BUF = ((REUSE ? (PB > TB ? PB : TB) : PB + TB) + 15) / 16 * 16;
  // + twiddles (n complex) + up to four n-point spectral rows
  static constexpr int SMEM = BUF + table_bytes<T, N1, N2>(4);
  static_assert(2 * NW * CL >= n && 2 * (NW - 1) * CL < n, "the waves must cover the plane's complex lines");
};

// Round trips along y, then z (second step: z, then y) of one x-plane per
// block.  g.ntiles = planes per field (n_x), g.nfields = fields; PairIO as
// for the line launches with the plane spectral tables above.
// DIRECT: the first y round trips load their column pairs straight from
// global memory (no staging phase: each wave starts as soon as its own
// columns arrive) and the last ones store theirs straight back (no final
// barrier or staging pass).  A wave's column pairs are 2 CL elements of a
// row, so this suits splits whose CL pairs span 64 bytes or more.
// WPE > 0: compiled for at least WPE waves per SIMD (a register budget of
// 512 / WPE), where that lets more blocks share a CU without spills.
template <typename T, int N1, int N2, int NW, bool REUSE, bool DIRECT = false, int WPE = 0>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 64 * NW), amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
k_plane(LineGeom g, const PairIO* __restrict__ fio) {
  typedef typename Cx<T>::type C;
  typedef typename Vec16<T>::type V;
  typedef PlaneShape<T, N1, N2, NW, REUSE> S;
  constexpr int n = S::n, CL = S::CL, RS = S::RS;
  constexpr int VW = 16 / (int)sizeof(T);
  constexpr int CPR = n / VW;          // 16-byte chunks per plane row
  constexpr int NCH = n * CPR;
  constexpr int IT = (NCH + 64 * NW - 1) / (64 * NW);   // chunks per thread, all in flight at once
  constexpr bool CPART = IT * 64 * NW != NCH;             // the last round of chunks is partial
  constexpr bool PART = 2 * NW * CL != n;                 // lanes past the plane's lines
  static_assert(n % VW == 0 && RS % VW == 0, "plane rows must be whole 16-byte chunks");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* P = (T*)smem_raw;
  const int f = blockIdx.x / g.ntiles, x = blockIdx.x - f * g.ntiles;
  const PairIO fd = fio[f];
  const bool fused = fd.emul2 != nullptr;
  const long long base = (long long)x * n * n;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int p = lane / N2, t = lane % N2;
  const int cl = wave * CL + p;        // this lane's complex line: rows / columns 2cl, 2cl+1
  const bool lv = !PART || 2 * cl < n;
  C* L = (C*)(smem_raw + (REUSE ? 0 : S::PB)) + wave * TBuf<N1, N2>::wave + p * TBuf<N1, N2>::line;
  const C* __restrict__ tw = (const C*)g.tw;
  C* twl = (C*)(smem_raw + S::BUF);
  T* em = (T*)(twl + n);               // e_z, e_y of the first step, then of the second
  // global row r of the round trips (0, 1: z, y of the first step; 2, 3: of
  // the second): the factors, or for lane-pair splits their (a, b) rows
  auto grow = [&](int r) -> const T* {
    const T* b = (const T*)(r < 2 ? fd.emul : fd.emul2);
    return b + (pair_split<N1, N2>() ? 2 * n : 0) + (r & 1) * n;
  };
  C vd[DIRECT ? N1 : 1];
  if (DIRECT) {
    // the first y round trips' column pairs, times pre, all in flight
    const T* __restrict__ src = (const T*)fd.in;
    const T* __restrict__ pre = (const T*)fd.pre;
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const long long ga = base + (long long)(t + N2 * r) * n + 2 * cl;
      if (lv) vd[r] = *(const C*)(src + ga);
      else { vd[r].x = (T)0; vd[r].y = (T)0; }
    }
    if (pre && lv) {
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        const C w = *(const C*)(pre + base + (long long)(t + N2 * r) * n + 2 * cl);
        vd[r].x *= w.x;
        vd[r].y *= w.y;
      }
    }
  }
  for (int i = tid; i < n; i += 64 * NW) twl[i] = tw[i];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (r >= 2 && !fused) break;
    const T* src = grow(r);
    for (int i = tid; i < n; i += 64 * NW) em[r * n + i] = src[i];
  }
  // chunk i of this thread: row = c / CPR, column (c % CPR) * VW, c = tid + i 64 NW;
  // false past the plane's last chunk
  // The staging load and the final store walk the same chunks.  The final
  // store reads the thread index through an opaque move, so the compiler
  // recomputes the chunk addresses there instead of keeping them live
  // through every round trip (FP32 (12, 8): 112 -> 88 VGPRs, FP64 (16, 8):
  // 246 -> 184).
  int tidc = tid;
  auto chunk = [&](int i, int* row, int* col) -> bool {
    const int c = tidc + i * 64 * NW;
    *row = c / CPR;
    *col = (c - *row * CPR) * VW;
    return !CPART || c < NCH;
  };
  if (DIRECT) {
    // the first y round trips, each wave as soon as its own columns arrive:
    // twiddles and factors from global memory (their LDS copies are not yet
    // visible), then the plane into LDS
    line_roundtrip<T, N1, N2, true>(vd, L, tw, tw, grow(1), t);
    if (REUSE) __syncthreads();
    else lds_sync<true>();
    if (lv) {
#pragma unroll
      for (int r = 0; r < N1; ++r) *(C*)(P + (t + N2 * r) * RS + 2 * cl) = vd[r];
    }
  } else {
    const T* __restrict__ src = (const T*)fd.in;
    const T* __restrict__ pre = (const T*)fd.pre;
    V v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      int row, col;
      if (chunk(i, &row, &col)) v[i] = *(const V*)(src + base + (long long)row * n + col);
    }
    if (pre) {
#pragma unroll
      for (int i = 0; i < IT; ++i) {
        int row, col;
        if (!chunk(i, &row, &col)) continue;
        const V w = *(const V*)(pre + base + (long long)row * n + col);
        vmul(v[i], w);
      }
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      int row, col;
      if (chunk(i, &row, &col)) vstore_lds(P + row * RS + col, v[i]);
    }
  }
  __syncthreads();

  auto ypass = [&](const T* ey) {
    C v[N1];
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      if (lv) v[r] = *(const C*)(P + (t + N2 * r) * RS + 2 * cl);
      else { v[r].x = (T)0; v[r].y = (T)0; }
    }
    if (REUSE) __syncthreads();
    line_roundtrip<T, N1, N2, true>(v, L, tw, twl, ey, t);
    if (REUSE) __syncthreads();
    else lds_sync<true>();
    if (lv) {
#pragma unroll
      for (int r = 0; r < N1; ++r) *(C*)(P + (t + N2 * r) * RS + 2 * cl) = v[r];
    }
  };

  // The step pair runs as y, z | z, y (the launch list's own op order): the
  // step boundary sits between the two z round trips of a wave's own rows,
  // so post, the store of mid and pre2 happen in registers, and each wave's
  // w-factor loads and stores overlap the other waves' arithmetic.
  if (!DIRECT) {
    ypass(em + n);
    __syncthreads();
  }
  {
    C v[N1];
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const int k = t + N2 * r;
      v[r].x = lv ? P[(2 * cl) * RS + k] : (T)0;
      v[r].y = lv ? P[(2 * cl + 1) * RS + k] : (T)0;
    }
    if (REUSE) __syncthreads();
    line_roundtrip<T, N1, N2, true>(v, L, tw, twl, em, t);
    // register r of this lane: points t + N2 r of rows 2cl (x) and 2cl + 1 (y)
    const long long ra = base + (long long)(2 * cl) * n + t;
    const T* __restrict__ post = (const T*)fd.post;
    T* __restrict__ dst = (T*)(fused ? fd.mid : fd.out);
    T wx[N1], wy[N1];
    if (post && lv) {
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        wx[r] = post[ra + N2 * r];
        wy[r] = post[ra + n + N2 * r];
      }
#pragma unroll
      for (int r = 0; r < N1; ++r) { v[r].x *= wx[r]; v[r].y *= wy[r]; }
    }
    if (lv) {
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        dst[ra + N2 * r] = v[r].x;
        dst[ra + n + N2 * r] = v[r].y;
      }
    }
    if (!fused) return;    // uniform over the block
    const T* __restrict__ pre2 = (const T*)fd.pre2;
    if (pre2 && lv) {
      if (pre2 != post) {
#pragma unroll
        for (int r = 0; r < N1; ++r) {
          wx[r] = pre2[ra + N2 * r];
          wy[r] = pre2[ra + n + N2 * r];
        }
      }
#pragma unroll
      for (int r = 0; r < N1; ++r) { v[r].x *= wx[r]; v[r].y *= wy[r]; }
    }
    line_roundtrip<T, N1, N2, true>(v, L, tw, twl, em + 2 * n, t);
    if (REUSE) __syncthreads();
    else lds_sync<true>();
    if (lv) {
#pragma unroll
      for (int r = 0; r < N1; ++r) {
        const int k = t + N2 * r;
        P[(2 * cl) * RS + k] = v[r].x;
        P[(2 * cl + 1) * RS + k] = v[r].y;
      }
    }
  }
  __syncthreads();
  if (DIRECT) {
    // the last y round trips, straight to global memory (times post2)
    C v[N1];
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      if (lv) v[r] = *(const C*)(P + (t + N2 * r) * RS + 2 * cl);
      else { v[r].x = (T)0; v[r].y = (T)0; }
    }
    if (REUSE) __syncthreads();
    line_roundtrip<T, N1, N2, true>(v, L, tw, twl, em + 3 * n, t);
    if (!lv) return;
    const T* __restrict__ post2 = (const T*)fd.post2;
    T* __restrict__ out = (T*)fd.out;
#pragma unroll
    for (int r = 0; r < N1; ++r) {
      const long long ga = base + (long long)(t + N2 * r) * n + 2 * cl;
      C o = v[r];
      if (post2) { const C w = *(const C*)(post2 + ga); o.x *= w.x; o.y *= w.y; }
      *(C*)(out + ga) = o;
    }
    return;
  }
  ypass(em + 3 * n);
  __syncthreads();
  asm volatile("v_mov_b32 %0, %1" : "=v"(tidc) : "v"(tid));
  {
    const T* __restrict__ post2 = (const T*)fd.post2;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      int row, col;
      if (!chunk(i, &row, &col)) continue;
      const long long ga = base + (long long)row * n + col;
      V v = vload_lds<T>(P + row * RS + col);
      if (post2) { const V w = *(const V*)(post2 + ga); vmul(v, w); }
      *(V*)((T*)fd.out + ga) = v;
    }
  }
}

}  // namespace psx
