// In-LDS mixed-radix Stockham FFT over a tile of lines.
//
// A "line plan" factors a transform length n into radices {8,4,2,5,3} plus
// generic small primes (7, 11, 13).  The tile of L lines lives in LDS with a
// leading dimension ld (>= n); two LDS buffers ping-pong between stages so a
// stage never reads what it writes.  Twiddles come from a per-length table
// tw[k] = exp(-2*pi*i*k/n) computed on the host in double precision and
// rounded once to float, which keeps the FP32 transform error at the level
// of the arithmetic rather than the twiddle generation.
//
// Stage s with radix R and Ns = prod(radix[0..s-1]) maps butterfly j
// (0 <= j < n/R) from inputs a[j + r*n/R] to outputs
// b[(j/Ns)*Ns*R + j%Ns + r*Ns] after twiddling input r by W_{Ns*R}^{r*(j%Ns)}
// (the self-sorting Stockham formulation: no bit reversal pass).
#pragma once
#include "tm_common.h"

#define TM_MAX_STAGES 14
#define TM_MAX_GENERIC_RADIX 13

// Workgroup barrier ordering LDS only: waits for this wave's LDS operations,
// not for its outstanding global loads and stores, so prefetched loads and
// earlier stores stay in flight across the barrier.  (__syncthreads() also
// drains vector memory.)
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

struct LinePlan {
  int n;
  int nst;
  int radix[TM_MAX_STAGES];
  const float2* tw;  // n entries, exp(-2 pi i k / n)
};

// x * (s*i) with s = -1 (forward) or +1 (inverse)
template <int DIR>
__device__ __forceinline__ float2 rot90(float2 x) {
  return DIR < 0 ? make_float2(x.y, -x.x) : make_float2(-x.y, x.x);
}

template <int DIR>
__device__ __forceinline__ float2 tw_load(const float2* tw, int idx) {
  float2 w = tw[idx];
  return DIR < 0 ? w : make_float2(w.x, -w.y);
}

template <int DIR>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = c_add(a, b);
  v[1] = c_sub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
  float2 a0 = c_add(v0, v2), a1 = c_sub(v0, v2);
  float2 a2 = c_add(v1, v3), a3 = rot90<DIR>(c_sub(v1, v3));
  v0 = c_add(a0, a2);
  v1 = c_add(a1, a3);
  v2 = c_sub(a0, a2);
  v3 = c_sub(a1, a3);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  const float h = 0.70710678118654752440f;
  float2 a0 = c_add(v[0], v[4]), b0 = c_sub(v[0], v[4]);
  float2 a1 = c_add(v[1], v[5]), b1 = c_sub(v[1], v[5]);
  float2 a2 = c_add(v[2], v[6]), b2 = c_sub(v[2], v[6]);
  float2 a3 = c_add(v[3], v[7]), b3 = c_sub(v[3], v[7]);
  const float s = (float)DIR;
  // b1 *= (1 + s i)/sqrt2 ; b2 *= s i ; b3 *= (-1 + s i)/sqrt2
  b1 = make_float2(h * (b1.x - s * b1.y), h * (b1.y + s * b1.x));
  b2 = rot90<DIR>(b2);
  b3 = make_float2(h * (-b3.x - s * b3.y), h * (-b3.y + s * b3.x));
  dft4<DIR>(a0, a1, a2, a3);
  dft4<DIR>(b0, b1, b2, b3);
  v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
  v[1] = b0; v[3] = b1; v[5] = b2; v[7] = b3;
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* v) {
  const float s3 = 0.86602540378443864676f;
  float2 t1 = c_add(v[1], v[2]);
  float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  float2 t3 = rot90<DIR>(c_scale(c_sub(v[1], v[2]), s3));
  v[0] = c_add(v[0], t1);
  v[1] = c_add(t2, t3);
  v[2] = c_sub(t2, t3);
}

template <int DIR>
__device__ __forceinline__ void dft5(float2* v) {
  const float c1 = 0.30901699437494742410f;   // cos(2pi/5)
  const float c2 = -0.80901699437494742410f;  // cos(4pi/5)
  const float s1 = 0.95105651629515357212f;   // sin(2pi/5)
  const float s2 = 0.58778525229247312917f;   // sin(4pi/5)
  float2 a1 = c_add(v[1], v[4]), b1 = c_sub(v[1], v[4]);
  float2 a2 = c_add(v[2], v[3]), b2 = c_sub(v[2], v[3]);
  float2 x0 = v[0];
  float2 r1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  float2 r2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  float2 i1 = rot90<DIR>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  float2 i2 = rot90<DIR>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
  v[1] = c_add(r1, i1);
  v[4] = c_sub(r1, i1);
  v[2] = c_add(r2, i2);
  v[3] = c_sub(r2, i2);
}

// Generic odd-prime DFT: out[q] = sum_r v[r] W_R^{rq}; W_R^m = tw[m*(n/R)].
template <int DIR>
__device__ __forceinline__ void dft_generic(float2* v, int R, const float2* tw, int n) {
  float2 out[TM_MAX_GENERIC_RADIX];
  const int step = n / R;
  for (int q = 0; q < R; ++q) {
    float2 acc = v[0];
    int m = 0;
    for (int r = 1; r < R; ++r) {
      m += q;
      if (m >= R) m -= R;
      acc = c_add(acc, c_mul(v[r], tw_load<DIR>(tw, m * step)));
    }
    out[q] = acc;
  }
  for (int q = 0; q < R; ++q) v[q] = out[q];
}

// One radix stage for butterfly j of one line.
template <int DIR, int R>
__device__ __forceinline__ void stage_fixed(const float2* __restrict__ src, float2* __restrict__ dst,
                                            int j, int nb, int Ns, int ts, const float2* __restrict__ tw,
                                            int es) {
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src[(j + r * nb) * es];
  const int k = j % Ns;
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], tw_load<DIR>(tw, r * k * ts));
  }
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 3) dft3<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5) dft5<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  const int base = (j - k) * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[(base + r * Ns) * es] = v[r];
}

template <int DIR>
__device__ __forceinline__ void stage_generic(const float2* __restrict__ src, float2* __restrict__ dst,
                                              int j, int nb, int Ns, int ts, int R, const float2* __restrict__ tw,
                                              int n, int es) {
  float2 v[TM_MAX_GENERIC_RADIX];
  for (int r = 0; r < R; ++r) v[r] = src[(j + r * nb) * es];
  const int k = j % Ns;
  if (Ns > 1)
    for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], tw_load<DIR>(tw, r * k * ts));
  dft_generic<DIR>(v, R, tw, n);
  const int base = (j - k) * R + k;
  for (int r = 0; r < R; ++r) dst[(base + r * Ns) * es] = v[r];
}

// Transform L lines in LDS; element e of line l lives at a[l*ls + e*es].
// LFAST selects the thread->butterfly mapping: consecutive threads take
// consecutive lines (for the element-major [e][l] layout, ls = 1, es = L) or
// consecutive butterflies of one line (line-major [l][e] layout, es = 1).
// Both keep neighbouring lanes on neighbouring LDS words for the loads.
// All threads of the block must call this (it synchronises); the caller
// synchronises before the call so the tile is fully loaded.  Returns the
// buffer holding the result (a or b).
template <int DIR, bool LFAST>
__device__ float2* lds_fft_lines(float2* a, float2* b, int L, int ls, int es, const LinePlan& p,
                                 int tid, int nthr) {
  const int n = p.n;
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.radix[s];
    const int nb = n / R;
    const int ts = n / (Ns * R);
    const int total = nb * L;
    for (int t = tid; t < total; t += nthr) {
      int l, j;
      if (LFAST) {
        j = t / L;
        l = t - j * L;
      } else {
        l = t / nb;
        j = t - l * nb;
      }
      const float2* src = a + l * ls;
      float2* dst = b + l * ls;
      switch (R) {
        case 8: stage_fixed<DIR, 8>(src, dst, j, nb, Ns, ts, p.tw, es); break;
        case 4: stage_fixed<DIR, 4>(src, dst, j, nb, Ns, ts, p.tw, es); break;
        case 2: stage_fixed<DIR, 2>(src, dst, j, nb, Ns, ts, p.tw, es); break;
        case 5: stage_fixed<DIR, 5>(src, dst, j, nb, Ns, ts, p.tw, es); break;
        case 3: stage_fixed<DIR, 3>(src, dst, j, nb, Ns, ts, p.tw, es); break;
        default: stage_generic<DIR>(src, dst, j, nb, Ns, ts, R, p.tw, n, es); break;
      }
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}
