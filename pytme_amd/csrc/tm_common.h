// Shared helpers for the pytme_amd HIP library (gfx950 / CDNA4).
//
// Every exported entry point returns an int status (TM_OK on success) and
// records a human-readable message retrievable with tm_last_error().  The
// Python wrapper maps TM_ERR_OOM to MemoryError, TM_ERR_VALUE to ValueError
// and everything else to RuntimeError.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#define TM_OK 0
#define TM_ERR_VALUE 1
#define TM_ERR_OOM 2
#define TM_ERR_HIP 3

void tm_set_error(const char* fmt, ...);

#define TM_CHECK_HIP(expr)                                                   \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      tm_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),    \
                   __FILE__, __LINE__);                                      \
      return _e == hipErrorOutOfMemory ? TM_ERR_OOM : TM_ERR_HIP;            \
    }                                                                        \
  } while (0)

#define TM_CHECK_LAUNCH() TM_CHECK_HIP(hipGetLastError())

#define TM_REQUIRE(cond, ...)        \
  do {                               \
    if (!(cond)) {                   \
      tm_set_error(__VA_ARGS__);     \
      return TM_ERR_VALUE;           \
    }                                \
  } while (0)

static inline int64_t tm_ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- complex helpers (float2 = interleaved complex64) ----
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_add(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 c_scale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }

// ---- mirror-pair order of half-spectrum rows -------------------------------
// A half-spectrum row of M + 1 bins (real length 2M) stored with its first
// nm positions (a multiple of 2h) in tiles of 2h positions, tile j holding
// bins h*j .. h*j + h - 1 in its first half and their mirrors M - h*j ..
// M - h*j - h + 1 in its second half, so position c and c ^ h of a tile hold
// a pair (k, M - k); the remaining M + 1 - nm positions hold the middle bins
// nm/2 .. M - nm/2 in order.  A transform along another axis that sees whole
// tiles then has both bins of every pair in lanes l and l ^ h.
//
// h == 1 (radix-4 units; M = 4 G with G odd, nm = 4 (G - 1)): unit u of 8
// positions holds the first-stage groups j = u + 1 and G - j of the row's
// M-point inverse whose first stage is radix 4 (bins j + G r at positions
// 8 u + r and (G - j) + G r at 8 u + 4 + r, r = 0..3), so positions c and
// 7 - c of a unit hold a pair (k, M - k) and each group sits in one lane
// quad; the remaining positions hold bins 0, G, 2G, 3G and 4G = M.
__host__ __device__ __forceinline__ int tm_pair_bin(int q, int nm, int M, int h) {
  if (h == 1) {
    const int G = M / 4;
    if (q >= nm) return G * (q - nm);
    const int c = q & 7, j = (q >> 3) + 1;
    return c < 4 ? j + G * c : (G - j) + G * (c - 4);
  }
  if (q >= nm) return nm / 2 + (q - nm);
  const int j = q / (2 * h), c = q - j * 2 * h;
  return c < h ? h * j + c : M - h * j - (c - h);
}
__host__ __device__ __forceinline__ int tm_pair_pos(int k, int nm, int M, int h) {
  if (h == 1) {
    const int G = M / 4, j0 = k % G, r = k / G;
    if (j0 == 0) return nm + r;
    return j0 <= (G - 1) / 2 ? 8 * (j0 - 1) + r : 8 * (G - j0 - 1) + 4 + r;
  }
  const int np = nm / 2;
  if (k < np) return (k / h) * 2 * h + k % h;
  if (M - k < np) return ((M - k) / h) * 2 * h + h + (M - k) % h;
  return nm + (k - np);
}

// The M-point complex input of a row's C2R (real length N2 = 2M) at bin k:
// Z[k] = (X[k] + X*[M-k]) + i W^k (X[k] - X*[M-k]), W = exp(2 pi i / N2),
// from a = X[k], b = X[M-k] and w = W^k; the imaginary parts of the DC and
// Nyquist bins are dropped (k == 0), as numpy's irfft does.  The inverse
// M-point transform of Z holds the outputs 2n and 2n + 1 in its real and
// imaginary parts.
__device__ __forceinline__ float2 tm_c2r_pre(float2 a, float2 b, float2 w, bool dc) {
  a.y = dc ? 0.f : a.y;
  b.y = dc ? 0.f : b.y;
  const float2 e = make_float2(a.x + b.x, a.y - b.y);
  const float2 d = c_mul(make_float2(a.x - b.x, a.y + b.y), w);
  return make_float2(e.x - d.y, e.y + d.x);
}

// Lane m ^ X within a lane quad (DPP quad_perm).
template <int X>
__device__ __forceinline__ float tm_quad_xor(float v) {
  constexpr int sel = (0 ^ X) | ((1 ^ X) << 2) | ((2 ^ X) << 4) | ((3 ^ X) << 6);
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), sel, 0xf, 0xf, false));
}

// Inverse radix-4 DFT across a lane quad: lane m holds z_m on entry and
// y_r = sum_n z_n i^(r n) on return for r = (m >> 1) | ((m & 1) << 1) (a
// radix-2 step over lanes m, m ^ 2, lane 3's difference times i, then one
// over m, m ^ 1).  Every lane of the quad must take part.
__device__ __forceinline__ float2 tm_quad_idft4(float2 z, int m) {
  const float sa = (m & 2) ? -1.f : 1.f, sb = (m & 1) ? -1.f : 1.f;
  float2 t = make_float2(tm_quad_xor<2>(z.x) + sa * z.x, tm_quad_xor<2>(z.y) + sa * z.y);
  if (m == 3) t = make_float2(-t.y, t.x);
  return make_float2(tm_quad_xor<1>(t.x) + sb * t.x, tm_quad_xor<1>(t.y) + sb * t.y);
}

// ---- streaming global accesses --------------------------------------------
// Non-temporal cache policy for data that is larger than every cache and
// touched once per pass (the correlation spectra of the fused scan, full
// volumes of the plain FFT passes).  A hint only: results are unchanged.
// ON is false for tiles whose row segments are narrower than a 128-byte
// cache line: with the hint, partially written lines stream at well under
// half the rate (a 1024-point pass over 64-byte segments: 7.0 -> 11.0 ms).
#ifndef TM_STREAM_HINTS
#define TM_STREAM_HINTS 1
#endif
typedef float tm_f2v __attribute__((ext_vector_type(2)));
template <bool ON = true>
__device__ __forceinline__ float2 tm_ld_stream(const float2* p) {
#if TM_STREAM_HINTS
  if constexpr (ON) {
    const tm_f2v v = __builtin_nontemporal_load(reinterpret_cast<const tm_f2v*>(p));
    return make_float2(v.x, v.y);
  }
#endif
  return *p;
}
// Two consecutive complex values (16-byte aligned) in one non-temporal load.
typedef float tm_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 tm_ld_stream2(const float2* p) {
#if TM_STREAM_HINTS
  const tm_f4v v = __builtin_nontemporal_load(reinterpret_cast<const tm_f4v*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
#else
  return *reinterpret_cast<const float4*>(p);
#endif
}
template <bool ON = true>
__device__ __forceinline__ void tm_st_stream(float2* p, float2 a) {
#if TM_STREAM_HINTS
  if constexpr (ON) {
    const tm_f2v v = {a.x, a.y};
    __builtin_nontemporal_store(v, reinterpret_cast<tm_f2v*>(p));
    return;
  }
#endif
  *p = a;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block i on XCD i % 8).
// Renumbering so that consecutive logical tiles run on one XCD keeps the
// neighbouring column tiles of a row (which share DRAM pages) together.
#ifndef TM_XCD_REMAP
#define TM_XCD_REMAP 1
#endif
__device__ __forceinline__ int tm_xcd_block() {
  const int n = gridDim.x, b = blockIdx.x;
  if (!TM_XCD_REMAP || (n & 7)) return b;
  return (b & 7) * (n >> 3) + (b >> 3);
}

// ---- optional per-kernel timing (HIP events around launches) -------------
// When enabled with tm_timing_enable(1), TM_TIMED launches record a start and
// stop event on the launch stream and remember the kernel's algorithmic HBM
// bytes, so the host can report time and bytes per kernel name.
bool tm_timing_on();
hipEvent_t tm_timing_event();
void tm_timing_push(const char* name, hipEvent_t a, hipEvent_t b, double bytes);

#define TM_TIMED(name, bytes, stream, ...)                      \
  do {                                                          \
    hipEvent_t _ta = nullptr;                                   \
    if (tm_timing_on()) {                                       \
      _ta = tm_timing_event();                                  \
      hipEventRecord(_ta, (stream));                            \
    }                                                           \
    __VA_ARGS__;                                                \
    if (_ta) {                                                  \
      hipEvent_t _tb = tm_timing_event();                       \
      hipEventRecord(_tb, (stream));                            \
      tm_timing_push((name), _ta, _tb, (double)(bytes));        \
    }                                                           \
  } while (0)
