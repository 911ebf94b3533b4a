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
