// FFT plan structure and pass launchers shared by the FFT entry points and
// the fused scan kernels.
#pragma once
#include "fft_core.h"
#include "tm_common.h"

#define TM_FFT_THREADS 256
#define TM_FFT_MAX_LEN 4096

struct tm_line {
  LinePlan lp;
  float2* tw_dev;
};

struct tm_fft_plan {
  int ndim;
  int64_t shape[3];   // real shape
  int64_t cshape[3];  // half-spectrum shape
  int64_t batch;
  tm_line lines[3];
};

// C2C along the middle axis of [outer][lin][ld] -> [outer][n][ld] for the
// first `inner` <= ld inner indices (lin <= n zero-padded; dir -1 forward,
// +1 inverse unnormalised).  Inner indices inner..ld-1 are not touched.
// `variant` picks the radix plan of the compile-time kernels (fft_ct.h:
// ct_factor_v); the generic path ignores it.
int tm_launch_c2c(const LinePlan& lp, const float2* in, int lin, float2* out, int64_t outer, int64_t inner,
                  int64_t ld, int dir, float scale, hipStream_t s, const char* name, int variant = 0);
// In-place inverse C2C along the middle axis of [outer][n][ld] (first
// `inner` lines, whole tiles of 2 ph lines) whose lines are mirror-pair
// rows (tm_pair_bin with h = ph) of a half spectrum of M + 1 bins: each
// output is stored as the C2R input tm_c2r_pre of its bin (twx: the 2M
// twiddles of the row transform).  Lengths 128, 256, 512 (ph 8), 1024 (ph 4).
// ph 1: radix-4 units (tm_pair_bin h = 1) for lengths 128, 256 and 512; the
// stored values are then the first-stage outputs of the row's inverse.
int tm_launch_c2c_pairs(const LinePlan& lp, const float2* twx, int M, int ph, float2* buf, int64_t outer,
                        int64_t inner, int64_t ld, hipStream_t s, const char* name, int variant = 0);
// Compute units of the current device (cached per device).
int tm_device_cus();
// The ph of tm_launch_c2c_pairs for length n, 0 if it has no kernel.
int tm_c2c_pair_h(int n);
// R2C of nrows contiguous rows of lin <= n reals (zero-padded to n) into
// rows of n/2+1 bins at a row stride of ldo >= n/2+1 complex; pair_h > 0
// stores the bins in the mirror-pair order (tm_pair_bin) with pair_nm
// paired positions.
int tm_launch_r2c_rows(const LinePlan& lp, const float* in, int lin, float2* out, int64_t nrows, int64_t ldo,
                       hipStream_t s, const char* name, int pair_nm = 0, int pair_h = 0);
// The same transform by the compile-time kernels (one half-length complex
// transform per row and the split into bins) for n = 128, 200 and 256 in
// radix plan `variant`; other lengths go to tm_launch_r2c_rows.
int tm_launch_r2c_rows_ct(const LinePlan& lp, const float* in, int lin, float2* out, int64_t nrows, int64_t ldo,
                          hipStream_t s, const char* name, int pair_nm, int pair_h, int variant);
// C2R of nrows contiguous half-spectrum rows, times scale.
int tm_launch_c2r_rows(const LinePlan& lp, const float2* in, float* out, int64_t nrows, float scale, hipStream_t s,
                       const char* name);
// Whole n-D transforms of `batch` contiguous arrays with plan p.
int tm_fft_r2c_batch(const tm_fft_plan* p, int64_t batch, const float* in, float2* out, hipStream_t s);
int tm_fft_c2r_batch(const tm_fft_plan* p, int64_t batch, float2* in, float* out, hipStream_t s);
