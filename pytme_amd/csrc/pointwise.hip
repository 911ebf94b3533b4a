// Elementwise and reduction kernels of the per-rotation scoring loop:
// spectral products, FLC normalisation, max/argmax over rotations, fills and
// deterministic FP64 reductions.  All are grid-stride loops over flat
// contiguous arrays; the volumes are HBM-bound so each kernel streams its
// operands exactly once.
#include <stdarg.h>

#include "tm_common.h"

#define TM_PW_THREADS 256
#define TM_RED_BLOCKS 1024

static unsigned pw_blocks(int64_t n) {
  int64_t b = tm_ceil_div(n, (int64_t)TM_PW_THREADS * 4);
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (unsigned)b;
}

__global__ void k_fill_f32(float* __restrict__ p, float v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i32(int* __restrict__ p, int v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// out = a * b (complex), b broadcast with period nb (nb == n: same shape)
__global__ void k_cmul_cc(const float2* __restrict__ a, const float2* __restrict__ b, float2* __restrict__ out,
                          int64_t n, int64_t nb) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = c_mul(a[i], b[nb == n ? i : i % nb]);
}

// out = a * f (complex times real), f broadcast with period nf
__global__ void k_cmul_cr(const float2* __restrict__ a, const float* __restrict__ f, float2* __restrict__ out,
                          int64_t n, int64_t nf) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = c_scale(a[i], f[nf == n ? i : i % nf]);
}

// op: 0 add, 1 subtract, 2 multiply, 3 divide.  b broadcast with period nb
// (nb == 1 is a scalar array); if b is null the host scalar s is used.
__global__ void k_binary_f32(int op, const float* __restrict__ a, const float* __restrict__ b, float s,
                             float* __restrict__ out, int64_t n, int64_t nb) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float x = a[i];
    const float y = b ? b[nb == n ? i : i % nb] : s;
    float r;
    switch (op) {
      case 0: r = x + y; break;
      case 1: r = x - y; break;
      case 2: r = x * y; break;
      default: r = x / y; break;
    }
    out[i] = r;
  }
}

// Reference norm_scores (tme/backends/npfftw_backend.py:537-587), numpy
// semantics: sigma := 1 where sigma < eps.
__device__ __forceinline__ float flc_norm(float x, float ex2, float ex, float n, float eps) {
#pragma clang fp contract(off)
  float m = ex / n;
  m = m * m;
  float v = ex2 / n - m;
  v = fmaxf(v, 0.f);
  float sd = sqrtf(v);
  sd = sd < eps ? 1.f : sd;
  return x / (sd * n);
}

__global__ void k_norm_scores(const float* __restrict__ arr, const float* __restrict__ exp_sq,
                              const float* __restrict__ sq_exp, const double* __restrict__ n_dev, double n_host,
                              float eps, float* __restrict__ out, int64_t n, int64_t n_ex) {
  const float nf = (float)(n_dev ? n_dev[0] : n_host);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = n_ex == n ? i : i % n_ex;
    const float x = arr ? arr[i] : 1.f;
    out[i] = flc_norm(x, exp_sq[j], sq_exp[j], nf, eps);
  }
}

// Reference max_score_over_rotations (npfftw_backend.py:503-535): strict '>'
// so the earliest rotation wins an exact tie.
__global__ void k_max_update(const float* __restrict__ scores, float* __restrict__ max_scores,
                             int* __restrict__ rot, int rot_index, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float s = scores[i];
    if (s > max_scores[i]) {
      max_scores[i] = s;
      rot[i] = rot_index;
    }
  }
}

// ---- deterministic FP64 reductions --------------------------------------
// Stage 1: a fixed grid of TM_RED_BLOCKS blocks; block k sums elements
// k*T + t, k*T + t + G*T, ... (fixed order per thread), then a fixed tree.
// Stage 2: one thread sums the partials in index order.  The result depends
// only on n, never on timing.

template <int K>
__device__ __forceinline__ void block_tree(double (&v)[K], double (*sh)[TM_PW_THREADS]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < K; ++k) sh[k][tid] = v[k];
  __syncthreads();
  for (int s = TM_PW_THREADS / 2; s > 0; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int k = 0; k < K; ++k) sh[k][tid] += sh[k][tid + s];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = sh[k][0];
}

// mode 0: sum x ; mode 1: (sum x*m, sum x*x*m)
__global__ void __launch_bounds__(TM_PW_THREADS)
k_reduce_stage1(int mode, const float* __restrict__ x, const float* __restrict__ m, int64_t n,
                double* __restrict__ partials) {
  __shared__ double sh[2][TM_PW_THREADS];
  double v[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * TM_PW_THREADS;
  for (int64_t i = (int64_t)blockIdx.x * TM_PW_THREADS + threadIdx.x; i < n; i += stride) {
    const double a = x[i];
    if (mode == 0) {
      v[0] += a;
    } else {
      const double w = m[i];
      v[0] += a * w;
      v[1] += a * a * w;
    }
  }
  block_tree<2>(v, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = v[0];
    partials[2 * blockIdx.x + 1] = v[1];
  }
}

// mode 0: out[0] = sum.  mode 1: (mean, std) of normalize_template with
// n = n_dev[0] (or n_host):  mean = S1/n,  std = sqrt(max(S2/n - mean^2, 0)).
__global__ void k_reduce_stage2(int mode, const double* __restrict__ partials, int nparts,
                                const double* __restrict__ n_dev, double n_host, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s0 = 0.0, s1 = 0.0;
  for (int i = 0; i < nparts; ++i) {
    s0 += partials[2 * i];
    s1 += partials[2 * i + 1];
  }
  if (mode == 0) {
    out[0] = s0;
  } else {
    const double nn = n_dev ? n_dev[0] : n_host;
    const double mean = s0 / nn;
    double var = s1 / nn - mean * mean;
    var = var > 0.0 ? var : 0.0;
    out[0] = mean;
    out[1] = sqrt(var);
  }
}

// t <- (t - mean) / std * m, mean/std rounded to float32 once.
__global__ void k_normalize_apply(float* __restrict__ t, const float* __restrict__ m,
                                  const double* __restrict__ ms, int64_t n) {
  const float mean = (float)ms[0];
  const float sd = (float)ms[1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    t[i] = (t[i] - mean) / sd * m[i];
}

extern "C" {

int64_t tm_reduce_workspace(void) { return 2 * (int64_t)TM_RED_BLOCKS + 2; }

int tm_fill_f32(float* p, float v, int64_t n, void* stream) {
  TM_REQUIRE(p || n == 0, "null pointer");
  if (n == 0) return TM_OK;
  hipLaunchKernelGGL(k_fill_f32, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, (hipStream_t)stream, p, v, n);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_fill_i32(int* p, int v, int64_t n, void* stream) {
  TM_REQUIRE(p || n == 0, "null pointer");
  if (n == 0) return TM_OK;
  hipLaunchKernelGGL(k_fill_i32, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, (hipStream_t)stream, p, v, n);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_cmul_cc(const void* a, const void* b, void* out, int64_t n, int64_t nb, void* stream) {
  TM_REQUIRE(a && b && out, "null pointer");
  TM_REQUIRE(nb >= 1 && n % nb == 0, "broadcast period %lld does not divide %lld", (long long)nb, (long long)n);
  TM_TIMED("cmul_cc", 24.0 * n, (hipStream_t)stream,
           hipLaunchKernelGGL(k_cmul_cc, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, (hipStream_t)stream,
                              (const float2*)a, (const float2*)b, (float2*)out, n, nb));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_cmul_cr(const void* a, const float* f, void* out, int64_t n, int64_t nf, void* stream) {
  TM_REQUIRE(a && f && out, "null pointer");
  TM_REQUIRE(nf >= 1 && n % nf == 0, "broadcast period %lld does not divide %lld", (long long)nf, (long long)n);
  TM_TIMED("cmul_cr", 20.0 * n, (hipStream_t)stream,
           hipLaunchKernelGGL(k_cmul_cr, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, (hipStream_t)stream,
                              (const float2*)a, f, (float2*)out, n, nf));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_binary_f32(int op, const float* a, const float* b, float s, float* out, int64_t n, int64_t nb,
                  void* stream) {
  TM_REQUIRE(a && out, "null pointer");
  TM_REQUIRE(op >= 0 && op <= 3, "unknown op %d", op);
  TM_REQUIRE(!b || (nb >= 1 && n % nb == 0), "bad broadcast period");
  hipLaunchKernelGGL(k_binary_f32, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, (hipStream_t)stream, op, a, b, s,
                     out, n, nb);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_norm_scores(const float* arr, const float* exp_sq, const float* sq_exp, const double* n_obs_dev,
                   double n_obs_host, float eps, float* out, int64_t n, int64_t n_ex, void* stream) {
  TM_REQUIRE(exp_sq && sq_exp && out, "null pointer");
  TM_REQUIRE(n_ex >= 1 && n % n_ex == 0, "bad broadcast period");
  TM_TIMED("norm_scores", 16.0 * n, (hipStream_t)stream,
           hipLaunchKernelGGL(k_norm_scores, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, (hipStream_t)stream,
                              arr, exp_sq, sq_exp, n_obs_dev, n_obs_host, eps, out, n, n_ex));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_max_update(const float* scores, float* max_scores, int* rot, int rot_index, int64_t n, void* stream) {
  TM_REQUIRE(scores && max_scores && rot, "null pointer");
  hipLaunchKernelGGL(k_max_update, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, (hipStream_t)stream, scores,
                     max_scores, rot, rot_index, n);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// out_dev[0] = sum(x) in FP64, deterministic; workspace >= tm_reduce_workspace() doubles.
int tm_sum_f64(const float* x, int64_t n, double* out_dev, double* workspace, void* stream) {
  TM_REQUIRE(x && out_dev && workspace, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_reduce_stage1, dim3(TM_RED_BLOCKS), dim3(TM_PW_THREADS), 0, s, 0, x, (const float*)nullptr,
                     n, workspace);
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(64), 0, s, 0, workspace, TM_RED_BLOCKS, (const double*)nullptr,
                     0.0, out_dev);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// Reference normalize_template (tme/matching_utils.py:61-99) over a flat
// array: moments in FP64 (deterministic), the affine update in FP32.
int tm_normalize_template(float* t, const float* m, int64_t n, const double* n_obs_dev, double n_obs_host,
                          double* workspace, void* stream) {
  TM_REQUIRE(t && m && workspace, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  double* ms = workspace + 2 * TM_RED_BLOCKS;
  hipLaunchKernelGGL(k_reduce_stage1, dim3(TM_RED_BLOCKS), dim3(TM_PW_THREADS), 0, s, 1, (const float*)t, m, n,
                     workspace);
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(64), 0, s, 1, workspace, TM_RED_BLOCKS, n_obs_dev, n_obs_host,
                     ms);
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_normalize_apply, dim3(pw_blocks(n)), dim3(TM_PW_THREADS), 0, s, t, m, (const double*)ms, n);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

}  // extern "C"
