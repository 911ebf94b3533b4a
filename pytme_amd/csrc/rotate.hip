// Trilinear (spline order 1) rigid rotation of a template and its mask.
//
// Semantics follow scipy.ndimage.affine_transform(order=1, mode="constant",
// cval=0) with a homogeneous (ndim+1)x(ndim+1) matrix M, as used by the
// reference numpy backend (tme/backends/npfftw_backend.py:357-388):
//   input coordinate  c_i = sum_j M[i][j] * o_j + M[i][3]
//   if any c_i < 0 or c_i > n_i - 1   -> output 0
//   else value = sum over the 8 corners of v * w0 * w1 * w2 with
//        w = (1 - t, t), t = c - floor(c)
// Coordinates, weights and the corner accumulation are evaluated in double
// precision with explicitly rounded multiplies/adds (no FMA contraction) in
// the same order scipy's C loop uses, then rounded once to float32, so the
// rotated template is bit-identical to scipy's in all but rare double
// rounding ties.
//
// The rotated mask's sum (the FLC n_obs) is reduced deterministically in
// FP64: one partial per workgroup, summed in a fixed order by
// k_sum_partials.
#include <math.h>

#include "tm_common.h"

#define TM_ROT_THREADS 256

struct RotBox {
  int s0, s1, s2;     // input (template) shape; also the output box shape
  int64_t ds0, ds1;   // output strides (elements) of axes 0 and 1; axis 2 stride 1
  int64_t dbatch;     // output batch stride (elements)
};

__device__ __forceinline__ double dmul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double dadd(double a, double b) { return __dadd_rn(a, b); }

// block reduction of a double (deterministic tree) -> thread 0
__device__ __forceinline__ double block_sum_d(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = sh[tid] + sh[tid + s];
    __syncthreads();
  }
  return sh[0];
}

__device__ __forceinline__ bool axis_weights(double c, int n, int& i0, int& i1, double& w0, double& w1) {
  if (!(c >= 0.0) || c > (double)(n - 1)) return false;
  const double f = floor(c);
  const double t = dadd(c, -f);
  i0 = (int)f;
  i1 = i0 + 1 < n ? i0 + 1 : i0;  // weight of the upper corner is 0 when i0 == n-1
  w0 = dadd(1.0, -t);
  w1 = t;
  return true;
}

struct RotPose {
  double center[3];       // rotation centre (voxel coordinates)
  double translation[3];  // subtracted after rotation (reference convention)
};

// Homogeneous pull matrix of the reference (npfftw_backend.py:326-355):
//   M = T(-t) . T(c) . R^{-1} . T(-c)   ->   in = R^{-1}(o - c) + c - t
// R^{-1} is formed in double from the float32 R via the adjugate.
__device__ __forceinline__ void pose_matrix(const float* __restrict__ R, const RotPose& pose, double* M) {
  const double a = R[0], b = R[1], c = R[2], d = R[3], e = R[4], f = R[5], g = R[6], h = R[7], i = R[8];
  const double A = e * i - f * h, B = f * g - d * i, C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  const double r = 1.0 / det;
  const double inv[9] = {A * r, (c * h - b * i) * r, (b * f - c * e) * r,
                         B * r, (a * i - c * g) * r, (c * d - a * f) * r,
                         C * r, (b * g - a * h) * r, (a * e - b * d) * r};
  for (int k = 0; k < 3; ++k) {
    M[4 * k + 0] = inv[3 * k + 0];
    M[4 * k + 1] = inv[3 * k + 1];
    M[4 * k + 2] = inv[3 * k + 2];
    const double rc = inv[3 * k + 0] * pose.center[0] + inv[3 * k + 1] * pose.center[1] +
                      inv[3 * k + 2] * pose.center[2];
    M[4 * k + 3] = pose.center[k] - pose.translation[k] - rc;
  }
}

// Output bricks of 8 (x) x 8 (y) x 4 (z) voxels, one per block of 256
// threads: the block stages the source region its brick samples (the
// bounding box of the brick's mapped corners, one voxel of margin, clipped
// to the template) in LDS and interpolates from there; a region above
// TM_ROT_LDS voxels, or a corner outside it, is read from global memory.
// The per-voxel arithmetic is the same either way.
#define TM_ROT_BX 8
#define TM_ROT_BY 8
#define TM_ROT_BZ 4
#define TM_ROT_LDS 4096

__device__ __forceinline__ double map_coord(const double* M, int i, double d0, double d1, double d2) {
  double t = 0.0;
  t = dadd(t, dmul(d0, M[4 * i + 0]));
  t = dadd(t, dmul(d1, M[4 * i + 1]));
  t = dadd(t, dmul(d2, M[4 * i + 2]));
  return dadd(t, M[4 * i + 3]);
}

__global__ void __launch_bounds__(TM_ROT_THREADS)
k_rotate_trilinear(const float* __restrict__ src, const float* __restrict__ srcm,
                   const float* __restrict__ rots, const double* __restrict__ mats, RotPose pose, RotBox box,
                   float* __restrict__ dst, float* __restrict__ dstm, double* __restrict__ partials) {
  __shared__ double sh[TM_ROT_THREADS];
  __shared__ double Msh[12];
  __shared__ int lo[3], ext[3];
  __shared__ float st[TM_ROT_LDS], sm[TM_ROT_LDS];
  const int b = blockIdx.y;
  const int nbx = (box.s2 + TM_ROT_BX - 1) / TM_ROT_BX, nby = (box.s1 + TM_ROT_BY - 1) / TM_ROT_BY;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
  const int x0 = bx * TM_ROT_BX, y0 = by * TM_ROT_BY, z0 = bz * TM_ROT_BZ;
  const int sz[3] = {box.s0, box.s1, box.s2};
  if (threadIdx.x == 0) {
    if (rots) {
      pose_matrix(rots + 9 * b, pose, Msh);
    } else {
      for (int k = 0; k < 12; ++k) Msh[k] = mats[12 * b + k];
    }
    const int zc[2] = {z0, min(z0 + TM_ROT_BZ, box.s0) - 1}, yc[2] = {y0, min(y0 + TM_ROT_BY, box.s1) - 1},
              xc[2] = {x0, min(x0 + TM_ROT_BX, box.s2) - 1};
    for (int i = 0; i < 3; ++i) {
      double mn = INFINITY, mx = -INFINITY;
      for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q)
          for (int r = 0; r < 2; ++r) {
            const double c = map_coord(Msh, i, zc[p], yc[q], xc[r]);
            mn = fmin(mn, c);
            mx = fmax(mx, c);
          }
      // corners used: floor(c) and floor(c) + 1 for c in [mn, mx], inside [0, n-1]
      int l = 0, h = -1;
      if (mn == mn && mx == mx && fabs(mn) < 1e9 && fabs(mx) < 1e9) {
        l = max((int)floor(mn) - 1, 0);
        h = min((int)floor(mx) + 2, sz[i] - 1);
      }
      lo[i] = l;
      ext[i] = h >= l ? h - l + 1 : 0;
    }
  }
  __syncthreads();
  const int e0 = ext[0], e1 = ext[1], e2 = ext[2];
  const int vol = e0 * e1 * e2;
  const bool staged = vol > 0 && vol <= TM_ROT_LDS;  // uniform
  if (staged) {
    for (int i = threadIdx.x; i < vol; i += TM_ROT_THREADS) {
      const int z = i / (e1 * e2), r = i - z * e1 * e2, y = r / e2, x = r - y * e2;
      const int64_t g = ((int64_t)(lo[0] + z) * box.s1 + (lo[1] + y)) * box.s2 + (lo[2] + x);
      st[i] = src[g];
      if (srcm) sm[i] = srcm[g];
    }
  }
  __syncthreads();
  const int o2 = x0 + (threadIdx.x & 7), o1 = y0 + ((threadIdx.x >> 3) & 7), o0 = z0 + (threadIdx.x >> 6);
  float val = 0.f, valm = 0.f;
  if (o0 < box.s0 && o1 < box.s1 && o2 < box.s2) {
    const double d0 = o0, d1 = o1, d2 = o2;
    double c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) c[i] = map_coord(Msh, i, d0, d1, d2);
    int a0, a1, b0, b1, f0, f1;
    double wa0, wa1, wb0, wb1, we0, we1;
    const bool in = axis_weights(c[0], box.s0, a0, a1, wa0, wa1) &&
                    axis_weights(c[1], box.s1, b0, b1, wb0, wb1) &&
                    axis_weights(c[2], box.s2, f0, f1, we0, we1);
    if (in) {
      const int ia[2] = {a0, a1}, ib[2] = {b0, b1}, ie[2] = {f0, f1};
      const double wa[2] = {wa0, wa1}, wb[2] = {wb0, wb1}, we[2] = {we0, we1};
      // the staged region holds every corner of this voxel (else: global)
      const bool lds = staged && a0 >= lo[0] && a1 < lo[0] + e0 && b0 >= lo[1] && b1 < lo[1] + e1 &&
                       f0 >= lo[2] && f1 < lo[2] + e2;
      double acc = 0.0, accm = 0.0;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int r = 0; r < 2; ++r) {
            float sv, mv = 0.f;
            if (lds) {
              const int li = ((ia[p] - lo[0]) * e1 + (ib[q] - lo[1])) * e2 + (ie[r] - lo[2]);
              sv = st[li];
              if (srcm) mv = sm[li];
            } else {
              const int64_t idx = ((int64_t)ia[p] * box.s1 + ib[q]) * box.s2 + ie[r];
              sv = src[idx];
              if (srcm) mv = srcm[idx];
            }
            double x = (double)sv;
            x = dmul(dmul(dmul(x, wa[p]), wb[q]), we[r]);
            acc = dadd(acc, x);
            if (srcm) {
              double xm = (double)mv;
              xm = dmul(dmul(dmul(xm, wa[p]), wb[q]), we[r]);
              accm = dadd(accm, xm);
            }
          }
      val = (float)acc;
      valm = (float)accm;
    }
    const int64_t o = (int64_t)b * box.dbatch + (int64_t)o0 * box.ds0 + (int64_t)o1 * box.ds1 + o2;
    dst[o] = val;
    if (dstm) dstm[o] = valm;
  }
  if (partials) {
    const double s = block_sum_d((double)valm, sh);
    if (threadIdx.x == 0) partials[(int64_t)b * gridDim.x + blockIdx.x] = s;
  }
}

// blocks (output bricks) per rotation
static int64_t rot_blocks(const int* shape) {
  return tm_ceil_div(shape[2], TM_ROT_BX) * tm_ceil_div(shape[1], TM_ROT_BY) * tm_ceil_div(shape[0], TM_ROT_BZ);
}

// out[b] = sum_{i < nparts} partials[b*nparts + i], in index order.
__global__ void k_sum_partials(const double* __restrict__ partials, int nparts, double* __restrict__ out,
                               int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  double s = 0.0;
  for (int i = 0; i < nparts; ++i) s += partials[(int64_t)b * nparts + i];
  out[b] = s;
}

// Internal launcher shared with the fused scan: poses are float32 rotations
// (9 per batch entry) about the geometric box centre.
int tm_rotate_launch(const float* src, const float* srcm, const int* shape, const float* rots, int nbatch,
                     float* dst, float* dstm, int64_t ds0, int64_t ds1, int64_t dbatch, double* n_obs,
                     double* workspace, hipStream_t s) {
  TM_REQUIRE(nbatch >= 1 && nbatch <= 65535, "rotation batch %d outside [1, 65535]", nbatch);
  RotBox box;
  box.s0 = shape[0];
  box.s1 = shape[1];
  box.s2 = shape[2];
  box.ds0 = ds0;
  box.ds1 = ds1;
  box.dbatch = dbatch;
  RotPose pose;
  for (int k = 0; k < 3; ++k) {
    pose.center[k] = (shape[k] - 1) / 2.0;
    pose.translation[k] = 0.0;
  }
  const int64_t nvox = (int64_t)box.s0 * box.s1 * box.s2;
  const int64_t nblk = rot_blocks(shape);
  TM_TIMED("rotate_trilinear", (srcm ? 8.0 : 4.0) * nvox * nbatch, s,
           hipLaunchKernelGGL(k_rotate_trilinear, dim3((unsigned)nblk, nbatch), dim3(TM_ROT_THREADS), 0, s, src,
                              srcm, rots, (const double*)nullptr, pose, box, dst, dstm, n_obs ? workspace : nullptr));
  TM_CHECK_LAUNCH();
  if (n_obs) {
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)tm_ceil_div(nbatch, 64)), dim3(64), 0, s, workspace,
                       (int)nblk, n_obs, nbatch);
    TM_CHECK_LAUNCH();
  }
  return TM_OK;
}

extern "C" {

// Number of doubles the workspace of tm_rotate_trilinear must hold when
// n_obs is requested.
int64_t tm_rotate_workspace(const int* shape, int nbatch) { return rot_blocks(shape) * (int64_t)nbatch; }

// Rotate src (and optionally the mask srcm) of shape[3] by nbatch poses and
// write the shape[0..2] box at dst + b*dst_batch_stride with row strides
// dst_strides[0..1] (the last axis is contiguous).  A pose is either a
// float32 3x3 rotation R (rots_dev, 9 floats per batch entry) combined with
// center/translation (host, 3 doubles each; translation may be null), or a
// precomputed row-major 3x4 double pull matrix (mats_dev, 12 per entry).
// When n_obs_out (device, nbatch doubles) is non-null, the FP64 sum of each
// rotated mask is written there; this needs a device workspace of
// tm_rotate_workspace() doubles.
int tm_rotate_trilinear(const float* src, const float* srcm, const int* shape, const float* rots_dev,
                        const double* mats_dev, const double* center, const double* translation, int nbatch,
                        float* dst, float* dstm, const int64_t* dst_strides, int64_t dst_batch_stride,
                        double* n_obs_out, double* workspace, void* stream) {
  TM_REQUIRE(src && dst && shape && dst_strides, "null argument to tm_rotate_trilinear");
  TM_REQUIRE((rots_dev != nullptr) != (mats_dev != nullptr), "exactly one of rots/mats must be given");
  TM_REQUIRE(!rots_dev || center, "rotation poses need a center");
  TM_REQUIRE(nbatch >= 1 && nbatch <= 65535, "rotation batch %d outside [1, 65535]", nbatch);
  TM_REQUIRE(shape[0] >= 1 && shape[1] >= 1 && shape[2] >= 1, "bad template shape");
  TM_REQUIRE(!n_obs_out || (srcm && dstm && workspace), "n_obs requires a mask, a mask output and a workspace");
  hipStream_t s = (hipStream_t)stream;
  RotBox box;
  box.s0 = shape[0];
  box.s1 = shape[1];
  box.s2 = shape[2];
  box.ds0 = dst_strides[0];
  box.ds1 = dst_strides[1];
  box.dbatch = dst_batch_stride;
  RotPose pose;
  for (int k = 0; k < 3; ++k) {
    pose.center[k] = center ? center[k] : 0.0;
    pose.translation[k] = translation ? translation[k] : 0.0;
  }
  const int64_t nblk = rot_blocks(shape);
  TM_REQUIRE(nblk <= 0x7fffffff, "template too large");
  hipLaunchKernelGGL(k_rotate_trilinear, dim3((unsigned)nblk, nbatch), dim3(TM_ROT_THREADS), 0, s, src, srcm,
                     rots_dev, mats_dev, pose, box, dst, dstm, n_obs_out ? workspace : nullptr);
  TM_CHECK_LAUNCH();
  if (n_obs_out) {
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)tm_ceil_div(nbatch, 64)), dim3(64), 0, s, workspace,
                       (int)nblk, n_obs_out, nbatch);
    TM_CHECK_LAUNCH();
  }
  return TM_OK;
}

}  // extern "C"
