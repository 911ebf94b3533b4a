// Local FLC score moments and their gradient for batches of poses (the
// refinement objective of tme/matching_optimization.py:488-548 evaluated at
// many candidate peaks at once, plus its derivative).
//
// A pose p places the standardised template g and its mask m with the
// corner at the integer voxel v_p of the target f and samples them at
//     q(o) = R_p^T (o - c) + c + s_p,   c = floor(T / 2),
// for every template voxel o (order-1 interpolation, zero outside
// [0, T-1] per axis, as scipy's map_coordinates in constant mode).  The
// forward kernel returns the six moments
//     n = sum m,  sum g m,  sum g^2 m,  sum g m f,  sum m f,  sum m f^2
// (f = 0 where the template box leaves the target) from which the host
// assembles the score; the backward kernel takes the score's derivatives
// w_k = dS/d(moment k) and returns dS/d(R^T) (3x3) and dS/ds (3),
// recomputing the interpolation weights instead of storing them.
//
// Layout: template and mask interleaved with their +x neighbour as float4
// [T0][T1][T2] (g(x), m(x), g(x1), m(x1)), x1 = min(x + 1, T2 - 1), x
// fastest: one 16-byte gather fetches both x corners of a trilinear sample,
// so a sample takes 4 gathers instead of 8 (1.8 MB for 48^3, L2-resident);
// the target is the caller's float32
// [N0][N1][N2] volume in HBM.  Grid: (blocks per pose, poses); each block
// owns a contiguous range of template voxels, every thread accumulates its
// voxels (moments in FP64), the block reduces in LDS and writes one partial
// per pose and block; a second kernel sums the partials in block order, so
// results are deterministic.
#include <math.h>
#include <stdlib.h>

#include <string.h>

#include <algorithm>

#include "tm_common.h"

#define TM_RF_THREADS 256
#define TM_RF_NJ 78  // moments (6) + Jacobian (6 x 12) per pose

struct tm_flc_local_plan {
  int T[3];
  int64_t nvox;
  int max_poses;
  int nblk;  // blocks per pose
  float rmask;  // largest distance of a nonzero mask voxel from the centre T // 2
  float4* gm;
  double* partial;  // [max_poses][nblk][12]
};

namespace {

struct Sample {
  float g, m;
  float dg[3], dm[3];  // d/dq (z, y, x); filled when GRAD
};

template <bool GRAD>
__device__ __forceinline__ Sample trilinear(const float4* __restrict__ gm, int T0, int T1, int T2, float qz, float qy,
                                            float qx) {
  Sample r;
  const bool valid = qz >= 0.f && qy >= 0.f && qx >= 0.f && qz <= (float)(T0 - 1) && qy <= (float)(T1 - 1) &&
                     qx <= (float)(T2 - 1);
  const float fz0 = floorf(qz), fy0 = floorf(qy), fx0 = floorf(qx);
  int z0 = (int)fz0, y0 = (int)fy0, x0 = (int)fx0;
  const float wz = qz - fz0, wy = qy - fy0, wx = qx - fx0;
  // out-of-range samples read voxel 0 and are zeroed afterwards
  z0 = valid ? z0 : 0;
  y0 = valid ? y0 : 0;
  x0 = valid ? x0 : 0;
  const int z1 = min(z0 + 1, T0 - 1), y1 = min(y0 + 1, T1 - 1);
  const int r00 = (z0 * T1 + y0) * T2, r01 = (z0 * T1 + y1) * T2, r10 = (z1 * T1 + y0) * T2,
            r11 = (z1 * T1 + y1) * T2;
  // each gather holds the x0 and x1 = min(x0 + 1, T2 - 1) corners
  const float4 a00 = gm[r00 + x0], a01 = gm[r01 + x0], a10 = gm[r10 + x0], a11 = gm[r11 + x0];
  const float2 v000 = make_float2(a00.x, a00.y), v001 = make_float2(a00.z, a00.w);
  const float2 v010 = make_float2(a01.x, a01.y), v011 = make_float2(a01.z, a01.w);
  const float2 v100 = make_float2(a10.x, a10.y), v101 = make_float2(a10.z, a10.w);
  const float2 v110 = make_float2(a11.x, a11.y), v111 = make_float2(a11.z, a11.w);
  auto lerp = [](float a, float b, float t) { return a + t * (b - a); };
  // along x
  const float g00 = lerp(v000.x, v001.x, wx), g01 = lerp(v010.x, v011.x, wx);
  const float g10 = lerp(v100.x, v101.x, wx), g11 = lerp(v110.x, v111.x, wx);
  const float m00 = lerp(v000.y, v001.y, wx), m01 = lerp(v010.y, v011.y, wx);
  const float m10 = lerp(v100.y, v101.y, wx), m11 = lerp(v110.y, v111.y, wx);
  const float g0 = lerp(g00, g01, wy), g1 = lerp(g10, g11, wy);
  const float m0 = lerp(m00, m01, wy), m1 = lerp(m10, m11, wy);
  r.g = valid ? lerp(g0, g1, wz) : 0.f;
  r.m = valid ? lerp(m0, m1, wz) : 0.f;
  if (GRAD) {
    const float s = valid ? 1.f : 0.f;
    r.dg[0] = s * (g1 - g0);
    r.dm[0] = s * (m1 - m0);
    r.dg[1] = s * lerp(g01 - g00, g11 - g10, wz);
    r.dm[1] = s * lerp(m01 - m00, m11 - m10, wz);
    const float gx0 = lerp(v001.x - v000.x, v011.x - v010.x, wy), gx1 = lerp(v101.x - v100.x, v111.x - v110.x, wy);
    const float mx0 = lerp(v001.y - v000.y, v011.y - v010.y, wy), mx1 = lerp(v101.y - v100.y, v111.y - v110.y, wy);
    r.dg[2] = s * lerp(gx0, gx1, wz);
    r.dm[2] = s * lerp(mx0, mx1, wz);
  }
  return r;
}

// NV = 6 (moments) or 12 (gradient): block-reduce acc[NV] in FP64 and store
template <int NV>
__device__ __forceinline__ void block_store(double* acc, double* out) {
  __shared__ double red[TM_RF_THREADS / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < TM_RF_THREADS / 64; ++w) v += red[w][threadIdx.x];
    out[threadIdx.x] = v;
  }
}

// MODE 0: the six moments; 1: dScore/d(R^T, shift) for given weights
// dScore/d(moment); 2: the moments and their Jacobian d(moment)/d(R^T,
// shift) (6 x 12) in one pass, so that the gradient of any function of the
// moments needs no second pass over the samples.
template <int MODE>
__global__ void __launch_bounds__(TM_RF_THREADS)
k_flc_local(const float* __restrict__ target, int N0, int N1, int N2, const float4* __restrict__ gm, int T0, int T1,
            int T2, const float* __restrict__ poses, const int* __restrict__ corners,
            const double* __restrict__ weights, float rskip, int nblk, double* __restrict__ partial) {
  constexpr bool BWD = MODE == 1, JAC = MODE == 2;
  constexpr int NV = MODE == 0 ? 6 : (MODE == 1 ? 12 : TM_RF_NJ);
  const int p = blockIdx.y, blk = blockIdx.x;
  const float* P = poses + 12 * p;  // R^T row-major (9), s (3)
  const float r00 = P[0], r01 = P[1], r02 = P[2], r10 = P[3], r11 = P[4], r12 = P[5], r20 = P[6], r21 = P[7],
              r22 = P[8];
  const float cz = (float)(T0 / 2), cy = (float)(T1 / 2), cx = (float)(T2 / 2);
  const float bz = cz + P[9], by = cy + P[10], bx = cx + P[11];
  const int vz = corners[3 * p], vy = corners[3 * p + 1], vx = corners[3 * p + 2];
  float w[6];
  if (BWD) {
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = (float)weights[6 * p + k];
  }
  // voxels are walked in 4x4x4 cubes, one cube per wave instruction: the
  // rotated samples of a cube fall into a compact region of the template
  // brick, so the 4 trilinear gathers of a wave touch few cache lines (a
  // row of 64 voxels would sample a line across the brick).  Every term of
  // the moments and of their gradient carries the interpolated mask, so a
  // cube whose samples are all farther than rskip = rmask + sqrt(3) from the
  // template centre adds exact zeros and is skipped: |q - c| >= |o - c| -
  // |shift| for q = R^T (o - c) + c + shift.
  const int ttx = (T2 + 3) >> 2, tty = (T1 + 3) >> 2, ttz = (T0 + 3) >> 2;
  const int ncube = ttx * tty * ttz;
  const int per = (ncube + nblk - 1) / nblk;
  const int c0 = per * blk, c1 = min(c0 + per, ncube);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  const float shift = sqrtf(P[9] * P[9] + P[10] * P[10] + P[11] * P[11]);
  const float reach = rskip + shift + 2.6f;  // + the half diagonal of a cube (1.5 sqrt 3) and rounding
  const float reach2 = reach * reach;
  // MODE 0/1: FP32 partial sums flushed into FP64 every 32 samples; MODE 2
  // (78 sums) keeps FP32 per thread (a thread sees at most 64 samples per
  // pose: nblk = ceil(T^3 / (256 * 64))) and reduces in FP64
  double acc[JAC ? 1 : NV];
  float ga[NV];
#pragma unroll
  for (int k = 0; k < (JAC ? 1 : NV); ++k) acc[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NV; ++k) ga[k] = 0.f;
  int nacc = 0;
  int cube = c0 + wv;
  int ux = cube % ttx, uy = (cube / ttx) % tty, uz = cube / (ttx * tty);
  for (; cube < c1; cube += 4) {
    if (cube != c0 + wv) {
      ux += 4;
      while (ux >= ttx) {
        ux -= ttx;
        if (++uy >= tty) {
          uy = 0;
          ++uz;
        }
      }
    }
    const float ez = 4.f * uz + 1.5f - cz, ey = 4.f * uy + 1.5f - cy, ex = 4.f * ux + 1.5f - cx;
    if (ez * ez + ey * ey + ex * ex > reach2) continue;  // uniform over the wave
    const int ix = 4 * ux + lx, iy = 4 * uy + ly, iz = 4 * uz + lz;
    if (ix >= T2 || iy >= T1 || iz >= T0) continue;
    const float dz = (float)iz - cz, dy = (float)iy - cy, dx = (float)ix - cx;
    const float qz = r00 * dz + r01 * dy + r02 * dx + bz;
    const float qy = r10 * dz + r11 * dy + r12 * dx + by;
    const float qx = r20 * dz + r21 * dy + r22 * dx + bx;
    const Sample sm = trilinear<MODE != 0>(gm, T0, T1, T2, qz, qy, qx);
    const int pz = vz + iz, py = vy + iy, px = vx + ix;
    const bool inside = pz >= 0 && py >= 0 && px >= 0 && pz < N0 && py < N1 && px < N2;
    const float f = inside ? target[((int64_t)pz * N1 + py) * N2 + px] : 0.f;
    if (JAC) {
      const float g = sm.g, m = sm.m, gmv = g * m, mf = m * f;
      ga[0] += m;
      ga[1] += gmv;
      ga[2] += gmv * g;
      ga[3] += gmv * f;
      ga[4] += mf;
      ga[5] += mf * f;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        // d(moment)/dq_a of n, sum gm, sum g^2 m, sum g m f, sum m f, sum m f^2
        const float dm = sm.dm[a], dg = sm.dg[a];
        const float G1 = g * dm + m * dg;
        const float G[6] = {dm, G1, g * (g * dm + 2.f * m * dg), f * G1, f * dm, f * f * dm};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          // q = R^T d + c + shift: d q_a / d (R^T)_ab = d_b, d q / d shift = I
          float* J = ga + 6 + 12 * k;
          J[3 * a + 0] += G[k] * dz;
          J[3 * a + 1] += G[k] * dy;
          J[3 * a + 2] += G[k] * dx;
          J[9 + a] += G[k];
        }
      }
    } else if (!BWD) {
      const float gmv = sm.g * sm.m, mf = sm.m * f;
      ga[0] += sm.m;
      ga[1] += gmv;
      ga[2] += gmv * sm.g;
      ga[3] += gmv * f;
      ga[4] += mf;
      ga[5] += mf * f;
      // flush the float partial sums into FP64 every 32 voxels
      if (++nacc == 32) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          acc[k] += (double)ga[k];
          ga[k] = 0.f;
        }
        nacc = 0;
      }
    } else {
      const float ag = w[1] * sm.m + 2.f * w[2] * sm.g * sm.m + w[3] * sm.m * f;
      const float am = w[0] + w[1] * sm.g + w[2] * sm.g * sm.g + w[3] * sm.g * f + w[4] * f + w[5] * f * f;
      const float gq0 = ag * sm.dg[0] + am * sm.dm[0];
      const float gq1 = ag * sm.dg[1] + am * sm.dm[1];
      const float gq2 = ag * sm.dg[2] + am * sm.dm[2];
      // q = R^T d + c + s: d q_a / d (R^T)_ab = d_b, d q / d s = I
      ga[0] += gq0 * dz;
      ga[1] += gq0 * dy;
      ga[2] += gq0 * dx;
      ga[3] += gq1 * dz;
      ga[4] += gq1 * dy;
      ga[5] += gq1 * dx;
      ga[6] += gq2 * dz;
      ga[7] += gq2 * dy;
      ga[8] += gq2 * dx;
      ga[9] += gq0;
      ga[10] += gq1;
      ga[11] += gq2;
      // flush the float partial sums into FP64 every 32 voxels
      if (++nacc == 32) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          acc[k] += (double)ga[k];
          ga[k] = 0.f;
        }
        nacc = 0;
      }
    }
  }
  if (JAC) {
    double out[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = (double)ga[k];
    block_store<NV>(out, partial + ((int64_t)p * nblk + blk) * TM_RF_NJ);
  } else {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += (double)ga[k];
    block_store<NV>(acc, partial + ((int64_t)p * nblk + blk) * TM_RF_NJ);
  }
}

__global__ void k_flc_local_reduce(const double* __restrict__ partial, int nposes, int nblk, int nv,
                                   double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nposes * nv) return;
  const int p = i / nv, k = i - p * nv;
  double v = 0.0;
  for (int b = 0; b < nblk; ++b) v += partial[((int64_t)p * nblk + b) * TM_RF_NJ + k];
  out[i] = v;
}

// largest distance of a nonzero mask voxel from the centre (T // 2 per axis),
// as order-preserving bits of a non-negative float
__global__ void k_mask_radius(const float* __restrict__ mask, int T0, int T1, int T2, unsigned* __restrict__ out) {
  const int64_t n = (int64_t)T0 * T1 * T2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (mask[i] == 0.f) continue;
    const int x = (int)(i % T2), y = (int)((i / T2) % T1), z = (int)(i / ((int64_t)T1 * T2));
    const float dz = (float)(z - T0 / 2), dy = (float)(y - T1 / 2), dx = (float)(x - T2 / 2);
    atomicMax(out, __float_as_uint(sqrtf(dz * dz + dy * dy + dx * dx)));
  }
}

__global__ void k_interleave(const float* __restrict__ a, const float* __restrict__ b, int64_t n, int T2,
                             float4* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t i1 = (int)(i % T2) == T2 - 1 ? i : i + 1;
  out[i] = make_float4(a[i], b[i], a[i1], b[i1]);
}

int local_launch(tm_flc_local_plan* p, int mode, const float* target, const int64_t* tshape, const float* poses,
                 const int* corners, int nposes, const double* weights, double* out, hipStream_t s) {
  const bool bwd = mode == 1;
  TM_REQUIRE(p && target && tshape && poses && corners && out, "null argument to the local FLC kernels");
  TM_REQUIRE(nposes >= 1 && nposes <= p->max_poses, "%d poses outside [1, %d]", nposes, p->max_poses);
  TM_REQUIRE(!bwd || weights, "the gradient needs the score weights");
  for (int i = 0; i < 3; ++i) TM_REQUIRE(tshape[i] >= 1 && tshape[i] < (1 << 30), "target axis %d out of range", i);
  const dim3 grid((unsigned)p->nblk, (unsigned)nposes);
  const double bytes = (double)nposes * p->nvox * 4.0;  // target window (template brick is L2-resident)
  const float rskip = p->rmask + 1.7321f;  // + sqrt(3): the trilinear cell's reach
  const int T0 = p->T[0], T1 = p->T[1], T2 = p->T[2];
  // LDS reserved but unused, to hold the kernels below their register
  // occupancy: 4 blocks per CU for the moments, 5 for the gradient (fewer
  // gather footprints sharing each CU's cache; 4096 poses x 48^3: 1.82 ->
  // 1.70 ms and 1.80 -> 1.76 ms, measured from 3 to 7 blocks)
  const size_t occ_fwd = 36 * 1024, occ_bwd = 28 * 1024;
  if (mode == 1)
    TM_TIMED("flc_local_bwd", bytes, s,
             hipLaunchKernelGGL(k_flc_local<1>, grid, dim3(TM_RF_THREADS), occ_bwd, s, target, (int)tshape[0],
                                (int)tshape[1], (int)tshape[2], p->gm, T0, T1, T2, poses, corners, weights, rskip,
                                p->nblk, p->partial));
  else if (mode == 2)
    TM_TIMED("flc_local_fwd_jac", bytes, s,
             hipLaunchKernelGGL(k_flc_local<2>, grid, dim3(TM_RF_THREADS), 0, s, target, (int)tshape[0],
                                (int)tshape[1], (int)tshape[2], p->gm, T0, T1, T2, poses, corners,
                                (const double*)nullptr, rskip, p->nblk, p->partial));
  else
    TM_TIMED("flc_local_fwd", bytes, s,
             hipLaunchKernelGGL(k_flc_local<0>, grid, dim3(TM_RF_THREADS), occ_fwd, s, target, (int)tshape[0],
                                (int)tshape[1], (int)tshape[2], p->gm, T0, T1, T2, poses, corners,
                                (const double*)nullptr, rskip, p->nblk, p->partial));
  TM_CHECK_LAUNCH();
  const int nv = mode == 1 ? 12 : (mode == 2 ? TM_RF_NJ : 6);
  hipLaunchKernelGGL(k_flc_local_reduce, dim3((unsigned)tm_ceil_div((int64_t)nposes * nv, 256)), dim3(256), 0, s,
                     p->partial, nposes, p->nblk, nv, out);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// ---------------------------------------------------------------------------
// fused refinement step (one thread per pose)

// R = Rz(c) Ry(b) Rz(a) for zyz angles (a, b, c) (scipy "zyz" extrinsic) and
// dR/da, dR/db, dR/dc
__device__ void zyz_and_derivs(double a, double b, double c, double R[9], double D[3][9]) {
  const double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cc = cos(c), sc = sin(c);
  R[0] = cc * cb * ca - sc * sa;
  R[1] = -cc * cb * sa - sc * ca;
  R[2] = cc * sb;
  R[3] = sc * cb * ca + cc * sa;
  R[4] = -sc * cb * sa + cc * ca;
  R[5] = sc * sb;
  R[6] = -sb * ca;
  R[7] = sb * sa;
  R[8] = cb;
  // d/da
  D[0][0] = -cc * cb * sa - sc * ca;
  D[0][1] = -cc * cb * ca + sc * sa;
  D[0][2] = 0.0;
  D[0][3] = -sc * cb * sa + cc * ca;
  D[0][4] = -sc * cb * ca - cc * sa;
  D[0][5] = 0.0;
  D[0][6] = sb * sa;
  D[0][7] = sb * ca;
  D[0][8] = 0.0;
  // d/db
  D[1][0] = -cc * sb * ca;
  D[1][1] = cc * sb * sa;
  D[1][2] = cc * cb;
  D[1][3] = -sc * sb * ca;
  D[1][4] = sc * sb * sa;
  D[1][5] = sc * cb;
  D[1][6] = -cb * ca;
  D[1][7] = cb * sa;
  D[1][8] = -sb;
  // d/dc
  D[2][0] = -sc * cb * ca - cc * sa;
  D[2][1] = sc * cb * sa - cc * ca;
  D[2][2] = -sc * sb;
  D[2][3] = cc * cb * ca - sc * sa;
  D[2][4] = -cc * cb * sa - sc * ca;
  D[2][5] = cc * sb;
  D[2][6] = 0.0;
  D[2][7] = 0.0;
  D[2][8] = 0.0;
}

// poses float32 [n][12] (R^T row-major, shift) and corners int32 [n][3] from
// translations t and zyz angles a (float64 [n][3], radians): v = trunc(t),
// s = t - v; the shift is -R^T s (rigid) or s (the reference's sampling)
__global__ void k_refine_prepare(const double* __restrict__ t, const double* __restrict__ ang, int n, int rigid,
                                 float* __restrict__ poses, int* __restrict__ corners) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double R[9], D[3][9];
  zyz_and_derivs(ang[3 * i], ang[3 * i + 1], ang[3 * i + 2], R, D);
  double sv[3];
  for (int k = 0; k < 3; ++k) {
    const double v = trunc(t[3 * i + k]);
    corners[3 * i + k] = (int)v;
    sv[k] = t[3 * i + k] - v;
  }
  float* P = poses + 12 * i;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) P[3 * r + c] = (float)R[3 * c + r];  // R^T
  for (int r = 0; r < 3; ++r) {
    double sh = 0.0;  // (R^T s)_r in float64, then rounded like the rest of the pose
    for (int c = 0; c < 3; ++c) sh += R[3 * c + r] * sv[c];
    P[9 + r] = rigid ? (float)(-sh) : (float)sv[r];
  }
}

struct RefineAdam {
  double lr_t, lr_a, beta1, beta2, eps, bc1, bc2sqrt, bound, eps32;
  int rigid, update;
};

// score of pose i from its six moments M; with want_w also the weights
// w = dScore / d(n, A, B, C, D, E)
__device__ double score_and_weights(const double* M, double eps32, bool want_w, double* w) {
  const double nn = M[0], A = M[1], Bm = M[2], C = M[3], Dd = M[4], E = M[5];
  const double mu = A / nn;
  const double var_t = fmax(Bm / nn - mu * mu, 0.0), var_f = fmax(E / nn - (Dd / nn) * (Dd / nn), 0.0);
  const double sd_t = sqrt(var_t), sd_f = sqrt(var_f);
  const double P = C - mu * Dd, Q = sd_t * sd_f * nn;
  const bool flat = sd_f < eps32;
  const double score = flat ? 0.0 : P / Q;
  if (!want_w) return score;
  for (int k = 0; k < 6; ++k) w[k] = 0.0;
  if (!flat) {
    const double dPdn = Dd * A / (nn * nn), dPdA = -Dd / nn, dPdC = 1.0, dPdD = -mu;
    const double dsdt_dB = 1.0 / (2.0 * sd_t * nn), dsdt_dA = -(A / (nn * nn)) / sd_t,
                 dsdt_dn = (-Bm / (nn * nn) + 2.0 * A * A / (nn * nn * nn)) / (2.0 * sd_t);
    const double dsdf_dE = 1.0 / (2.0 * sd_f * nn), dsdf_dD = -(Dd / (nn * nn)) / sd_f,
                 dsdf_dn = (-E / (nn * nn) + 2.0 * Dd * Dd / (nn * nn * nn)) / (2.0 * sd_f);
    const double dQdn = nn * (sd_f * dsdt_dn + sd_t * dsdf_dn) + sd_t * sd_f;
    const double dQdA = nn * sd_f * dsdt_dA, dQdB = nn * sd_f * dsdt_dB;
    const double dQdD = nn * sd_t * dsdf_dD, dQdE = nn * sd_t * dsdf_dE;
    const double S = score;
    w[0] = (dPdn - S * dQdn) / Q;
    w[1] = (dPdA - S * dQdA) / Q;
    w[2] = (-S * dQdB) / Q;
    w[3] = dPdC / Q;
    w[4] = (dPdD - S * dQdD) / Q;
    w[5] = (-S * dQdE) / Q;
  }
  return score;
}

__device__ void track_best(int i, double score, const double* t, const double* ang, double* best_s, double* best_t,
                           double* best_a) {
  if (score > best_s[i]) {
    best_s[i] = score;
    for (int k = 0; k < 3; ++k) {
      best_t[3 * i + k] = t[3 * i + k];
      best_a[3 * i + k] = ang[3 * i + k];
    }
  }
}

// g = dScore/d(R^T) row-major (9), dScore/d(shift) (3) of pose i, back to
// (t, a) (shift = -R^T s (rigid) or s; s = t - trunc(t); R^T from a), then
// one Adam step on the loss -score (torch's update): m <- lerp(m, grad,
// 1 - beta1), v <- beta2 v + (1 - beta2) grad^2, p <- p - lr / bc1 * m /
// (sqrt(v) / sqrt(bc2) + eps); t stays within bound of t0 when bound >= 0
__device__ void chain_and_adam(int i, const double* g, double* t, double* ang, const double* t0, double* mom1,
                               double* mom2, const RefineAdam& h) {
  double R[9], Dr[3][9];
  zyz_and_derivs(ang[3 * i], ang[3 * i + 1], ang[3 * i + 2], R, Dr);
  double Rt[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Rt[3 * r + c] = R[3 * c + r];
  double sv[3];
  for (int k = 0; k < 3; ++k) sv[k] = t[3 * i + k] - trunc(t[3 * i + k]);
  double gRt[9], gs[3];
  for (int k = 0; k < 9; ++k) gRt[k] = g[k];
  if (h.rigid) {
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int r = 0; r < 3; ++r) v -= g[9 + r] * Rt[3 * r + c];
      gs[c] = v;
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) gRt[3 * r + c] -= g[9 + r] * sv[c];
  } else {
    for (int c = 0; c < 3; ++c) gs[c] = g[9 + c];
  }
  double ga[3];
  for (int k = 0; k < 3; ++k) {
    double v = 0.0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) v += gRt[3 * r + c] * Dr[k][3 * c + r];  // dS/dR[c][r] = dS/dRt[r][c]
    ga[k] = v;
  }
  for (int k = 0; k < 6; ++k) {
    const bool is_t = k < 3;
    const double grad = -(is_t ? gs[k] : ga[k - 3]);
    double* m1 = mom1 + 6 * i + k;
    double* m2 = mom2 + 6 * i + k;
    *m1 = *m1 + (1.0 - h.beta1) * (grad - *m1);
    *m2 = h.beta2 * *m2 + (1.0 - h.beta2) * grad * grad;
    const double denom = sqrt(*m2) / h.bc2sqrt + h.eps;
    const double step = (is_t ? h.lr_t : h.lr_a) / h.bc1;
    double* p = is_t ? t + 3 * i + k : ang + 3 * i + (k - 3);
    *p -= step * *m1 / denom;
    if (is_t && h.bound >= 0.0) *p = fmin(fmax(*p, t0[3 * i + k] - h.bound), t0[3 * i + k] + h.bound);
  }
}

// one-pass form: score, its gradient from the moments and their Jacobian
// (mj [n][78]), best-pose tracking and the Adam step
__global__ void k_refine_adam(const double* __restrict__ mj, double* __restrict__ t, double* __restrict__ ang,
                              const double* __restrict__ t0, double* __restrict__ mom1, double* __restrict__ mom2,
                              double* __restrict__ best_s, double* __restrict__ best_t, double* __restrict__ best_a,
                              int n, RefineAdam h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* M = mj + 78 * i;
  double w[6];
  const double score = score_and_weights(M, h.eps32, h.update, w);
  track_best(i, score, t, ang, best_s, best_t, best_a);
  if (!h.update) return;
  double g[12];
  for (int j = 0; j < 12; ++j) {
    double v = 0.0;
    for (int k = 0; k < 6; ++k) v += w[k] * M[6 + 12 * k + j];
    g[j] = v;
  }
  chain_and_adam(i, g, t, ang, t0, mom1, mom2, h);
}

// two-pass form, first half: score from the moments (m [n][6]), best-pose
// tracking and, when update, the weights dScore/d(moment) for the gradient
// kernel (w [n][6])
__global__ void k_refine_weights(const double* __restrict__ m, const double* __restrict__ t,
                                 const double* __restrict__ ang, double* __restrict__ best_s,
                                 double* __restrict__ best_t, double* __restrict__ best_a, double* __restrict__ w,
                                 int n, int update, double eps32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double wi[6];
  const double score = score_and_weights(m + 6 * i, eps32, update != 0, wi);
  track_best(i, score, t, ang, best_s, best_t, best_a);
  if (update)
    for (int k = 0; k < 6; ++k) w[6 * i + k] = wi[k];
}

// two-pass form, second half: the Adam step from the gradient kernel's
// dScore/d(R^T, shift) (g [n][12])
__global__ void k_refine_adam_grad(const double* __restrict__ g, double* __restrict__ t, double* __restrict__ ang,
                                   const double* __restrict__ t0, double* __restrict__ mom1,
                                   double* __restrict__ mom2, int n, RefineAdam h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  chain_and_adam(i, g + 12 * i, t, ang, t0, mom1, mom2, h);
}

}  // namespace

extern "C" {

// Plan for poses of one standardised template and its mask (T^3 float32,
// device).  The template and mask are copied (interleaved) into the plan.
int tm_flc_local_plan_create(const float* template_std, const float* mask, const int* tshape, int max_poses,
                             tm_flc_local_plan** out) {
  TM_REQUIRE(template_std && mask && tshape && out, "null argument to tm_flc_local_plan_create");
  TM_REQUIRE(max_poses >= 1 && max_poses <= (1 << 20), "max_poses %d out of range", max_poses);
  for (int i = 0; i < 3; ++i) TM_REQUIRE(tshape[i] >= 1 && tshape[i] <= 1024, "template axis %d out of range", i);
  tm_flc_local_plan* p = (tm_flc_local_plan*)calloc(1, sizeof(tm_flc_local_plan));
  TM_REQUIRE(p, "host allocation failed");
  for (int i = 0; i < 3; ++i) p->T[i] = tshape[i];
  p->nvox = (int64_t)tshape[0] * tshape[1] * tshape[2];
  p->max_poses = max_poses;
  // at most 64 voxels per thread (the FP32 per-thread sums of the Jacobian
  // pass rely on it), at least one block per pose
  int64_t nb = tm_ceil_div(p->nvox, (int64_t)TM_RF_THREADS * 64);
  p->nblk = (int)(nb < 1 ? 1 : nb);
  hipError_t e = hipMalloc((void**)&p->gm, sizeof(float4) * p->nvox);
  if (e == hipSuccess)
    e = hipMalloc((void**)&p->partial, sizeof(double) * TM_RF_NJ * (size_t)max_poses * p->nblk);
  if (e != hipSuccess) {
    tm_set_error("hipMalloc failed: %s", hipGetErrorString(e));
    if (p->gm) hipFree(p->gm);
    if (p->partial) hipFree(p->partial);
    free(p);
    return e == hipErrorOutOfMemory ? TM_ERR_OOM : TM_ERR_HIP;
  }
  hipLaunchKernelGGL(k_interleave, dim3((unsigned)tm_ceil_div(p->nvox, 256)), dim3(256), 0, 0, template_std, mask,
                     p->nvox, p->T[2], p->gm);
  e = hipGetLastError();
  // the mask radius (the partial buffer's first word serves as the counter)
  unsigned rbits = 0u;
  if (e == hipSuccess) e = hipMemcpy(p->partial, &rbits, sizeof(unsigned), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mask_radius, dim3((unsigned)std::min<int64_t>(tm_ceil_div(p->nvox, 256), 1024)), dim3(256), 0,
                       0, mask, p->T[0], p->T[1], p->T[2], (unsigned*)p->partial);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&rbits, p->partial, sizeof(unsigned), hipMemcpyDeviceToHost);
  memcpy(&p->rmask, &rbits, sizeof(float));
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    tm_set_error("preparing the template brick: %s", hipGetErrorString(e));
    hipFree(p->gm);
    hipFree(p->partial);
    free(p);
    return TM_ERR_HIP;
  }
  *out = p;
  return TM_OK;
}

int tm_flc_local_plan_destroy(tm_flc_local_plan* p) {
  if (p) {
    hipFree(p->gm);
    hipFree(p->partial);
    free(p);
  }
  return TM_OK;
}

// Moments of n_poses poses: poses float32 [n][12] = R^T row-major, s;
// corners int32 [n][3] = v; out float64 [n][6].
int tm_flc_local_fwd(tm_flc_local_plan* p, const float* target, const int64_t* target_shape, const float* poses,
                     const int* corners, int n_poses, double* moments, void* stream) {
  return local_launch(p, 0, target, target_shape, poses, corners, n_poses, nullptr, moments, (hipStream_t)stream);
}

// Moments and their Jacobian in one pass: out float64 [n][78] = the six
// moments, then for each moment k the 12 derivatives d(moment_k)/d(R^T)
// row-major (9) and d(moment_k)/ds (3).
int tm_flc_local_fwd_jac(tm_flc_local_plan* p, const float* target, const int64_t* target_shape, const float* poses,
                         const int* corners, int n_poses, double* out, void* stream) {
  return local_launch(p, 2, target, target_shape, poses, corners, n_poses, nullptr, out, (hipStream_t)stream);
}

// Gradient: weights float64 [n][6] = dS/d(moment); out float64 [n][12] =
// dS/d(R^T) row-major (9), dS/ds (3).
int tm_flc_local_bwd(tm_flc_local_plan* p, const float* target, const int64_t* target_shape, const float* poses,
                     const int* corners, int n_poses, const double* weights, double* grad, void* stream) {
  return local_launch(p, 1, target, target_shape, poses, corners, n_poses, weights, grad, (hipStream_t)stream);
}

// Fused refinement step helpers (no counterpart in the reference): pose
// preparation and the score / gradient / Adam / best-pose update, in the
// one-pass form (tm_refine_adam after tm_flc_local_fwd_jac) and the
// two-pass form (tm_refine_weights after tm_flc_local_fwd, then
// tm_flc_local_bwd with those weights and tm_refine_adam_grad).
int tm_refine_prepare(const double* t, const double* angles, int n, int rigid, float* poses, int* corners,
                      void* stream) {
  TM_REQUIRE(t && angles && poses && corners && n >= 1, "bad arguments to tm_refine_prepare");
  hipLaunchKernelGGL(k_refine_prepare, dim3((unsigned)tm_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, t,
                     angles, n, rigid, poses, corners);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

static RefineAdam adam_params(int rigid, int update, double lr_t, double lr_a, double beta1, double beta2,
                              double eps, int adam_step, double bound, double eps32) {
  RefineAdam h;
  h.lr_t = lr_t;
  h.lr_a = lr_a;
  h.beta1 = beta1;
  h.beta2 = beta2;
  h.eps = eps;
  h.bc1 = 1.0 - pow(beta1, (double)adam_step);
  h.bc2sqrt = sqrt(1.0 - pow(beta2, (double)adam_step));
  h.bound = bound;
  h.eps32 = eps32;
  h.rigid = rigid;
  h.update = update;
  return h;
}

int tm_refine_adam(const double* moments_jac, double* t, double* angles, const double* t0, double* adam_m,
                   double* adam_v, double* best_score, double* best_t, double* best_angles, int n, int rigid,
                   int update, double lr_t, double lr_a, double beta1, double beta2, double eps, int adam_step,
                   double bound, double eps32, void* stream) {
  TM_REQUIRE(moments_jac && t && angles && t0 && adam_m && adam_v && best_score && best_t && best_angles && n >= 1,
             "bad arguments to tm_refine_adam");
  TM_REQUIRE(adam_step >= 1, "Adam steps count from 1");
  const RefineAdam h = adam_params(rigid, update, lr_t, lr_a, beta1, beta2, eps, adam_step, bound, eps32);
  hipLaunchKernelGGL(k_refine_adam, dim3((unsigned)tm_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     moments_jac, t, angles, t0, adam_m, adam_v, best_score, best_t, best_angles, n, h);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_refine_weights(const double* moments, const double* t, const double* angles, double* best_score,
                      double* best_t, double* best_angles, double* weights, int n, int update, double eps32,
                      void* stream) {
  TM_REQUIRE(moments && t && angles && best_score && best_t && best_angles && n >= 1 && (weights || !update),
             "bad arguments to tm_refine_weights");
  hipLaunchKernelGGL(k_refine_weights, dim3((unsigned)tm_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     moments, t, angles, best_score, best_t, best_angles, weights, n, update, eps32);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_refine_adam_grad(const double* grad, double* t, double* angles, const double* t0, double* adam_m,
                        double* adam_v, int n, int rigid, double lr_t, double lr_a, double beta1, double beta2,
                        double eps, int adam_step, double bound, void* stream) {
  TM_REQUIRE(grad && t && angles && t0 && adam_m && adam_v && n >= 1, "bad arguments to tm_refine_adam_grad");
  TM_REQUIRE(adam_step >= 1, "Adam steps count from 1");
  const RefineAdam h = adam_params(rigid, 1, lr_t, lr_a, beta1, beta2, eps, adam_step, bound, 0.0);
  hipLaunchKernelGGL(k_refine_adam_grad, dim3((unsigned)tm_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     grad, t, angles, t0, adam_m, adam_v, n, h);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

}  // extern "C"
