// Batched fused exhaustive scan: one call scores a batch of B rotations over
// the whole padded tomogram and folds them into the max/argmax state.
//
// Data layout (C order, complex64 interleaved), fast shape (N0, N1, N2),
// N2c = N2/2 + 1, template box (T0, T1, T2):
//   TR[b][T0][T1][T2], MR[b][...]  rotated template / mask (float32)
//   Q1[q][b][T0][T1][N2c]          x-transforms of the template boxes
//   P [q][b][T0][N1][N2c]          x,y-transforms (only T0 z-planes exist)
//   Q [o][b][N0][N1][N2c]          correlation spectra after the z pass
// with q in {mask, template} (FLC) or {template} (CORR) and o in
// {S1, S2, X} (FLC) or {X} (CORR).
//
// Per batch:
//   rotate      TR, MR (+ FP64 sum of each rotated mask)       T^3 work
//   filter      batched T^3 rfft * filter * irfft (optional)   T^3 work
//   normalize   FP64 masked moments, t = (t - mean)/std * m    T^3 work
//   x pass      R2C of the T0*T1 template rows, zero-padded    small
//   y pass      C2C of T1 -> N1 zero-padded lines              ~T0/N0 of a pass
//   pass A      per (ky, kx) column: forward z-FFT of the pruned template
//               columns, multiply with the resident target spectra,
//               inverse z-FFT of every product -> Q           reads spectra once per batch
//   pass B      inverse y-FFT of Q in place                    one HBM pass
//   pass C      per (z, y) row: C2R of the NQ rows, FLC normalisation (or
//               the FLCSphericalMask denominator), max/argmax over the
//               batch in registers, one read-modify-write of the state
// The target spectra are read once per batch instead of once per rotation
// and the score volume never exists in HBM.
#include <math.h>
#include <algorithm>
#include <stdlib.h>
#include <string.h>

#include "fft_core.h"
#include "fft_ct.h"
#include "fft_plan.h"
#include <type_traits>

#include "tm_common.h"

#define TM_FU_THREADS 256
#define TM_FU_MAXP 16  // positions per thread: tile elements <= 4096 = 16 * 256
#define TM_MOM_BLOCKS 32
// pass C on radix-4 units keeps the twiddle powers of its last stage (1)
// or of both its stages (2) in registers for all of a block's transforms
#ifndef TM_PC_HOLD_TWIDDLES
#define TM_PC_HOLD_TWIDDLES 2
#endif
// 1: pass A's pruned forward transforms skip their broadcast first stage
#ifndef TM_PA_BCAST
#define TM_PA_BCAST 1
#endif

int tm_rotate_launch(const float* src, const float* srcm, const int* shape, const float* rots, int nbatch,
                     float* dst, float* dstm, int64_t ds0, int64_t ds1, int64_t dbatch, double* n_obs,
                     double* workspace, hipStream_t s);
extern "C" int tm_cmul_cr(const void* a, const float* f, void* out, int64_t n, int64_t nf, void* stream);
extern "C" int64_t tm_rotate_workspace(const int* shape, int nbatch);

// --------------------------------------------------------------------------
// batched normalize_template (tme/matching_utils.py:61-99)

// partials[b][blk][2] = (sum t*m, sum t*t*m) over a fixed split of the box
__global__ void __launch_bounds__(TM_FU_THREADS)
k_moments_batched(const float* __restrict__ t, const float* __restrict__ m, int64_t n, int64_t m_bstride,
                  double* __restrict__ partials) {
  __shared__ double sh[2][TM_FU_THREADS];
  const int b = blockIdx.y;
  const float* tb = t + (int64_t)b * n;
  const float* mb = m + (int64_t)b * m_bstride;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TM_FU_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * TM_FU_THREADS) {
    const double a = tb[i], w = mb[i];
    s0 += a * w;
    s1 += a * a * w;
  }
  const int tid = threadIdx.x;
  sh[0][tid] = s0;
  sh[1][tid] = s1;
  __syncthreads();
  for (int s = TM_FU_THREADS / 2; s > 0; s >>= 1) {
    if (tid < s) {
      sh[0][tid] += sh[0][tid + s];
      sh[1][tid] += sh[1][tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double* p = partials + 2 * ((int64_t)b * gridDim.x + blockIdx.x);
    p[0] = sh[0][0];
    p[1] = sh[1][0];
  }
}

// ms[b] = (mean, std) with n = n_dev[b] (per rotation) or n_host
__global__ void k_moments_finalize(const double* __restrict__ partials, int nparts, const double* __restrict__ n_dev,
                                   double n_host, int nbatch, double* __restrict__ ms) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  double s0 = 0.0, s1 = 0.0;
  for (int i = 0; i < nparts; ++i) {
    s0 += partials[2 * ((int64_t)b * nparts + i)];
    s1 += partials[2 * ((int64_t)b * nparts + i) + 1];
  }
  const double nn = n_dev ? n_dev[b] : n_host;
  const double mean = s0 / nn;
  double var = s1 / nn - mean * mean;
  var = var > 0.0 ? var : 0.0;
  ms[2 * b] = mean;
  ms[2 * b + 1] = sqrt(var);
}

__global__ void k_normalize_batched(float* __restrict__ t, const float* __restrict__ m, int64_t n, int64_t m_bstride,
                                    const double* __restrict__ ms) {
  const int b = blockIdx.y;
  const float mean = (float)ms[2 * b];
  const float sd = (float)ms[2 * b + 1];
  float* tb = t + (int64_t)b * n;
  const float* mb = m + (int64_t)b * m_bstride;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    tb[i] = (tb[i] - mean) / sd * mb[i];
}

// --------------------------------------------------------------------------
// pass A: z-axis forward (pruned) x target spectra x z-axis inverse

template <int NQ>
__global__ void __launch_bounds__(TM_FU_THREADS)
k_pass_a(const float2* __restrict__ P, const float2* __restrict__ Ft, const float2* __restrict__ Ft2,
         float2* __restrict__ Q, LinePlan lp, int T0, int N1, int N2c, int N2s, int B, int TI, int tiles_per_row) {
  extern __shared__ float2 lds[];
  const int N0 = lp.n;
  const int tid = threadIdx.x;
  const int ky = blockIdx.x / tiles_per_row;
  const int kx0 = (blockIdx.x - ky * tiles_per_row) * TI;
  const int ti = (N2c - kx0) < TI ? (N2c - kx0) : TI;
  const int npos = N0 * ti;
  const int64_t plane = (int64_t)N1 * N2s;
  const int64_t col = (int64_t)ky * N2s + kx0;
  float2* A = lds;
  float2* Bf = lds + npos;

  float2 F[TM_FU_MAXP], F2[TM_FU_MAXP], M[TM_FU_MAXP];
#pragma unroll
  for (int i = 0; i < TM_FU_MAXP; ++i) {
    const int p = tid + i * TM_FU_THREADS;
    F[i] = F2[i] = M[i] = make_float2(0.f, 0.f);
    if (p < npos) {
      const int z = p / ti, l = p - z * ti;
      F[i] = Ft[z * plane + col + l];
      if (NQ == 3) F2[i] = Ft2[z * plane + col + l];
    }
  }

  for (int b = 0; b < B; ++b) {
    if (NQ == 3) {
      // mask spectrum along z (inputs beyond T0 are zero)
      const float2* src = P + ((int64_t)(0 * B + b) * T0) * plane + col;
      for (int p = tid; p < npos; p += TM_FU_THREADS) {
        const int z = p / ti, l = p - z * ti;
        A[p] = z < T0 ? src[z * plane + l] : make_float2(0.f, 0.f);
      }
      __syncthreads();
      float2* R = lds_fft_lines<-1, true>(A, Bf, ti, 1, ti, lp, tid, TM_FU_THREADS);
#pragma unroll
      for (int i = 0; i < TM_FU_MAXP; ++i) {
        const int p = tid + i * TM_FU_THREADS;
        if (p < npos) M[i] = R[p];
      }
      __syncthreads();
    }
    {
      // template spectrum, X = F * t
      const float2* src = P + ((int64_t)((NQ == 3 ? 1 : 0) * B + b) * T0) * plane + col;
      for (int p = tid; p < npos; p += TM_FU_THREADS) {
        const int z = p / ti, l = p - z * ti;
        A[p] = z < T0 ? src[z * plane + l] : make_float2(0.f, 0.f);
      }
      __syncthreads();
      float2* R = lds_fft_lines<-1, true>(A, Bf, ti, 1, ti, lp, tid, TM_FU_THREADS);
#pragma unroll
      for (int i = 0; i < TM_FU_MAXP; ++i) {
        const int p = tid + i * TM_FU_THREADS;
        if (p < npos) R[p] = c_mul(F[i], R[p]);
      }
      __syncthreads();
      float2* O = lds_fft_lines<1, true>(R, R == A ? Bf : A, ti, 1, ti, lp, tid, TM_FU_THREADS);
      float2* dst = Q + ((int64_t)((NQ - 1) * B + b) * N0) * plane + col;
      for (int p = tid; p < npos; p += TM_FU_THREADS) {
        const int z = p / ti, l = p - z * ti;
        dst[z * plane + l] = O[p];
      }
      __syncthreads();
    }
    if (NQ == 3) {
#pragma unroll 1
      for (int o = 0; o < 2; ++o) {
#pragma unroll
        for (int i = 0; i < TM_FU_MAXP; ++i) {
          const int p = tid + i * TM_FU_THREADS;
          if (p < npos) A[p] = c_mul(o == 0 ? F[i] : F2[i], M[i]);
        }
        __syncthreads();
        float2* O = lds_fft_lines<1, true>(A, Bf, ti, 1, ti, lp, tid, TM_FU_THREADS);
        float2* dst = Q + ((int64_t)(o * B + b) * N0) * plane + col;
        for (int p = tid; p < npos; p += TM_FU_THREADS) {
          const int z = p / ti, l = p - z * ti;
          dst[z * plane + l] = O[p];
        }
        __syncthreads();
      }
    }
  }
}

// --------------------------------------------------------------------------
// pass C: C2R along x, score, max/argmax over the batch, state update

__device__ __forceinline__ float fu_flc_norm(float x, float ex2, float ex, float n, float eps) {
#pragma clang fp contract(off)
  float m = ex / n;
  m = m * m;
  float v = ex2 / n - m;
  v = fmaxf(v, 0.f);
  float sd = sqrtf(v);
  sd = sd < eps ? 1.f : sd;
  return x / (sd * n);
}

// The factor r with fu_flc_norm(x * scale, ex2 * scale, ex * scale, n, eps)
// ~= x * r, from the unscaled inverse-transform outputs and sn = scale / n:
// r = sn / sd, or sn where sd = sqrt(v) < eps (v < eps^2).  The hardware
// reciprocal square root (one quarter-rate instruction instead of a square
// root and a reciprocal) and the folded scale replace the correctly rounded
// divisions: a few ulp in the score.
__device__ __forceinline__ float fu_flc_factor(float ex2, float ex, float sn, float eps2) {
  const float m = ex * sn;
  const float v = fmaxf(ex2 * sn - m * m, 0.f);  // fmaxf also maps NaN to 0
  return v < eps2 ? sn : sn * __builtin_amdgcn_rsqf(v);
}

template <int NQ>
__global__ void __launch_bounds__(TM_FU_THREADS)
k_pass_c(const float2* __restrict__ Q, LinePlan lp, int64_t nrows, int64_t ldq, int RL, int B,
         const double* __restrict__ n_obs,
         const float* __restrict__ inv_den, const int* __restrict__ rot_idx, float scale, float eps,
         float* __restrict__ max_scores, int* __restrict__ max_rot) {
  extern __shared__ float2 lds[];
  const int N2 = lp.n, N2c = N2 / 2 + 1;
  const int ld = N2 + 1;
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * RL;
  const int rl = (int)((nrows - r0) < RL ? (nrows - r0) : RL);
  const int npos = rl * N2;
  float2* A = lds;
  float2* Bf = lds + RL * ld;

  float best[TM_FU_MAXP], v1[TM_FU_MAXP], v2[TM_FU_MAXP];
  int arg[TM_FU_MAXP];
#pragma unroll
  for (int i = 0; i < TM_FU_MAXP; ++i) {
    best[i] = -INFINITY;
    arg[i] = -1;
    v1[i] = v2[i] = 0.f;
  }
  for (int b = 0; b < B; ++b) {
#pragma unroll 1
    for (int o = 0; o < NQ; ++o) {
      const float2* src = Q + ((int64_t)(o * B + b) * nrows + r0) * ldq;
      for (int f = tid; f < N2c * rl; f += TM_FU_THREADS) {
        const int l = f / N2c, e = f - l * N2c;
        const float2 v = src[(int64_t)l * ldq + e];
        A[l * ld + e] = v;
        if (e > 0 && N2 - e >= N2c) A[l * ld + (N2 - e)] = c_conj(v);
      }
      __syncthreads();
      float2* R = lds_fft_lines<1, false>(A, Bf, rl, ld, 1, lp, tid, TM_FU_THREADS);
#pragma unroll
      for (int i = 0; i < TM_FU_MAXP; ++i) {
        const int p = tid + i * TM_FU_THREADS;
        if (p < npos) {
          const int l = p / N2, e = p - l * N2;
          const float x = R[l * ld + e].x * scale;
          if (o == NQ - 1) {
            float s;
            if (NQ == 3) {
              s = fu_flc_norm(x, v2[i], v1[i], (float)n_obs[b], eps);
            } else {
              s = inv_den ? x * inv_den[(r0 + l) * N2 + e] : x;
            }
            if (s > best[i]) {
              best[i] = s;
              arg[i] = rot_idx[b];
            }
          } else if (o == 0) {
            v1[i] = x;
          } else {
            v2[i] = x;
          }
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < TM_FU_MAXP; ++i) {
    const int p = tid + i * TM_FU_THREADS;
    if (p < npos) {
      const int l = p / N2, e = p - l * N2;
      const int64_t g = (r0 + l) * N2 + e;
      if (best[i] > max_scores[g]) {
        max_scores[g] = best[i];
        max_rot[g] = arg[i];
      }
    }
  }
}


// --------------------------------------------------------------------------
// compile-time specialised pass A / pass C

// Pass A with the z transforms in registers: a block owns L (128 B) kx
// columns of one ky; spectra are multiplied in registers between the forward
// (pruned) and inverse transforms.  KIND 1: X = F t (template slot q_in,
// output slot o_x); KIND 0 (FLC): S1 = F m, S2 = F2 m (mask slot q_in, output
// slots 0 and 1); KIND 2 (FLC): all three per rotation (mask slot 0,
// template slot 1, outputs 0, 1, 2) with the two target tiles read once.
// The main launches split the FLC work into KIND 0 + KIND 1, which keeps
// the resident spectra and transform buffers within the register file (a
// KIND 2 main kernel with deferred stores spills at 128 VGPRs: 11.6 ms
// against 6.6 + 4.1 ms on 512x512x200 at B = 64); the short tail launches
// take KIND 2 (1.04 against 0.73 + 0.41 ms).
//
// The spectra rows are padded to a multiple of L bins, so every load and
// store is unpredicated (padding columns are transformed too and never
// read).  PR: T0 <= N0 / R_IN, so only the first element of each first-stage
// butterfly can be non-zero and the others are compile-time zeros.  The next
// rotation's input columns are loaded before this rotation's transforms.
// The X kernel's 512-thread blocks (KIND 1) are held to 128 VGPRs for
// lengths up to 512 so two blocks share a CU (a couple of spilled registers
// are cheaper than half the occupancy).
//
// Split rows (TAIL): the output rows hold ldo bins (a multiple of L) and the
// last ntl bins of every row go to the dense tail array Qt[o][b][ky][z][ntl];
// the tail tiles are their own launch (kx_base = ldo) whose stores are
// predicated to the valid columns, so the main launch stays unpredicated; a
// tail tile packs L / ntl rows (ky) of ntl columns each.
template <int N0, int L, int NT, int KIND, bool PR, bool TAIL = false, int V = 0, bool DEFER = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu((KIND == 1 && PR && N0 <= 512) ? 4 : 1, 8)))
k_pass_a_ct(const float2* __restrict__ P, const float2* __restrict__ Ft, const float2* __restrict__ Ft2,
            float2* __restrict__ Q, float2* __restrict__ Qt, const float2* __restrict__ tw, int T0, int N1, int N2s,
            int ldo, int B, int tiles_per_row, int kx_base, int ntl, int q_in, int o_x) {
  using PF = CtPlan<N0, L, NT, true, false, V>;
  using PI = CtPlan<N0, L, NT, true, true, V>;
  static_assert(PF::E_OUT == PI::E_IN && PF::R_OUT == PI::R_IN && PF::NST == PI::NST, "ownership mismatch");
  static_assert(!DEFER || (!TAIL && KIND != 2 && PF::NST >= 2 && PI::NST == PF::NST),
                "deferred stores need staged transforms (and one or two outputs per rotation)");
  static_assert((N0 / PF::R_IN * L) % NT == 0 && (N0 / PF::R_OUT * L) % NT == 0 &&
                    (N0 / PI::R_OUT * L) % NT == 0,
                "every butterfly slot must be valid");
  constexpr int EI = PF::E_IN, EO = PF::E_OUT, EW = EI > EO ? EI : EO;
  constexpr int BI = EI / PF::R_IN;      // first-stage butterflies per thread
  constexpr int NX = PR ? BI : EI;       // input loads per thread
  extern __shared__ float2 lds[];
  __shared__ float2 tws[N0];  // twiddles in LDS: no vector-memory waits inside the transforms
  const int tid = threadIdx.x;
  for (int i = tid; i < N0; i += NT) tws[i] = tw[i];
  // the transforms run over two exchange buffers (ct_fft2: no barrier
  // inside a stage nor before a transform); with an even stage count the
  // last stage reads bx, so the next transform starts in the other buffer
  float2 *bx = lds, *by = lds + N0 * L;
  auto after_transform = [&]() {
    if constexpr (ct2_last_in_x<PF>()) {
      float2* t = bx;
      bx = by;
      by = t;
    }
  };
  const int bid = tm_xcd_block();
  const int ky = TAIL ? bid * (L / ntl) : bid / tiles_per_row;
  const int kx0 = kx_base + (TAIL ? 0 : (bid - ky * tiles_per_row) * L);
  // TAIL: lane column l holds row ky + l / ntl, column kx0 + l % ntl; its
  // input offset from col is tl; lanes past the packed rows read a valid
  // column and store nothing
  int tl = 0, tky = ky, tc = 0;
  bool tok = true;
  if constexpr (TAIL) {
    const int lane = tid % L, j = lane / ntl;
    tok = j < L / ntl && ky + j < N1;
    tky = tok ? ky + j : ky;
    tc = tok ? lane - j * ntl : 0;
    tl = (tky - ky) * N2s + tc;
  }
  const int plane = N1 * N2s;  // per-slab offsets fit in 32 bits (checked by the launcher)
  const int64_t col = (int64_t)ky * N2s + kx0;
  const int plane_o = N1 * ldo;  // output rows of ldo bins
  const int64_t col_o = (int64_t)ky * ldo + kx0;
  constexpr bool TWO = KIND != 1;  // both target spectra
  const float2* Ftc = Ft + col;
  const float2* Ft2c = TWO ? Ft2 + col : nullptr;

  float2 F[EO], F2[TWO ? EO : 1];
#pragma unroll
  for (int i = 0; i < EO / PF::R_OUT; ++i)
#pragma unroll
    for (int r = 0; r < PF::R_OUT; ++r) {
      int l, e;
      PF::out_pos(tid, i, r, l, e);
      const int s = i * PF::R_OUT + r;
      const int lo = TAIL ? tl : l;
      F[s] = Ftc[e * plane + lo];
      if constexpr (TWO) F2[s] = Ft2c[e * plane + lo];
    }

  // input element of load slot k: rows e >= T0 are clamped to a valid row
  // (loaded, then replaced by zero)
  auto in_elem = [&](int k, int& l, int& e) {
    if constexpr (PR) {
      int j;
      PF::template slot<PF::R_IN>(tid, k, l, j);
      e = j;
    } else {
      PF::in_pos(tid, k / PF::R_IN, k % PF::R_IN, l, e);
    }
  };
  float2 x[NX], xn[NX];
  float2 xt[KIND == 2 ? NX : 1], xtn[KIND == 2 ? NX : 1];  // KIND 2: the template's input columns
  auto load_in = [&](int q, int b, float2* dst) {
    const float2* src = P + ((int64_t)(q * B + b) * T0) * (int64_t)plane + col;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      int l, e;
      in_elem(k, l, e);
      dst[k] = src[(e < T0 ? e : T0 - 1) * plane + (TAIL ? tl : l)];
    }
  };

  auto inv_store = [&](int o, int b, float2* w) {
    ct_fft2<PI, 1, L>(bx, by, tws, w, w, tid);
    after_transform();
    if constexpr (TAIL) {
      float2* dst = Qt + ((int64_t)(o * B + b) * N1 + tky) * (int64_t)(N0 * ntl) + tc;
#pragma unroll
      for (int i = 0; i < PI::E_OUT / PI::R_OUT; ++i)
#pragma unroll
        for (int r = 0; r < PI::R_OUT; ++r) {
          int l, e;
          PI::out_pos(tid, i, r, l, e);
          if (tok) dst[e * ntl] = w[i * PI::R_OUT + r];
        }
    } else {
      float2* dst = Q + ((int64_t)(o * B + b) * N0) * (int64_t)plane_o + col_o;
#pragma unroll
      for (int i = 0; i < PI::E_OUT / PI::R_OUT; ++i)
#pragma unroll
        for (int r = 0; r < PI::R_OUT; ++r) {
          int l, e;
          PI::out_pos(tid, i, r, l, e);
          // plain stores: with the non-temporal hint this pass ran slower
          // (512x512x200, 64 rotations: 8.29 + 4.51 ms vs 8.02 + 4.29 ms)
          dst[e * plane_o + l] = w[i * PI::R_OUT + r];
        }
    }
  };

  // rotations of this block: all of them, or a share of them where the
  // launch splits the batch over gridDim.y (the few tail blocks)
  const int b0 = (int)((int64_t)blockIdx.y * B / gridDim.y), b1 = (int)((int64_t)(blockIdx.y + 1) * B / gridDim.y);

  // Deferred stores (main launches, DEFER): the last output of rotation b
  // (S2, or X) is held in registers and stored in pieces after the LDS
  // stages of rotation b + 1's forward and first inverse transform, S1 after
  // those of S2's inverse.  Stored in one burst after its own transform, the
  // outputs of the CU's waves leave together (the block's waves meet at every
  // stage barrier) and the next transforms wait behind them; spread out, the
  // stores overlap the transforms (512x512x200, 64 rotations: S1/S2 kernel
  // 7.95 -> 6.53 ms, X kernel 4.15 -> 4.05 ms in 1024-thread blocks).
  constexpr int NE = PI::E_OUT, NH = PF::NST - 1;  // outputs per thread, hook points per transform
  float2 held[DEFER ? NE : 1];
  float2* held_dst = nullptr;
  auto out_ptr = [&](int o, int b) { return Q + ((int64_t)(o * B + b) * N0) * (int64_t)plane_o + col_o; };
  // piece K of N of a thread's outputs v
  auto store_piece = [&](float2* dst, const float2* v, auto k, auto n) {
    constexpr int K = decltype(k)::value, N = decltype(n)::value;
#pragma unroll
    for (int q = K * NE / N; q < (K + 1) * NE / N; ++q) {
      int l, e;
      PI::out_pos(tid, q / PI::R_OUT, q % PI::R_OUT, l, e);
      dst[e * plane_o + l] = v[q];
    }
  };
  using CN2 = std::integral_constant<int, 2 * NH>;
  using CN1 = std::integral_constant<int, NH>;

  // the first-stage input registers of a transform from the loaded columns
  auto expand = [&](const float2* xin, float2* w) {
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      int l, e;
      in_elem(k, l, e);
      const float2 v = e < T0 ? xin[k] : make_float2(0.f, 0.f);
      if constexpr (PR) {
#pragma unroll
        for (int r = 0; r < PF::R_IN; ++r) w[k * PF::R_IN + r] = r == 0 ? v : make_float2(0.f, 0.f);
      } else {
        w[k] = v;
      }
    }
  };

  // forward transform of the loaded columns into w.  Pruned columns: the
  // first stage only broadcasts each loaded value, so it is skipped and the
  // second stage reads the values themselves (ct_fft_bc0)
  auto forward = [&](const float2* xin, float2* w, auto hook) {
    if constexpr (PR && TM_PA_BCAST) {
      float2 xb[NX];
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        int l, e;
        in_elem(k, l, e);
        xb[k] = e < T0 ? xin[k] : make_float2(0.f, 0.f);
      }
      ct_fft2_bc0<PF, -1, L>(bx, by, tws, xb, w, tid, hook);
      after_transform();
    } else {
      expand(xin, w);
      ct_fft2<PF, -1, L>(bx, by, tws, w, w, tid, hook);
      after_transform();
    }
  };

  // one rotation; FIRST: no held output to store yet
  auto rotation = [&](int b, auto first) {
    constexpr bool FIRST = decltype(first)::value;
    const int bn = b + 1 < b1 ? b + 1 : b;  // the last rotation reloads itself, keeping the loop unpredicated
    load_in(KIND == 2 ? 0 : q_in, bn, xn);
    if constexpr (KIND == 2) load_in(1, bn, xtn);
    float2 w[EW];
    if constexpr (KIND == 2) {
      forward(x, w, ct_no_hook{});
      float2 t[EO];
#pragma unroll
      for (int s = 0; s < EO; ++s) {
        t[s] = c_mul(F2[s], w[s]);
        w[s] = c_mul(F[s], w[s]);
      }
      inv_store(0, b, w);
      inv_store(1, b, t);
      forward(xt, w, ct_no_hook{});
#pragma unroll
      for (int s = 0; s < EO; ++s) w[s] = c_mul(F[s], w[s]);
      inv_store(2, b, w);
    } else if constexpr (DEFER) {
      // the held output: first half over the forward's stages, second half
      // over the first inverse's
      auto fwd_hook = [&](auto st) {
        if constexpr (!FIRST) store_piece(held_dst, held, st, CN2{});
      };
      auto inv_hook = [&](auto st) {
        if constexpr (!FIRST) store_piece(held_dst, held, std::integral_constant<int, NH + decltype(st)::value>{}, CN2{});
      };
      forward(x, w, fwd_hook);
      if constexpr (KIND == 1) {
#pragma unroll
        for (int s = 0; s < EO; ++s) w[s] = c_mul(F[s], w[s]);
        ct_fft2<PI, 1, L>(bx, by, tws, w, w, tid, inv_hook);
        after_transform();
#pragma unroll
        for (int q = 0; q < NE; ++q) held[q] = w[q];
        held_dst = out_ptr(o_x, b);
      } else {
        float2 t[EO];
#pragma unroll
        for (int s = 0; s < EO; ++s) {
          t[s] = c_mul(F2[s], w[s]);
          w[s] = c_mul(F[s], w[s]);
        }
        ct_fft2<PI, 1, L>(bx, by, tws, w, w, tid, inv_hook);
        after_transform();
        float2* d1 = out_ptr(0, b);
        ct_fft2<PI, 1, L>(bx, by, tws, t, t, tid, [&](auto st) { store_piece(d1, w, st, CN1{}); });
        after_transform();
#pragma unroll
        for (int q = 0; q < NE; ++q) held[q] = t[q];
        held_dst = out_ptr(1, b);
      }
    } else {
      forward(x, w, ct_no_hook{});
      if constexpr (KIND == 1) {
#pragma unroll
        for (int s = 0; s < EO; ++s) w[s] = c_mul(F[s], w[s]);
        inv_store(o_x, b, w);
      } else {
        float2 t[EO];
#pragma unroll
        for (int s = 0; s < EO; ++s) {
          t[s] = c_mul(F2[s], w[s]);
          w[s] = c_mul(F[s], w[s]);
        }
        inv_store(0, b, w);
        inv_store(1, b, t);
      }
    }
#pragma unroll
    for (int k = 0; k < NX; ++k) x[k] = xn[k];
    ct_ready<NX>(x);
    if constexpr (KIND == 2) {
#pragma unroll
      for (int k = 0; k < NX; ++k) xt[k] = xtn[k];
      ct_ready<NX>(xt);
    }
  };

  load_in(KIND == 2 ? 0 : q_in, b0, x);
  if constexpr (KIND == 2) load_in(1, b0, xt);
  ct_ready<EO>(F);
  if constexpr (TWO) ct_ready<EO>(F2);
  ct_ready<NX>(x);
  if constexpr (KIND == 2) ct_ready<NX>(xt);
  if constexpr (DEFER) {
    if (b0 < b1) rotation(b0, std::true_type{});
    for (int b = b0 + 1; b < b1; ++b) rotation(b, std::false_type{});
    if (b0 < b1) store_piece(held_dst, held, std::integral_constant<int, 0>{}, std::integral_constant<int, 1>{});
  } else {
    for (int b = b0; b < b1; ++b) rotation(b, std::true_type{});
  }
}

// LDS row stride (complex elements) of a pass C tile row of M + 1 bins.
constexpr int pass_c_ld(int M) { return M + 1; }

// Pass C with the x inverse as a half-length complex transform: the N2 = 2M
// real outputs of a row are the real/imaginary parts of the M-point inverse
// of Z[k] = (X[k] + X*[M-k]) + i W^k (X[k] - X*[M-k]), W = exp(2 pi i/N2)
// (imaginary parts of the DC and Nyquist bins dropped, as numpy's irfft).
//
// A block owns RL rows and walks the (rotation, spectrum) tiles of the batch
// (FLC: S1, S2, X per rotation), keeping the running max/argmax of its rows
// in registers and read-modify-writing the state once at the end.  For FLC
// the local sum S1 is held only until S2 arrives; from then on the thread
// holds the factor scale / (sd * n), so the score is one multiply.  Tiles
// are prefetched PD deep into registers; every prefetch is issued (the tail
// re-reads the last tile), so waiting for the oldest one never waits for the
// younger ones.
//
// SPLIT: rows hold NM = 16 * floor(N2c / 16) main positions in Q (row
// stride ldq) and the remaining NTL bins in the tail array
// Qt[o][b][y][z][NTL] (row r = z * N1 + y), in the mirror-pair order of
// tm_pair_bin (h = ph): the main positions already hold the transform
// inputs Z (pass B computed them from each pair), the tail bins are the
// middle bins NM/2 .. M - NM/2, paired here.  Only the load slots that
// meet the tail (the last one or two) carry the pairing; Z is written to
// LDS at its bin and the transform reads it directly.
//
// R4 (split rows in the radix-4 units of tm_pair_bin, h = 1): the main
// positions hold the outputs of the first (radix-4) stage of the row's
// inverse, computed by pass B, and go to LDS where that stage would write
// them; the tail lanes pair the bins 0, G, 2G, 3G, M (G = M / 4) and run
// group 0's radix-4 stage across their lanes; the transform starts at its
// second stage.
template <int M, int RL, int NT, int NQ, int PD, bool SPLIT = false, int V = 0, bool R4 = false>
__global__ void __launch_bounds__(NT)
k_pass_c_ct(const float2* __restrict__ Q, const float2* __restrict__ Qt, const float2* __restrict__ tw2,
            int64_t nrows, int ldq, int N0, int N1, int ntl, int ph, int B,
            const double* __restrict__ n_obs, const float* __restrict__ inv_den, const int* __restrict__ rot_idx,
            float scale, float eps, float* __restrict__ max_scores, int* __restrict__ max_rot) {
  // the inverse plan: the forward plan's stages reversed, or with R4 the
  // order whose first stage is radix 4
  using PI = CtPlan<M, RL, NT, false, R4 ? CtPlan<M, RL, NT, false, false, V>::radix(0) != 4 : true, V>;
  constexpr int N2 = 2 * M, N2c = M + 1, LD = pass_c_ld(M);
  constexpr int EI = PI::E_IN, EO = PI::E_OUT, EW = EI > EO ? EI : EO;
  constexpr int NM = N2c / 16 * 16, NTL = N2c - NM, NP = NM / 2;  // split rows
  constexpr int G = M / 4;  // R4: first-stage groups
  static_assert(!R4 || (SPLIT && PI::radix(0) == 4 && M % 4 == 0 && G % 2 == 1 && NM == M - 4 && NTL == 5),
                "radix-4 units: M = 4 G with G odd, tail bins 0, G, 2G, 3G, M");
  // load slots per thread and tile.  Split rows: P4 slots of two
  // consecutive main positions (16-byte loads, NM is even and the rows
  // 16-byte aligned), then the tail slots of one bin each; otherwise one
  // bin per slot
  constexpr int P4 = SPLIT ? (RL * NM / 2 + NT - 1) / NT : 0;
  // tail lanes per row: one per tail bin, or with R4 one lane quad (lane c
  // pairs bins G c and M - G c, c = 0..3)
  constexpr int TLR = R4 ? 4 : NTL;
  constexpr int NPF = SPLIT ? P4 + (RL * TLR + NT - 1) / NT : (RL * N2c + NT - 1) / NT;
  static_assert(PD >= 1 && PD <= 3, "prefetch depth 1 to 3");
  // the tile loop is unrolled by U so that the spectrum index o and the
  // prefetch buffer of every step are compile-time constants
  constexpr int U = NQ % PD == 0 ? NQ : (PD % NQ == 0 ? PD : NQ * PD);
  extern __shared__ float2 lds[];
  __shared__ float2 tws[N2];  // twiddles in LDS: no vector-memory waits inside the transforms
  const int tid = threadIdx.x;
  for (int i = tid; i < N2; i += NT) tws[i] = tw2[i];
  // XCD-grouped row blocks (the grid is a multiple of 8; blocks past the
  // rows exit).  A block takes RL consecutive rows (z, y), or with split
  // rows the RL rows (z0 .. z0 + RL - 1, y) of one y, whose tail bins are
  // one contiguous run of Qt[y][z][ntl] (consecutive blocks take the next
  // z group of the same y and share its cache lines)
  const int64_t blk = tm_xcd_block();
  int64_t r0 = 0;
  int rl, zg0 = 0, yg = 0;
  if constexpr (SPLIT) {
    const int nzb = (N0 + RL - 1) / RL;
    if (blk >= (int64_t)N1 * nzb) return;
    yg = (int)(blk / nzb);
    zg0 = (int)(blk - (int64_t)yg * nzb) * RL;
    rl = N0 - zg0 < RL ? N0 - zg0 : RL;
  } else {
    r0 = blk * RL;
    if (r0 >= nrows) return;
    rl = (int)((nrows - r0) < RL ? (nrows - r0) : RL);
  }
  // row (z * N1 + y) of the block's local row l
  auto row_of = [&](int l) -> int64_t { return SPLIT ? (int64_t)(zg0 + l) * N1 + yg : r0 + l; };
  const int tile_n = rl * N2c;
  const int iters = B * NQ;

  float best0[EO], best1[EO], d0[EO], d1[EO];
  int arg0[EO], arg1[EO];
#pragma unroll
  for (int s = 0; s < EO; ++s) {
    best0[s] = best1[s] = -INFINITY;
    arg0[s] = arg1[s] = -1;
    d0[s] = d1[s] = 0.f;
  }
  // Each load slot's offset within a slab (o, b) and its LDS position are
  // the same for every tile: computed once.  Rows are consecutive at stride
  // ldq from row r0; split rows: main slots offset from the slab start of
  // Q, tail slots from the slab start of Qt.  Unpredicated: slots past the
  // tile (or past the block's rl rows) re-read a valid bin and write the
  // junk line RL.
  int offs[NPF], lpos[NPF];
  int lpos2[SPLIT ? P4 : 1];  // split rows: LDS position of the slot's second bin
  uint32_t boffs[NPF];  // byte offsets of offs (main slots: offs >= 0)
  int offs2[NPF];      // split rows, tail slots: the pair partner's offset
  float2 wtw[NPF];     // and W^k of its bin
  // split rows: slots from P4 on hold tail bins
  auto tail_slot = [](int p) { return SPLIT && p >= P4; };
#pragma unroll
  for (int p = 0; p < NPF; ++p) {
    if constexpr (SPLIT) {
      if (p < P4) {
        // positions q, q + 1 of row l (rows past rl re-read row rl - 1 and
        // write the junk line; slots past the tile re-read position 0)
        const int g = tid + p * NT, g2 = g < RL * NM / 2 ? g : 0;
        const int l = g2 / (NM / 2), q = 2 * (g2 - l * (NM / 2));
        const int lc = l < rl ? l : rl - 1;
        const bool ok = g < RL * NM / 2 && l < rl;
        offs[p] = (int)row_of(lc) * ldq + q;
        if constexpr (R4) {
          // first-stage output 4 jj + r of group jj; quad lane m holds r =
          // (m >> 1) | ((m & 1) << 1)
          auto stage0 = [](int qq) {
            const int c = qq & 7, m = c & 3, j = (qq >> 3) + 1;
            return 4 * (c < 4 ? j : G - j) + ((m >> 1) | ((m & 1) << 1));
          };
          lpos[p] = ok ? l * LD + stage0(q) : RL * LD;
          lpos2[p] = ok ? l * LD + stage0(q + 1) : RL * LD;
        } else {
          lpos[p] = ok ? l * LD + tm_pair_bin(q, NM, M, ph) : RL * LD;
          lpos2[p] = ok ? l * LD + tm_pair_bin(q + 1, NM, M, ph) : RL * LD;
        }
        offs2[p] = offs[p];
        wtw[p] = make_float2(0.f, 0.f);
      } else {
        // tail bin c of row (z, y) = (zg0 + l, yg) sits at Qt[y][z][c]
        const int t = tid + (p - P4) * NT, t2 = t < RL * TLR ? t : 0;
        const int l = t2 / TLR, c = t2 - l * TLR;
        const int lc = l < rl ? l : rl - 1;
        const int base = (yg * N0 + zg0 + lc) * NTL;
        offs[p] = base + c;
        offs2[p] = base + NTL - 1 - c;
        if constexpr (R4) {
          // bin G c; the quad's radix-4 leaves first-stage output
          // (c >> 1) | ((c & 1) << 1) of group 0 in lane c
          lpos[p] = t < RL * TLR && l < rl ? l * LD + ((c >> 1) | ((c & 1) << 1)) : RL * LD;
          wtw[p] = tw_load<1>(tw2, G * c);
        } else {
          lpos[p] = t < RL * NTL && l < rl ? l * LD + NP + c : RL * LD;
          wtw[p] = tw_load<1>(tw2, NP + c);
        }
      }
    } else {
      const int f = tid + p * NT;
      const int fc = f < tile_n ? f : tile_n - 1;
      const int l = fc / N2c, e = fc - l * N2c;
      offs[p] = l * ldq + e;
      lpos[p] = f < tile_n ? l * LD + e : RL * LD;
    }
    boffs[p] = (uint32_t)offs[p] * (uint32_t)sizeof(float2);
  }
  // FLC: scale / n of every rotation of the batch, once per block (the
  // same float32 division the epilogue would repeat per rotation)
  float* snl = reinterpret_cast<float*>(lds + (RL + 1) * LD);
  if constexpr (NQ == 3) {
    for (int b = tid; b < B; b += NT) snl[b] = scale / (float)n_obs[b];
  }
  // pf[p]: the slot's bin (main slots of split rows: the first of two, the
  // second in pf2[p]); pf2[p] of a tail slot: the pair partner's bin
  // R4: the last stage's twiddle powers of this thread, held for the
  // block's 3B transforms (read from the twiddle table in global memory:
  // the LDS copy is not yet ordered for every thread here)
  float2 wlast[R4 && TM_PC_HOLD_TWIDDLES ? EO : 1];
  if constexpr (R4 && TM_PC_HOLD_TWIDDLES) ct_last_twiddles<PI, 1, 2>(tw2, wlast, tid);
  // and, with TM_PC_HOLD_TWIDDLES 2, those of the second stage
  constexpr int W1N = PI::bpt(PI::radix(1)) * PI::radix(1);
  float2 wfirst[R4 && TM_PC_HOLD_TWIDDLES == 2 ? W1N : 1];
  if constexpr (R4 && TM_PC_HOLD_TWIDDLES == 2) ct_last_twiddles<PI, 1, 2, 1>(tw2, wfirst, tid);
  // tail slot p has live lanes in this wave (wave-uniform: the tail lanes
  // are the first RL * TLR of the slot, so other waves skip its loads and
  // its pairing)
  const int wave0 = __builtin_amdgcn_readfirstlane(tid);
  auto tail_live = [&](int p) { return wave0 + (p - P4) * NT < RL * TLR; };
  auto prefetch = [&](int it, float2* pf, float2* pf2) {
    const int b = it / NQ, o = it - (it / NQ) * NQ;
    const float2* src = Q + ((int64_t)(o * B + b) * nrows + r0) * ldq;  // r0 = 0 with split rows
    const float2* tsrc = SPLIT ? Qt + (int64_t)(o * B + b) * nrows * NTL : nullptr;
    // a uniform base and a 32-bit byte offset per slot
    const char* srcb = reinterpret_cast<const char*>(src);
#pragma unroll
    for (int p = 0; p < NPF; ++p) {
      if constexpr (SPLIT) {
        if (p < P4) {
          const float4 v = tm_ld_stream2(reinterpret_cast<const float2*>(srcb + boffs[p]));
          pf[p] = make_float2(v.x, v.y);
          pf2[p] = make_float2(v.z, v.w);
        } else if (tail_live(p)) {
          pf[p] = tm_ld_stream(tsrc + offs[p]);
          pf2[p] = tm_ld_stream(tsrc + offs2[p]);
        }
      } else {
        pf[p] = tm_ld_stream(reinterpret_cast<const float2*>(srcb + boffs[p]));
      }
    }
  };
  auto step = [&](int it, int o, float2* pf, float2* pf2) {
    const int b = it / NQ;
    lds_barrier();  // the previous transform has finished reading the tile
#pragma unroll
    for (int p = 0; p < NPF; ++p) {
      if constexpr (SPLIT) {
        if (p < P4) {
          lds[lpos[p]] = pf[p];
          lds[lpos2[p]] = pf2[p];
        } else if (!tail_live(p)) {
        } else if constexpr (R4) {
          // group 0 (bins 0, G, 2G, 3G): Z of the pair (G c, M - G c) in
          // lane c of the row's quad, then the radix-4 stage across the quad
          const int c = tid & 3;
          lds[lpos[p]] = tm_quad_idft4(tm_c2r_pre(pf[p], pf2[p], wtw[p], c == 0), c);
        } else {
          lds[lpos[p]] = tm_c2r_pre(pf[p], pf2[p], wtw[p], false);
        }
      } else {
        lds[lpos[p]] = pf[p];
      }
    }
    prefetch(it + PD < iters ? it + PD : iters - 1, pf, pf2);
    lds_barrier();
    float2 z[EW];
    if constexpr (R4) {
      ct_fft_from1<PI, 1, LD, 2>(lds, tws, z, tid, TM_PC_HOLD_TWIDDLES ? wlast : nullptr,
                                 TM_PC_HOLD_TWIDDLES == 2 ? wfirst : nullptr);
    } else {
#pragma unroll
      for (int i = 0; i < EI / PI::R_IN; ++i)
#pragma unroll
        for (int r = 0; r < PI::R_IN; ++r) {
          int l, k;
          if (!PI::in_pos(tid, i, r, l, k)) l = RL;  // idle slot: junk line (never stored)
          if constexpr (SPLIT) {
            z[i * PI::R_IN + r] = lds[l * LD + k];
          } else {
            z[i * PI::R_IN + r] = tm_c2r_pre(lds[l * LD + k], lds[l * LD + (M - k)], tw_load<1>(tws, k), k == 0);
          }
        }
      ct_fft<PI, 1, LD, 2>(lds, tws, z, z, tid);
    }
    const float sn = NQ == 3 ? snl[b] : 0.f;
    const int ridx = rot_idx[b];
    // Every output slot is updated (idle slots and rows past rl hold junk
    // that is never stored); the branch on o is uniform.  FLC works on the
    // unscaled outputs (the scale is in sn and in the held factor); the
    // correlation scores scale first.
    if (o == NQ - 1) {
#pragma unroll
      for (int i = 0; i < EO / PI::R_OUT; ++i)
#pragma unroll
        for (int r = 0; r < PI::R_OUT; ++r) {
          int l, m;
          const int s = i * PI::R_OUT + r;
          const bool ok = PI::out_pos(tid, i, r, l, m) && l < rl;
          float s0, s1;
          if (NQ == 3) {
            s0 = z[s].x * d0[s];
            s1 = z[s].y * d1[s];
          } else {
            s0 = z[s].x * scale;
            s1 = z[s].y * scale;
            if (inv_den) {
              const float2 d = *reinterpret_cast<const float2*>(inv_den + row_of(ok ? l : 0) * N2 + 2 * m);
              s0 *= d.x;
              s1 *= d.y;
            }
          }
          if (s0 > best0[s]) {
            best0[s] = s0;
            arg0[s] = ridx;
          }
          if (s1 > best1[s]) {
            best1[s] = s1;
            arg1[s] = ridx;
          }
        }
    } else if (o == 0) {
#pragma unroll
      for (int s = 0; s < EO; ++s) {
        d0[s] = z[s].x;  // S1 until S2 arrives
        d1[s] = z[s].y;
      }
    } else {
#pragma unroll
      for (int s = 0; s < EO; ++s) {
        d0[s] = fu_flc_factor(z[s].x, d0[s], sn, eps * eps);
        d1[s] = fu_flc_factor(z[s].y, d1[s], sn, eps * eps);
      }
    }
  };

  float2 pf[PD][NPF], pf2[PD][NPF];
#pragma unroll
  for (int q = 0; q < PD; ++q) prefetch(q < iters ? q : iters - 1, pf[q], pf2[q]);
  for (int it0 = 0; it0 < iters; it0 += U) {  // iters is a multiple of U (checked by the launcher)
#pragma unroll
    for (int u = 0; u < U; ++u) step(it0 + u, u % NQ, pf[u % PD], pf2[u % PD]);
  }
#pragma unroll
  for (int i = 0; i < EO / PI::R_OUT; ++i)
#pragma unroll
    for (int r = 0; r < PI::R_OUT; ++r) {
      int l, m;
      const int s = i * PI::R_OUT + r;
      if (!PI::out_pos(tid, i, r, l, m) || l >= rl) continue;
      const int64_t g = row_of(l) * N2 + 2 * m;
      float2 old = *reinterpret_cast<float2*>(max_scores + g);
      int2 oldr = *reinterpret_cast<int2*>(max_rot + g);
      bool upd = false;
      if (best0[s] > old.x) {
        old.x = best0[s];
        oldr.x = arg0[s];
        upd = true;
      }
      if (best1[s] > old.y) {
        old.y = best1[s];
        oldr.y = arg1[s];
        upd = true;
      }
      if (upd) {
        *reinterpret_cast<float2*>(max_scores + g) = old;
        *reinterpret_cast<int2*>(max_rot + g) = oldr;
      }
    }
}

// ntl > 0: split rows (main rows of ldo bins in Q, ntl tail bins in Qt); the
// tail tile (kx_base = ldo) is a second launch of each kind.
template <int V, int N0, int L, int NT, int NT0, bool PR>
static int launch_pass_a_ct_pr(bool flc, const float2* P, const float2* ft, const float2* ft2, float2* Q,
                               float2* Qt, const float2* tw, int T0, int N1, int N2s, int ldo, int nm, int ntl,
                               int tpr, int B, hipStream_t s) {
  const int64_t blocks = (int64_t)N1 * tpr;
  const size_t shm = 2 * (size_t)N0 * L * sizeof(float2);  // two exchange buffers
  // algorithmic bytes of a launch over `bins` bins of every row: the target
  // spectra it multiplies by (nin), the T0 pruned planes of nq template-side
  // inputs per rotation and its nout outputs per rotation
  auto bytes_of = [&](int bins, int nin, int nout, int nq = 1) {
    return 8.0 * bins * N1 * ((double)N0 * nin + (double)T0 * nq * B + (double)N0 * nout * B);
  };
  const unsigned tail_blocks = ntl > 0 ? (unsigned)tm_ceil_div(N1, L / ntl) : 0;  // L / ntl rows per tail tile
  // The tail tiles are fewer than the CUs: their rotations are split over
  // gridDim.y, s ways, s <= 8 chosen so that the blocks fill the rounds of
  // resident blocks best (ties: fewer splits, each split re-reads its
  // target tile).  512 x 512 x 200: 171 tiles, s = 7 fills 93 % of the
  // last round where s = 3 left it with one block.
  auto tail_split_for = [&](const void* kernel, int nt) -> unsigned {
    if (ntl == 0) return 1;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, nt, shm) != hipSuccess || per < 1) per = 1;
    const int64_t R = (int64_t)tm_device_cus() * per;
    unsigned best = 1;
    double best_u = 0.0;
    for (int64_t sp = 1; sp <= std::min<int64_t>(B, 8); ++sp) {
      const int64_t n = (int64_t)tail_blocks * sp;
      const double u = (double)n / (double)(tm_ceil_div(n, R) * R);
      if (u > best_u + 1e-9) {
        best_u = u;
        best = (unsigned)sp;
      }
    }
    return best;
  };
  // pruned main launches in 1024-thread blocks store deferred (8 outputs
  // per thread; with 16, in the smaller blocks of the 128- and 256-point
  // kernels or unpruned, the held outputs push the kernels past 128 VGPRs or
  // into spills); tail launches store directly, the X tail in blocks of at
  // most 512 threads
  constexpr bool DS = PR && NT0 >= 1024 && CtPlan<N0, L, NT0, true, false, V>::NST >= 2;
  constexpr bool DX = PR && NT >= 1024 && CtPlan<N0, L, NT, true, false, V>::NST >= 2;
  constexpr int NTT = NT > 512 ? 512 : NT;
  // FLC tails in one launch (KIND 2) for pruned transforms of up to 512
  // points (1024-point or unpruned ones would spill with five spectra)
  constexpr bool TAIL1 = N0 <= 512 && PR;
  if (flc) {
    TM_TIMED("fused_pass_a", bytes_of(nm, 2, 2), s,
             hipLaunchKernelGGL((k_pass_a_ct<N0, L, NT0, 0, PR, false, V, DS>), dim3((unsigned)blocks), dim3(NT0), shm, s, P, ft, ft2,
                                Q, Qt, tw, T0, N1, N2s, ldo, B, tpr, 0, ntl, 0, 0));
    TM_CHECK_LAUNCH();
    if (ntl > 0 && TAIL1) {
      TM_TIMED("fused_pass_a_tail", bytes_of(ntl, 2, 3, 2), s,
               hipLaunchKernelGGL((k_pass_a_ct<N0, L, NT0, 2, PR, true, V>),
                                  dim3(tail_blocks, tail_split_for((const void*)k_pass_a_ct<N0, L, NT0, 2, PR, true, V>, NT0)), dim3(NT0),
                                  shm, s, P, ft, ft2, Q, Qt, tw, T0, N1, N2s, ldo, B, 1, nm, ntl, 0, 2));
      TM_CHECK_LAUNCH();
    } else if (ntl > 0) {
      TM_TIMED("fused_pass_a_tail", bytes_of(ntl, 2, 2), s,
               hipLaunchKernelGGL((k_pass_a_ct<N0, L, NT0, 0, PR, true, V>),
                                  dim3(tail_blocks, tail_split_for((const void*)k_pass_a_ct<N0, L, NT0, 0, PR, true, V>, NT0)), dim3(NT0), shm, s, P,
                                  ft, ft2, Q, Qt, tw, T0, N1, N2s, ldo, B, 1, nm, ntl, 0, 0));
      TM_CHECK_LAUNCH();
    }
  }
  TM_TIMED(flc ? "fused_pass_a_x" : "fused_pass_a", bytes_of(nm, 1, 1), s,
           hipLaunchKernelGGL((k_pass_a_ct<N0, L, NT, 1, PR, false, V, DX>), dim3((unsigned)blocks), dim3(NT), shm, s, P, ft,
                              (const float2*)nullptr, Q, Qt, tw, T0, N1, N2s, ldo, B, tpr, 0, ntl, flc ? 1 : 0,
                              flc ? 2 : 0));
  TM_CHECK_LAUNCH();
  if (ntl > 0 && !(flc && TAIL1)) {
    TM_TIMED(flc ? "fused_pass_a_x_tail" : "fused_pass_a_tail", bytes_of(ntl, 1, 1), s,
             hipLaunchKernelGGL((k_pass_a_ct<N0, L, NTT, 1, PR, true, V>),
                                dim3(tail_blocks, tail_split_for((const void*)k_pass_a_ct<N0, L, NTT, 1, PR, true, V>, NTT)), dim3(NTT), shm, s, P, ft,
                                (const float2*)nullptr, Q, Qt, tw, T0, N1, N2s, ldo, B, 1, nm, ntl, flc ? 1 : 0,
                                flc ? 2 : 0));
    TM_CHECK_LAUNCH();
  }
  return TM_OK;
}

// NT: threads of the X kernel, NT0: threads of the FLC mask kernel
template <int V, int N0, int L, int NT, int NT0 = NT>
static int launch_pass_a_ct(bool flc, const float2* P, const float2* ft, const float2* ft2, float2* Q, float2* Qt,
                            const float2* tw, int T0, int N1, int N2c, int N2s, int ldo, int ntl, int B,
                            hipStream_t s) {
  TM_REQUIRE((int64_t)N0 * N1 * N2s < (1LL << 31), "pass A: z column stride too large");
  const int nm = N2c - ntl;  // main bins (split rows) or N2c
  const int tpr = ntl > 0 ? nm / L : (int)tm_ceil_div(N2c, L);
  TM_REQUIRE((int64_t)tpr * L <= N2s, "pass A: spectra rows not padded to whole tiles");
  TM_REQUIRE(ntl == 0 || (nm % L == 0 && nm <= ldo && ntl <= L && nm + L <= N2s),
             "pass A: split rows need main rows of whole tiles and one tail tile inside the padded input rows");
  using PF = CtPlan<N0, L, NT, true, false, V>;
  if (T0 <= N0 / PF::R_IN)
    return launch_pass_a_ct_pr<V, N0, L, NT, NT0, true>(flc, P, ft, ft2, Q, Qt, tw, T0, N1, N2s, ldo, nm, ntl, tpr, B,
                                                        s);
  return launch_pass_a_ct_pr<V, N0, L, NT, NT0, false>(flc, P, ft, ft2, Q, Qt, tw, T0, N1, N2s, ldo, nm, ntl, tpr,
                                                       B, s);
}

template <int V, int M, int RL, int NT, int PD3, int PD1, bool SPLIT = false>
static int launch_pass_c_ct(bool flc, const float2* Q, const float2* Qt, const float2* tw2, int64_t nrows, int ldq,
                            int N0, int N1, int ntl, int ph, int B, const double* n_obs, const float* inv_den,
                            const int* rot_idx, float scale, float eps, float* max_scores, int* max_rot, double bytes,
                            hipStream_t s) {
  TM_REQUIRE(!SPLIT || (ntl == M + 1 - (M + 1) / 16 * 16 && ldq >= M + 1 - ntl && Qt && nrows == (int64_t)N0 * N1 &&
                        (ph == 8 || ph == 4 || ph == 1)),
             "pass C: split rows need the tail array of the last N2c %% 16 bins, mirror pairs and nrows = N0 * N1");
  TM_REQUIRE(!SPLIT || nrows * ldq < (1LL << 31), "pass C: main offsets exceed 32 bits");
  // PD3 / PD1: prefetch depth of the FLC (3 spectra) / correlation kernels;
  // the step count B * NQ must be a multiple of the unroll lcm(NQ, PD)
  // a multiple of 8: XCD-grouped blocks; split rows: N1 y values of
  // ceil(N0 / RL) z groups
  const int64_t row_blocks = SPLIT ? (int64_t)N1 * tm_ceil_div(N0, RL) : tm_ceil_div(nrows, RL);
  const int64_t blocks = tm_ceil_div(row_blocks, 8) * 8;
  TM_REQUIRE(blocks <= 0x7fffffff, "pass C grid too large");
  TM_REQUIRE(!SPLIT || (int64_t)N0 * N1 * ntl < (1LL << 31), "pass C: tail offsets exceed 32 bits");
  // the tile + the junk line, then scale / n of the batch's rotations (FLC)
  const size_t shm = (size_t)(RL + 1) * pass_c_ld(M) * sizeof(float2) + (flc ? sizeof(float) * (size_t)B : 0);
  static_assert(PD3 == 1 || PD3 == 3, "FLC prefetch depth must divide or be a multiple of 3");
  auto go = [&](auto r4tag) -> int {
    constexpr bool R4 = decltype(r4tag)::value;
    if (flc) {
      TM_TIMED("fused_pass_c", bytes, s,
               hipLaunchKernelGGL((k_pass_c_ct<M, RL, NT, 3, PD3, SPLIT, V, R4>), dim3((unsigned)blocks), dim3(NT), shm, s, Q,
                                  Qt, tw2, nrows, ldq, N0, N1, ntl, ph, B, n_obs, (const float*)nullptr, rot_idx, scale, eps,
                                  max_scores, max_rot));
    } else if (B % PD1 == 0) {
      TM_TIMED("fused_pass_c", bytes, s,
               hipLaunchKernelGGL((k_pass_c_ct<M, RL, NT, 1, PD1, SPLIT, V, R4>), dim3((unsigned)blocks), dim3(NT), shm, s, Q,
                                  Qt, tw2, nrows, ldq, N0, N1, ntl, ph, B, (const double*)nullptr, inv_den, rot_idx, scale,
                                  eps, max_scores, max_rot));
    } else {
      TM_TIMED("fused_pass_c", bytes, s,
               hipLaunchKernelGGL((k_pass_c_ct<M, RL, NT, 1, 1, SPLIT, V, R4>), dim3((unsigned)blocks), dim3(NT), shm, s, Q,
                                  Qt, tw2, nrows, ldq, N0, N1, ntl, ph, B, (const double*)nullptr, inv_den, rot_idx, scale,
                                  eps, max_scores, max_rot));
    }
    TM_CHECK_LAUNCH();
    return TM_OK;
  };
  // radix-4 units: rows of M + 1 = 4 G + 1 bins, G odd, 5 tail bins
  if constexpr (SPLIT && M % 4 == 0 && (M / 4) % 2 == 1 && (M + 1) % 16 == 5) {
    if (ph == 1) return go(std::true_type{});
  }
  TM_REQUIRE(ph != 1, "pass C: radix-4 units need rows of 4 G + 1 bins with G odd and 5 tail bins");
  return go(std::false_type{});
}

// --------------------------------------------------------------------------
// plan

struct tm_scan_plan {
  int mode;  // 0 FLC, 1 CORR
  int nq_in, nq_out;
  int N[3], N2c, N2s;  // N2s: row stride of the workspace spectra (N2c rounded up to 16 bins = 128 B)
  int N2m, ntl;        // correlation spectra rows: N2m bins in Q, ntl tail bins in Qt (split rows) or N2s, 0
  int pair_h;          // split rows: x bins in mirror-pair order (tm_pair_bin, h = pair_h) in every spectrum of the scan
  int T[3], T2c;
  int max_batch;
  int normalize;
  double mask_sum;
  const float* templ;
  const float* mask;
  const float* filter;
  const float2* ft;   // target spectra at row stride N2s (the caller's or padded copies)
  const float2* ft2;
  float2 *ftp, *ft2p;  // padded copies, owned by the plan
  const float* inv_den;
  tm_fft_plan* fast;   // line plans of N0, N1, N2
  tm_fft_plan* tplan;  // template-size plan (filter)
  // workspace
  float *TR, *MR;
  float2 *specT, *Q1, *P, *Q, *Qt;
  double *n_obs, *rot_ws, *mom_ws, *ms;
  // graph replay of the batch sequence (kernels only) on a plan stream
  int a_cols;  // pass A column tile: 16 (default where the rows allow it) or 8
  int c_rows;  // pass C rows per block for N2 = 200: 5 (default, 128 threads) or 10 (256 threads)
  int radix;   // radix plan variant of the compile-time passes (fft_ct.h ct_factor_v): 0 or 1
  int use_graph;
  hipStream_t gs;
  hipEvent_t ev_in, ev_out;
  float* rbuf;  // rotations and analyzer indices of the batch, read by the graph
  int* ibuf;
  // captured batch sequences: one per (batch size, state arrays), two kept
  // (the full batch and a call's remainder batch), least recently used
  // replaced; a configuration runs eagerly once (warm), is captured on its
  // second occurrence and replayed after that
  struct GraphSlot {
    hipGraphExec_t exec;
    int nb, warm;
    float* ms;
    int* mr;
    uint64_t used;
  } gslot[2];
  uint64_t g_clock;
  int64_t g_captures, g_replays, g_eager;
};

extern "C" int tm_fft_plan_create(int ndim, const int64_t* shape, int64_t batch, tm_fft_plan** out);
extern "C" int tm_fft_plan_destroy(tm_fft_plan* p);

static void scan_plan_free(tm_scan_plan* p) {
  if (!p) return;
  void* bufs[] = {p->TR, p->MR, p->specT, p->Q1, p->P, p->Q, p->Qt, p->n_obs, p->rot_ws, p->mom_ws, p->ms, p->ftp, p->ft2p,
                  p->rbuf, p->ibuf};
  for (void* b : bufs)
    if (b) hipFree(b);
  for (auto& g : p->gslot)
    if (g.exec) hipGraphExecDestroy(g.exec);
  if (p->ev_in) hipEventDestroy(p->ev_in);
  if (p->ev_out) hipEventDestroy(p->ev_out);
  if (p->gs) hipStreamDestroy(p->gs);
  if (p->fast) tm_fft_plan_destroy(p->fast);
  if (p->tplan) tm_fft_plan_destroy(p->tplan);
  free(p);
}

#define TM_ALLOC(ptr, bytes)                               \
  do {                                                     \
    hipError_t _e = hipMalloc((void**)&(ptr), (bytes));    \
    if (_e != hipSuccess) {                                \
      tm_set_error("hipMalloc(%zu) failed: %s", (size_t)(bytes), hipGetErrorString(_e)); \
      scan_plan_free(p);                                   \
      return _e == hipErrorOutOfMemory ? TM_ERR_OOM : TM_ERR_HIP; \
    }                                                      \
  } while (0)

// Workspace spectra rows are padded to a multiple of 16 bins (128 B), so
// every 16-bin column tile of the z and y passes is one aligned 128-byte
// segment per row.
// (8-bin rows would save the padding traffic but 64-byte row segments
// stream measurably slower than 128-byte ones on MI355X.)
static inline int64_t pad_bins(int64_t n) { return (n + 15) / 16 * 16; }

// Row layout of the correlation spectra.  Split rows where the compile-time
// passes run (x length 200: 101 bins, 96 per row in Q and the last 5 in the
// dense tail array instead of 112-bin rows; x length 256: 129 bins, 128 + 1
// instead of 144; z length 128, 256 or 512, or 1024 for x length 256; y
// length 128, 256, 512 or 1024), so no pass moves the padding bins;
// otherwise rows padded to N2s bins.  Split rows keep the x bins of every
// spectrum of the scan in the mirror-pair order of tm_pair_bin: pass B's
// y transform then ends with the pre-processing of pass C's half-length x
// inverse (tm_c2r_pre), which needs the pair (k, M - k) of a row.
#ifndef TM_SPLIT_ROWS
#define TM_SPLIT_ROWS 1
#endif
// 1: split rows of the 200-point x axis in radix-4 units (tm_pair_bin h = 1)
// where the y axis has a 16-line kernel; 0: mirror pairs there too.
#ifndef TM_R4_UNITS
#define TM_R4_UNITS 1
#endif
// N2m is the row stride of Q: the main bins, plus 16 unused bins where they
// are a multiple of 64 (with a 1 KB stride the 1024-point pass B over 64-byte
// segments ran 27.3 ms, with 1088 B 28.5 ms: lines split between rows).
static inline void row_layout(int64_t N0, int64_t N1, int64_t N2, int64_t& N2m, int64_t& ntl) {
  const int64_t N2c = N2 / 2 + 1;
  const bool z_ct = N0 == 128 || N0 == 256 || N0 == 512;
  const bool y_ct = N1 == 128 || N1 == 256 || N1 == 512 || N1 == 1024;
  const bool split = TM_SPLIT_ROWS && y_ct && ((N2 == 200 && z_ct) || (N2 == 256 && (z_ct || N0 == 1024)));
  const int64_t nm = N2c / 16 * 16;
  N2m = split ? (nm % 64 == 0 ? nm + 16 : nm) : pad_bins(N2c);
  ntl = split ? N2c - nm : 0;
}

// Rows of n2c natural-order bins into rows of n2s positions in mirror-pair
// order (padding positions zero); one block per 8 rows.
__global__ void __launch_bounds__(256)
k_pair_rows(const float2* __restrict__ src, float2* __restrict__ dst, int64_t rows, int n2c, int n2s, int nm, int h) {
  const int64_t r = (int64_t)blockIdx.x * 8 + threadIdx.x / 32;
  if (r >= rows) return;
  for (int q = threadIdx.x % 32; q < n2s; q += 32)
    dst[r * n2s + q] = q < n2c ? src[r * n2c + tm_pair_bin(q, nm, n2c - 1, h)] : make_float2(0.f, 0.f);
}

// Copy the caller's target spectra (rows of N2c bins) into the padded copies
// (in mirror-pair order with split rows).
static int scan_copy_target(tm_scan_plan* p, const void* ft, const void* ft2, hipStream_t s) {
  if (p->N2s == p->N2c) {
    p->ft = (const float2*)ft;
    p->ft2 = (const float2*)ft2;
    return TM_OK;
  }
  const size_t rows = (size_t)p->N[0] * p->N[1];
  if (p->pair_h) {
    const int nm = p->N2c - p->ntl;
    hipLaunchKernelGGL(k_pair_rows, dim3((unsigned)tm_ceil_div((int64_t)rows, 8)), dim3(256), 0, s,
                       (const float2*)ft, p->ftp, (int64_t)rows, p->N2c, p->N2s, nm, p->pair_h);
    TM_CHECK_LAUNCH();
    if (p->mode == 0) {
      hipLaunchKernelGGL(k_pair_rows, dim3((unsigned)tm_ceil_div((int64_t)rows, 8)), dim3(256), 0, s,
                         (const float2*)ft2, p->ft2p, (int64_t)rows, p->N2c, p->N2s, nm, p->pair_h);
      TM_CHECK_LAUNCH();
    }
    p->ft = p->ftp;
    p->ft2 = p->mode == 0 ? p->ft2p : nullptr;
    return TM_OK;
  }
  const size_t spitch = sizeof(float2) * p->N2c, dpitch = sizeof(float2) * p->N2s;
  TM_CHECK_HIP(hipMemcpy2DAsync(p->ftp, dpitch, ft, spitch, spitch, rows, hipMemcpyDeviceToDevice, s));
  if (p->mode == 0)
    TM_CHECK_HIP(hipMemcpy2DAsync(p->ft2p, dpitch, ft2, spitch, spitch, rows, hipMemcpyDeviceToDevice, s));
  p->ft = p->ftp;
  p->ft2 = p->mode == 0 ? p->ft2p : nullptr;
  return TM_OK;
}

extern "C" {

int64_t tm_scan_workspace_bytes(int mode, const int64_t* fast_shape, const int* tshape, int max_batch,
                                int with_filter) {
  const int64_t N0 = fast_shape[0], N1 = fast_shape[1], N2c = fast_shape[2] / 2 + 1, N2s = pad_bins(N2c);
  const int64_t T0 = tshape[0], T1 = tshape[1], T2 = tshape[2];
  const int64_t nq_in = mode == 0 ? 2 : 1, nq_out = mode == 0 ? 3 : 1, B = max_batch;
  int64_t bytes = 0;
  bytes += 4 * B * T0 * T1 * T2 * (mode == 0 ? 2 : 1);
  if (with_filter) bytes += 8 * B * T0 * T1 * (T2 / 2 + 1);
  bytes += 8 * nq_in * B * T0 * T1 * N2s;
  bytes += 8 * nq_in * B * T0 * N1 * N2s;
  int64_t N2m = 0, ntl = 0;
  row_layout(N0, N1, fast_shape[2], N2m, ntl);
  bytes += 8 * nq_out * B * N0 * N1 * (N2m + ntl);
  if (N2s != N2c) bytes += 8 * (mode == 0 ? 2 : 1) * N0 * N1 * N2s;  // padded target spectra
  return bytes;
}

// Create a fused scan plan.  mode 0 = FLC (flc_setup/flc_scoring), mode 1 =
// correlation family (CC/LCC/FLCSphericalMask: corr_scoring).  Device
// pointers stay owned by the caller and must outlive the plan.
int tm_scan_plan_create(int mode, const int64_t* fast_shape, const int* tshape, const float* templ,
                        const float* mask, const float* filter, const void* ft_target, const void* ft_target2,
                        const float* inv_denominator, int normalize, double mask_sum, int max_batch,
                        tm_scan_plan** out) {
  TM_REQUIRE(out && fast_shape && tshape && templ && ft_target, "null argument to tm_scan_plan_create");
  TM_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (FLC) or 1 (correlation)");
  TM_REQUIRE(mode == 1 || (mask && ft_target2), "FLC needs a mask and the squared-target spectrum");
  TM_REQUIRE(!normalize || mask, "normalisation needs a mask");
  TM_REQUIRE(max_batch >= 1 && max_batch <= 4096, "max_batch %d outside [1, 4096]", max_batch);
  for (int i = 0; i < 3; ++i)
    TM_REQUIRE(tshape[i] >= 1 && tshape[i] <= fast_shape[i], "template axis %d (%d) exceeds the FFT shape", i,
               tshape[i]);
  TM_REQUIRE(fast_shape[0] <= 4096 && fast_shape[1] <= 4096 && fast_shape[2] <= 4096, "FFT axis longer than 4096");
  tm_scan_plan* p = (tm_scan_plan*)calloc(1, sizeof(tm_scan_plan));
  TM_REQUIRE(p, "host allocation failed");
  p->mode = mode;
  p->nq_in = mode == 0 ? 2 : 1;
  p->nq_out = mode == 0 ? 3 : 1;
  for (int i = 0; i < 3; ++i) {
    p->N[i] = (int)fast_shape[i];
    p->T[i] = tshape[i];
  }
  p->N2c = p->N[2] / 2 + 1;
  p->N2s = (int)pad_bins(p->N2c);
  {
    int64_t n2m = 0, ntl = 0;
    row_layout(p->N[0], p->N[1], p->N[2], n2m, ntl);
    p->N2m = (int)n2m;
    p->ntl = (int)ntl;
    // x length 200 (rows of 4 * 25 + 1 bins) with a y kernel of 16-line
    // tiles: radix-4 units (pass B also runs the first stage of pass C's x
    // inverse); otherwise mirror pairs
    const bool r4 = TM_R4_UNITS && ntl && p->N[2] == 200 && (p->N[1] == 128 || p->N[1] == 256 || p->N[1] == 512);
    p->pair_h = ntl ? (r4 ? 1 : tm_c2c_pair_h(p->N[1])) : 0;
    if (ntl && !p->pair_h) {
      tm_set_error("split rows without a mirror-pair kernel for y length %d", p->N[1]);
      free(p);
      return TM_ERR_VALUE;
    }
  }
  p->T2c = p->T[2] / 2 + 1;
  p->max_batch = max_batch;
  p->normalize = mode == 0 ? 1 : normalize;
  p->mask_sum = mask_sum;
  p->templ = templ;
  p->mask = mask;
  p->filter = filter;
  p->ft = (const float2*)ft_target;
  p->ft2 = (const float2*)ft_target2;
  p->inv_den = inv_denominator;
  int rc = tm_fft_plan_create(3, fast_shape, 1, &p->fast);
  if (rc != TM_OK) {
    scan_plan_free(p);
    return rc;
  }
  if (filter) {
    const int64_t ts[3] = {tshape[0], tshape[1], tshape[2]};
    rc = tm_fft_plan_create(3, ts, 1, &p->tplan);
    if (rc != TM_OK) {
      scan_plan_free(p);
      return rc;
    }
  }
  const int64_t B = max_batch, T3 = (int64_t)p->T[0] * p->T[1] * p->T[2];
  const int64_t N0 = p->N[0], N1 = p->N[1], N2s = p->N2s, T0 = p->T[0], T1 = p->T[1];
  TM_ALLOC(p->TR, sizeof(float) * B * T3);
  if (mode == 0) TM_ALLOC(p->MR, sizeof(float) * B * T3);
  if (filter) TM_ALLOC(p->specT, sizeof(float2) * B * T0 * T1 * p->T2c);
  TM_ALLOC(p->Q1, sizeof(float2) * p->nq_in * B * T0 * T1 * N2s);
  TM_ALLOC(p->P, sizeof(float2) * p->nq_in * B * T0 * N1 * N2s);
  TM_ALLOC(p->Q, sizeof(float2) * p->nq_out * B * N0 * N1 * p->N2m);
  if (p->ntl) TM_ALLOC(p->Qt, sizeof(float2) * p->nq_out * B * N0 * N1 * p->ntl);
  if (p->N2s != p->N2c) {
    TM_ALLOC(p->ftp, sizeof(float2) * N0 * N1 * N2s);
    if (mode == 0) TM_ALLOC(p->ft2p, sizeof(float2) * N0 * N1 * N2s);
  }
  TM_ALLOC(p->n_obs, sizeof(double) * B);
  TM_ALLOC(p->rot_ws, sizeof(double) * (tm_rotate_workspace(p->T, B) + 8));
  TM_ALLOC(p->mom_ws, sizeof(double) * 2 * TM_MOM_BLOCKS * B);
  TM_ALLOC(p->ms, sizeof(double) * 2 * B);
  // the spectra may still be in flight on the caller's stream
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) {
    rc = scan_copy_target(p, ft_target, ft_target2, nullptr);
    if (rc == TM_OK) e = hipStreamSynchronize(nullptr);
  }
  if (e != hipSuccess || rc != TM_OK) {
    if (e != hipSuccess) tm_set_error("copying the target spectra: %s", hipGetErrorString(e));
    scan_plan_free(p);
    return rc != TM_OK ? rc : TM_ERR_HIP;
  }
  *out = p;
  return TM_OK;
}

// Refresh the plan's copy of the target spectra (same shapes as at
// creation), ordered on `stream`.  Needed whenever the caller's spectra
// change while the plan is kept.
int tm_scan_plan_set_target(tm_scan_plan* p, const void* ft_target, const void* ft_target2, void* stream) {
  TM_REQUIRE(p && ft_target, "null argument to tm_scan_plan_set_target");
  TM_REQUIRE(p->mode == 1 || ft_target2, "FLC needs the squared-target spectrum");
  return scan_copy_target(p, ft_target, ft_target2, (hipStream_t)stream);
}

int tm_scan_plan_destroy(tm_scan_plan* p) {
  scan_plan_free(p);
  return TM_OK;
}

// Score rotations rots_dev[0..nbatch) (float32 3x3 each, device) with
// analyzer indices rot_index_dev[0..nbatch) (device int32) and fold them into
// max_scores / max_rot (fast_shape, device) with the strict '>' rule.
static int scan_batch_launch(tm_scan_plan* p, const float* rots_dev, const int* rot_index_dev, int nbatch,
                             float* max_scores, int* max_rot, hipStream_t s);

// Replay the batch's kernel sequence as a hipGraph on the plan's stream,
// ordered after and before the caller's stream by events.  The graph reads
// the rotations and indices from plan buffers, so only the state arrays and
// the batch size key it; the first batch of a configuration runs eagerly
// (it also settles every cached launch attribute), the second captures.
// Forget every captured batch sequence (the kernel sequence changed).
static int scan_graph_reset(tm_scan_plan* p) {
  for (auto& g : p->gslot) {
    if (g.exec) TM_CHECK_HIP(hipGraphExecDestroy(g.exec));
    g.exec = nullptr;
    g.warm = 0;
    g.used = 0;
  }
  return TM_OK;
}

static int scan_batch_graph(tm_scan_plan* p, const float* rots_dev, const int* rot_index_dev, int nbatch,
                            float* max_scores, int* max_rot, hipStream_t s) {
  TM_CHECK_HIP(hipEventRecord(p->ev_in, s));
  TM_CHECK_HIP(hipStreamWaitEvent(p->gs, p->ev_in, 0));
  TM_CHECK_HIP(hipMemcpyAsync(p->rbuf, rots_dev, sizeof(float) * 9 * nbatch, hipMemcpyDeviceToDevice, p->gs));
  TM_CHECK_HIP(hipMemcpyAsync(p->ibuf, rot_index_dev, sizeof(int) * nbatch, hipMemcpyDeviceToDevice, p->gs));
  tm_scan_plan::GraphSlot* slot = nullptr;
  for (auto& g : p->gslot)
    if (g.warm && g.nb == nbatch && g.ms == max_scores && g.mr == max_rot) slot = &g;
  int rc = TM_OK;
  if (!slot) {
    // a new configuration: replace the least recently used slot, run eagerly
    slot = p->gslot[0].used <= p->gslot[1].used ? &p->gslot[0] : &p->gslot[1];
    if (slot->exec) {
      TM_CHECK_HIP(hipGraphExecDestroy(slot->exec));
      slot->exec = nullptr;
    }
    slot->warm = 1;
    slot->nb = nbatch;
    slot->ms = max_scores;
    slot->mr = max_rot;
    ++p->g_eager;
    rc = scan_batch_launch(p, p->rbuf, p->ibuf, nbatch, max_scores, max_rot, p->gs);
  } else if (!slot->exec) {
    hipGraph_t g = nullptr;
    TM_CHECK_HIP(hipStreamBeginCapture(p->gs, hipStreamCaptureModeThreadLocal));
    rc = scan_batch_launch(p, p->rbuf, p->ibuf, nbatch, max_scores, max_rot, p->gs);
    const hipError_t e = hipStreamEndCapture(p->gs, &g);
    if (rc == TM_OK && e != hipSuccess) {
      tm_set_error("capturing the batch graph: %s", hipGetErrorString(e));
      rc = TM_ERR_HIP;
    }
    if (rc == TM_OK) {
      const hipError_t ie = hipGraphInstantiate(&slot->exec, g, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        tm_set_error("instantiating the batch graph: %s", hipGetErrorString(ie));
        slot->exec = nullptr;
        rc = TM_ERR_HIP;
      }
    }
    if (g) hipGraphDestroy(g);
    if (rc == TM_OK) rc = hipGraphLaunch(slot->exec, p->gs) == hipSuccess ? TM_OK : TM_ERR_HIP;
    ++p->g_captures;
  } else {
    TM_CHECK_HIP(hipGraphLaunch(slot->exec, p->gs));
    ++p->g_replays;
  }
  slot->used = ++p->g_clock;
  if (rc != TM_OK) return rc;
  TM_CHECK_HIP(hipEventRecord(p->ev_out, p->gs));
  TM_CHECK_HIP(hipStreamWaitEvent(s, p->ev_out, 0));
  return TM_OK;
}

int tm_scan_batch(tm_scan_plan* p, const float* rots_dev, const int* rot_index_dev, int nbatch, float* max_scores,
                  int* max_rot, void* stream) {
  TM_REQUIRE(p && rots_dev && rot_index_dev && max_scores && max_rot, "null argument to tm_scan_batch");
  TM_REQUIRE(nbatch >= 1 && nbatch <= p->max_batch, "batch %d outside [1, %d]", nbatch, p->max_batch);
  // per-kernel event timing records events inside the sequence: run eagerly
  if (p->use_graph && !tm_timing_on())
    return scan_batch_graph(p, rots_dev, rot_index_dev, nbatch, max_scores, max_rot, (hipStream_t)stream);
  return scan_batch_launch(p, rots_dev, rot_index_dev, nbatch, max_scores, max_rot, (hipStream_t)stream);
}

// Enable (1) or disable (0) the graph replay of the batch sequence.
int tm_scan_plan_use_graph(tm_scan_plan* p, int enable) {
  TM_REQUIRE(p, "null plan");
  if (enable && !p->gs) {
    TM_CHECK_HIP(hipStreamCreateWithFlags(&p->gs, hipStreamNonBlocking));
    TM_CHECK_HIP(hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming));
    TM_CHECK_HIP(hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming));
    TM_CHECK_HIP(hipMalloc((void**)&p->rbuf, sizeof(float) * 9 * p->max_batch));
    TM_CHECK_HIP(hipMalloc((void**)&p->ibuf, sizeof(int) * p->max_batch));
  }
  p->use_graph = enable ? 1 : 0;
  return TM_OK;
}

// --------------------------------------------------------------------------
// plan groups: one batch of several plans (templates of uneven size against
// one target) per call, replayed as one graph

struct tm_scan_group {
  int n;
  tm_scan_plan** plans;  // not owned
  hipStream_t gs;
  hipEvent_t ev_in, ev_out;
  hipGraphExec_t gexec;
  int warm;
  // key of the warm / captured configuration, per plan: batch size (0 =
  // no batch this call), state arrays and tile shapes
  int* k_nb;
  float** k_ms;
  int** k_mr;
  int* k_tiles;
};

int tm_scan_group_destroy(tm_scan_group* g);

int tm_scan_group_create(tm_scan_plan* const* plans, int n, tm_scan_group** out) {
  TM_REQUIRE(plans && out && n >= 1 && n <= 64, "plan group: 1 to 64 plans, got %d", n);
  for (int i = 0; i < n; ++i) {
    TM_REQUIRE(plans[i], "plan group: null plan %d", i);
    for (int j = 0; j < i; ++j) TM_REQUIRE(plans[j] != plans[i], "plan group: plan %d repeated", i);
  }
  tm_scan_group* g = (tm_scan_group*)calloc(1, sizeof(tm_scan_group));
  if (!g) {
    tm_set_error("plan group: out of host memory");
    return TM_ERR_OOM;
  }
  g->n = n;
  g->plans = (tm_scan_plan**)calloc(n, sizeof(tm_scan_plan*));
  g->k_nb = (int*)calloc(n, sizeof(int));
  g->k_ms = (float**)calloc(n, sizeof(float*));
  g->k_mr = (int**)calloc(n, sizeof(int*));
  g->k_tiles = (int*)calloc(3 * n, sizeof(int));
  if (!g->plans || !g->k_nb || !g->k_ms || !g->k_mr || !g->k_tiles) {
    tm_scan_group_destroy(g);
    tm_set_error("plan group: out of host memory");
    return TM_ERR_OOM;
  }
  for (int i = 0; i < n; ++i) g->plans[i] = plans[i];
  int rc = TM_OK;
  // the graph reads each plan's rotation buffers (allocated with its own graph state)
  for (int i = 0; i < n && rc == TM_OK; ++i) {
    const int was = plans[i]->use_graph;
    rc = tm_scan_plan_use_graph(plans[i], 1);
    plans[i]->use_graph = was;
  }
  if (rc == TM_OK && hipStreamCreateWithFlags(&g->gs, hipStreamNonBlocking) != hipSuccess) rc = TM_ERR_HIP;
  if (rc == TM_OK && hipEventCreateWithFlags(&g->ev_in, hipEventDisableTiming) != hipSuccess) rc = TM_ERR_HIP;
  if (rc == TM_OK && hipEventCreateWithFlags(&g->ev_out, hipEventDisableTiming) != hipSuccess) rc = TM_ERR_HIP;
  if (rc != TM_OK) {
    if (rc == TM_ERR_HIP) tm_set_error("plan group: creating its stream and events failed");
    tm_scan_group_destroy(g);
    return rc;
  }
  *out = g;
  return TM_OK;
}

int tm_scan_group_destroy(tm_scan_group* g) {
  if (!g) return TM_OK;
  if (g->gs) hipStreamSynchronize(g->gs);
  if (g->gexec) hipGraphExecDestroy(g->gexec);
  if (g->ev_in) hipEventDestroy(g->ev_in);
  if (g->ev_out) hipEventDestroy(g->ev_out);
  if (g->gs) hipStreamDestroy(g->gs);
  free(g->plans);
  free(g->k_nb);
  free(g->k_ms);
  free(g->k_mr);
  free(g->k_tiles);
  free(g);
  return TM_OK;
}

// One batch of every plan with nbatch[i] > 0, in plan order, into that
// plan's state arrays.  The first call of a configuration (batch sizes,
// state arrays, tile shapes) runs eagerly on the group stream, the second
// captures the whole sequence into one graph, later ones replay it; the
// group stream is event-ordered after and before the caller's stream.
// Per-kernel timing forces eager launches on the caller's stream.
int tm_scan_group_batch(tm_scan_group* g, const float* const* rots_dev, const int* const* rot_index_dev,
                        const int* nbatch, float* const* max_scores, int* const* max_rot, void* stream) {
  TM_REQUIRE(g && rots_dev && rot_index_dev && nbatch && max_scores && max_rot, "null argument to tm_scan_group_batch");
  int active = 0;
  for (int i = 0; i < g->n; ++i) {
    TM_REQUIRE(nbatch[i] >= 0 && nbatch[i] <= g->plans[i]->max_batch, "plan %d: batch %d outside [0, %d]", i,
               nbatch[i], g->plans[i]->max_batch);
    if (nbatch[i] == 0) continue;
    TM_REQUIRE(rots_dev[i] && rot_index_dev[i] && max_scores[i] && max_rot[i], "plan %d: null batch argument", i);
    ++active;
  }
  if (!active) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  if (tm_timing_on()) {
    for (int i = 0; i < g->n; ++i) {
      if (!nbatch[i]) continue;
      const int rc = scan_batch_launch(g->plans[i], rots_dev[i], rot_index_dev[i], nbatch[i], max_scores[i],
                                       max_rot[i], s);
      if (rc != TM_OK) return rc;
    }
    return TM_OK;
  }
  TM_CHECK_HIP(hipEventRecord(g->ev_in, s));
  TM_CHECK_HIP(hipStreamWaitEvent(g->gs, g->ev_in, 0));
  bool same = true;
  for (int i = 0; i < g->n; ++i) {
    tm_scan_plan* p = g->plans[i];
    if (nbatch[i]) {
      TM_CHECK_HIP(hipMemcpyAsync(p->rbuf, rots_dev[i], sizeof(float) * 9 * nbatch[i], hipMemcpyDeviceToDevice,
                                  g->gs));
      TM_CHECK_HIP(hipMemcpyAsync(p->ibuf, rot_index_dev[i], sizeof(int) * nbatch[i], hipMemcpyDeviceToDevice,
                                  g->gs));
    }
    float* ms = nbatch[i] ? max_scores[i] : nullptr;
    int* mr = nbatch[i] ? max_rot[i] : nullptr;
    if (g->k_nb[i] != nbatch[i] || g->k_ms[i] != ms || g->k_mr[i] != mr || g->k_tiles[3 * i] != p->a_cols ||
        g->k_tiles[3 * i + 1] != p->c_rows || g->k_tiles[3 * i + 2] != p->radix)
      same = false;
    g->k_nb[i] = nbatch[i];
    g->k_ms[i] = ms;
    g->k_mr[i] = mr;
    g->k_tiles[3 * i] = p->a_cols;
    g->k_tiles[3 * i + 1] = p->c_rows;
    g->k_tiles[3 * i + 2] = p->radix;
  }
  auto launch_all = [&]() -> int {
    for (int i = 0; i < g->n; ++i) {
      if (!nbatch[i]) continue;
      tm_scan_plan* p = g->plans[i];
      const int rc = scan_batch_launch(p, p->rbuf, p->ibuf, nbatch[i], max_scores[i], max_rot[i], g->gs);
      if (rc != TM_OK) return rc;
    }
    return TM_OK;
  };
  int rc = TM_OK;
  if (!same || !g->warm) {
    // a new configuration: drop the old graph and run eagerly once
    if (g->gexec) {
      TM_CHECK_HIP(hipGraphExecDestroy(g->gexec));
      g->gexec = nullptr;
    }
    g->warm = 1;
    rc = launch_all();
  } else if (!g->gexec) {
    hipGraph_t graph = nullptr;
    TM_CHECK_HIP(hipStreamBeginCapture(g->gs, hipStreamCaptureModeThreadLocal));
    rc = launch_all();
    const hipError_t e = hipStreamEndCapture(g->gs, &graph);
    if (rc == TM_OK && e != hipSuccess) {
      tm_set_error("capturing the group graph: %s", hipGetErrorString(e));
      rc = TM_ERR_HIP;
    }
    if (rc == TM_OK) {
      const hipError_t ie = hipGraphInstantiate(&g->gexec, graph, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        tm_set_error("instantiating the group graph: %s", hipGetErrorString(ie));
        g->gexec = nullptr;
        rc = TM_ERR_HIP;
      }
    }
    if (graph) hipGraphDestroy(graph);
    if (rc == TM_OK && hipGraphLaunch(g->gexec, g->gs) != hipSuccess) {
      tm_set_error("launching the group graph failed");
      rc = TM_ERR_HIP;
    }
  } else if (hipGraphLaunch(g->gexec, g->gs) != hipSuccess) {
    tm_set_error("launching the group graph failed");
    rc = TM_ERR_HIP;
  }
  if (rc != TM_OK) {
    g->warm = 0;  // start over with an eager run on the next call
    return rc;
  }
  TM_CHECK_HIP(hipEventRecord(g->ev_out, g->gs));
  TM_CHECK_HIP(hipStreamWaitEvent(s, g->ev_out, 0));
  return TM_OK;
}

// Tile shapes of the scan's passes (chosen by the batch autotuner):
// pass_a_columns 16 or 8 (0: 16 where the padded rows allow it), pass_c_rows
// 5 or 10 (0: 5; other than 5 only for an x length of 200).  A change drops
// the captured batch graph.  Results do not depend on the tiles.
int tm_scan_plan_set_tiles(tm_scan_plan* p, int pass_a_columns, int pass_c_rows) {
  TM_REQUIRE(p, "null plan");
  TM_REQUIRE(pass_a_columns == 0 || pass_a_columns == 8 || pass_a_columns == 16,
             "pass A column tiles are 8 or 16, got %d", pass_a_columns);
  TM_REQUIRE(pass_a_columns != 16 || tm_ceil_div(p->N2c, 16) * 16 <= p->N2s,
             "16-column pass A tiles need rows padded to 16 bins");
  TM_REQUIRE(pass_c_rows == 0 || pass_c_rows == 5 || (pass_c_rows == 10 && p->N[2] == 200),
             "pass C takes 5 rows per block (10 for an x length of 200), got %d", pass_c_rows);
  if (p->a_cols != pass_a_columns || p->c_rows != pass_c_rows) {
    p->a_cols = pass_a_columns;
    p->c_rows = pass_c_rows;
    const int rc = scan_graph_reset(p);
    if (rc != TM_OK) return rc;
  }
  return TM_OK;
}

// Counts of the plan's graph-mode batches: captured, replayed and run
// eagerly (the first occurrence of a configuration).
int tm_scan_plan_graph_stats(const tm_scan_plan* p, int64_t* captures, int64_t* replays, int64_t* eager) {
  TM_REQUIRE(p && captures && replays && eager, "null argument");
  *captures = p->g_captures;
  *replays = p->g_replays;
  *eager = p->g_eager;
  return TM_OK;
}

// Radix plan variant (0 or 1, fft_ct.h ct_factor_v) of the compile-time z,
// y and x passes of the batch.
int tm_scan_plan_set_radix(tm_scan_plan* p, int variant) {
  TM_REQUIRE(p, "null plan");
  TM_REQUIRE(variant == 0 || variant == 1, "radix plan variant %d is not 0 or 1", variant);
  if (p->radix != variant) {
    p->radix = variant;
    const int rc = scan_graph_reset(p);
    if (rc != TM_OK) return rc;
  }
  return TM_OK;
}

// Radices (in stage order) of the forward compile-time plan of length n in
// a variant; *count = 0 when n has no compile-time plan.
int tm_fft_radix_plan(int n, int variant, int* radices, int max_out, int* count) {
  TM_REQUIRE(radices && count && max_out >= 1, "null argument");
  TM_REQUIRE(variant == 0 || variant == 1, "radix plan variant %d is not 0 or 1", variant);
  *count = 0;
  if (n < 2 || n > 4096) return TM_OK;
  int prod = 1, k = 0;
  while (k < max_out) {
    const int r = ct_factor_v(n, k, variant);
    if (!r) break;
    radices[k++] = r;
    prod *= r;
  }
  if (prod == n) *count = k;
  return TM_OK;
}

}  // extern "C"

static int scan_batch_launch(tm_scan_plan* p, const float* rots_dev, const int* rot_index_dev, int nbatch,
                             float* max_scores, int* max_rot, hipStream_t s) {
  const int B = nbatch;
  const int T0 = p->T[0], T1 = p->T[1], T2 = p->T[2];
  const int64_t T3 = (int64_t)T0 * T1 * T2;
  const int N0 = p->N[0], N1 = p->N[1], N2 = p->N[2], N2c = p->N2c, N2s = p->N2s;
  const bool flc = p->mode == 0;
  int rc;

  // 1. rotate template (and mask) about the box centre
  rc = tm_rotate_launch(p->templ, flc ? p->mask : nullptr, p->T, rots_dev, B, p->TR, flc ? p->MR : nullptr,
                        (int64_t)T1 * T2, T2, T3, flc ? p->n_obs : nullptr, flc ? p->rot_ws : nullptr, s);
  if (rc != TM_OK) return rc;

  // 2. template-size Fourier filter
  if (p->filter) {
    rc = tm_fft_r2c_batch(p->tplan, B, p->TR, p->specT, s);
    if (rc != TM_OK) return rc;
    const int64_t nf = (int64_t)T0 * T1 * p->T2c;
    rc = tm_cmul_cr(p->specT, p->filter, p->specT, nf * B, nf, (void*)s);
    if (rc != TM_OK) return rc;
    rc = tm_fft_c2r_batch(p->tplan, B, p->specT, p->TR, s);
    if (rc != TM_OK) return rc;
  }

  // 3. normalise inside the (rotated / fixed) mask with FP64 moments
  if (p->normalize) {
    const float* m = flc ? p->MR : p->mask;
    const int64_t mstride = flc ? T3 : 0;
    TM_TIMED("fused_normalize", 12.0 * T3 * B, s,
             hipLaunchKernelGGL(k_moments_batched, dim3(TM_MOM_BLOCKS, B), dim3(TM_FU_THREADS), 0, s, p->TR, m, T3,
                                mstride, p->mom_ws));
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_moments_finalize, dim3((unsigned)tm_ceil_div(B, 64)), dim3(64), 0, s, p->mom_ws,
                       TM_MOM_BLOCKS, flc ? p->n_obs : (const double*)nullptr, p->mask_sum, B, p->ms);
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_normalize_batched, dim3((unsigned)tm_ceil_div(T3, 1024), B), dim3(TM_FU_THREADS), 0, s,
                       p->TR, m, T3, mstride, (const double*)p->ms);
    TM_CHECK_LAUNCH();
  }

  // 4. x pass (R2C of the zero-padded template rows) and y pass (T1 -> N1)
  const LinePlan& lx = p->fast->lines[2].lp;
  const LinePlan& ly = p->fast->lines[1].lp;
  const LinePlan& lz = p->fast->lines[0].lp;
  const int64_t q1_stride = (int64_t)B * T0 * T1 * N2s;
  const int pair_nm = p->pair_h ? N2c - p->ntl : 0;
  if (flc) {
    rc = tm_launch_r2c_rows_ct(lx, p->MR, T2, p->Q1, (int64_t)B * T0 * T1, N2s, s, "fused_x_pass", pair_nm,
                               p->pair_h, p->radix);
    if (rc != TM_OK) return rc;
  }
  rc = tm_launch_r2c_rows_ct(lx, p->TR, T2, p->Q1 + (flc ? q1_stride : 0), (int64_t)B * T0 * T1, N2s, s,
                             "fused_x_pass", pair_nm, p->pair_h, p->radix);
  if (rc != TM_OK) return rc;
  rc = tm_launch_c2c(ly, p->Q1, T1, p->P, (int64_t)p->nq_in * B * T0, N2c, N2s, -1, 1.0f, s, "fused_y_pass",
                     p->radix);
  if (rc != TM_OK) return rc;

  // 5. pass A: z-axis, spectral products
  bool done_a = true;
  // 16-bin column tiles when they fit the padded rows, else 8-bin tiles
  const bool l16 = tm_ceil_div(N2c, 16) * 16 <= N2s && p->a_cols != 8;
  auto pass_a = [&](auto vtag) -> int {
    constexpr int V = decltype(vtag)::value;
#define TM_PASS_A(N0_, L_, ...) \
  launch_pass_a_ct<V, N0_, L_, __VA_ARGS__>(flc, p->P, p->ft, p->ft2, p->Q, p->Qt, lz.tw, T0, N1, N2c, N2s, p->N2m, \
                                            p->ntl, B, s)
    switch (N0) {
      case 128: return l16 ? TM_PASS_A(128, 16, 256) : TM_PASS_A(128, 8, 128);
      case 256: return l16 ? TM_PASS_A(256, 16, 512) : TM_PASS_A(256, 8, 256);
      case 512: return l16 ? TM_PASS_A(512, 16, 1024, 1024) : TM_PASS_A(512, 8, 512);
      case 1024: return TM_PASS_A(1024, 8, 1024, 512);
#undef TM_PASS_A
      default: done_a = false; return TM_OK;
    }
  };
  rc = p->radix ? pass_a(std::integral_constant<int, 1>{}) : pass_a(std::integral_constant<int, 0>{});
  if (done_a && rc != TM_OK) return rc;
  if (!done_a) {
    int TI = 4096 / N0;
    if (TI > 16) TI = 16;
    if (TI > N2c) TI = N2c;
    TM_REQUIRE(TI >= 1 && (int64_t)N0 * TI <= TM_FU_MAXP * TM_FU_THREADS, "z axis too long for pass A");
    const int tpr = (int)tm_ceil_div(N2c, TI);
    const int64_t blocks = (int64_t)N1 * tpr;
    const size_t shm = 2 * (size_t)N0 * TI * sizeof(float2);
    // target spectra once per launch + pruned template columns + products
    const double C = (double)N0 * N1 * N2c;
    const double bytes = 8.0 * C * (flc ? 2 : 1) + 8.0 * p->nq_in * B * (double)T0 * N1 * N2c +
                         8.0 * C * p->nq_out * B;
    if (flc)
      TM_TIMED("fused_pass_a", bytes, s,
               hipLaunchKernelGGL(k_pass_a<3>, dim3((unsigned)blocks), dim3(TM_FU_THREADS), shm, s, p->P, p->ft,
                                  p->ft2, p->Q, lz, T0, N1, N2c, N2s, B, TI, tpr));
    else
      TM_TIMED("fused_pass_a", bytes, s,
               hipLaunchKernelGGL(k_pass_a<1>, dim3((unsigned)blocks), dim3(TM_FU_THREADS), shm, s, p->P, p->ft,
                                  (const float2*)nullptr, p->Q, lz, T0, N1, N2c, N2s, B, TI, tpr));
    TM_CHECK_LAUNCH();
  }

  // 6. pass B: inverse y in place
  if (p->ntl) {
    // split rows: the main rows (whole 16-bin tiles of mirror pairs, stored
    // as pass C's x-inverse inputs), then the tail array (lines (z, c) of
    // N0 * ntl, whole 128-byte segments; pass C pairs the tail bins itself)
    rc = tm_launch_c2c_pairs(ly, lx.tw, N2 / 2, p->pair_h, p->Q, (int64_t)p->nq_out * B * N0, N2c - p->ntl, p->N2m,
                             s, "fused_pass_b", p->radix);
    if (rc != TM_OK) return rc;
    rc = tm_launch_c2c(ly, p->Qt, N1, p->Qt, (int64_t)p->nq_out * B, (int64_t)N0 * p->ntl, (int64_t)N0 * p->ntl, 1,
                       1.0f, s, "fused_pass_b_tail", p->radix);
  } else {
    rc = tm_launch_c2c(ly, p->Q, N1, p->Q, (int64_t)p->nq_out * B * N0, N2c, N2s, 1, 1.0f, s, "fused_pass_b",
                       p->radix);
  }
  if (rc != TM_OK) return rc;

  // 7. pass C: inverse x, score, max/argmax
  bool done_c = false;
  {
    const int64_t nrows = (int64_t)N0 * N1;
    const double F = (double)nrows * N2;
    const double bytes = 8.0 * (double)nrows * N2c * p->nq_out * B + 16.0 * F + (p->inv_den ? 4.0 * F * B : 0.0);
    const float scale = (float)(1.0 / F);
    const float eps = 1.1920928955078125e-07f;
    done_c = true;
    auto pass_c = [&](auto vtag) -> int {
    constexpr int V = decltype(vtag)::value;
    switch (N2) {
      case 128: return launch_pass_c_ct<V, 64, 32, 256, 1, 2>(flc, p->Q, nullptr, lx.tw, nrows, N2s, N0, N1, 0, 0, B, p->n_obs, p->inv_den, rot_index_dev, scale, eps, max_scores, max_rot, bytes, s);
      case 200:
        if (p->ntl) {  // split rows: main bins in Q (row stride N2m), the tail bins in Qt
          if (p->c_rows == 10)
            return launch_pass_c_ct<V, 100, 10, 256, 3, 2, true>(flc, p->Q, p->Qt, lx.tw, nrows, p->N2m, N0, N1, p->ntl, p->pair_h, B, p->n_obs, p->inv_den, rot_index_dev, scale, eps, max_scores, max_rot, bytes, s);
          else
            return launch_pass_c_ct<V, 100, 5, 128, 3, 2, true>(flc, p->Q, p->Qt, lx.tw, nrows, p->N2m, N0, N1, p->ntl, p->pair_h, B, p->n_obs, p->inv_den, rot_index_dev, scale, eps, max_scores, max_rot, bytes, s);
        } else if (p->c_rows == 10) {
          return launch_pass_c_ct<V, 100, 10, 256, 3, 2>(flc, p->Q, nullptr, lx.tw, nrows, N2s, N0, N1, 0, 0, B, p->n_obs, p->inv_den, rot_index_dev, scale, eps, max_scores, max_rot, bytes, s);
        } else {
          return launch_pass_c_ct<V, 100, 5, 128, 3, 2>(flc, p->Q, nullptr, lx.tw, nrows, N2s, N0, N1, 0, 0, B, p->n_obs, p->inv_den, rot_index_dev, scale, eps, max_scores, max_rot, bytes, s);
        }
      case 256:
        if (p->ntl)  // split rows: 128 bins in Q, the Nyquist bin in Qt
          return launch_pass_c_ct<V, 128, 16, 256, 1, 2, true>(flc, p->Q, p->Qt, lx.tw, nrows, p->N2m, N0, N1, p->ntl, p->pair_h, B, p->n_obs, p->inv_den, rot_index_dev, scale, eps, max_scores, max_rot, bytes, s);
        else
          return launch_pass_c_ct<V, 128, 16, 256, 1, 2>(flc, p->Q, nullptr, lx.tw, nrows, N2s, N0, N1, 0, 0, B, p->n_obs, p->inv_den, rot_index_dev, scale, eps, max_scores, max_rot, bytes, s);
      default: done_c = false;
    }
    return TM_OK;
    };
    rc = p->radix ? pass_c(std::integral_constant<int, 1>{}) : pass_c(std::integral_constant<int, 0>{});
    if (done_c && rc != TM_OK) return rc;
  }
  if (!done_c) {
    int RL = 4096 / N2;
    if (RL > 16) RL = 16;
    const int64_t nrows = (int64_t)N0 * N1;
    if (nrows < RL) RL = (int)nrows;
    TM_REQUIRE(RL >= 1 && RL * N2 <= TM_FU_MAXP * TM_FU_THREADS, "x axis too long for pass C");
    const size_t shm = 2 * (size_t)RL * (N2 + 1) * sizeof(float2);
    const int64_t blocks = tm_ceil_div(nrows, RL);
    const double F = (double)nrows * N2;
    const double bytes = 8.0 * (double)nrows * N2c * p->nq_out * B + 16.0 * F + (p->inv_den ? 4.0 * F * B : 0.0);
    const float scale = (float)(1.0 / F);
    const float eps = 1.1920928955078125e-07f;
    if (flc)
      TM_TIMED("fused_pass_c", bytes, s,
               hipLaunchKernelGGL(k_pass_c<3>, dim3((unsigned)blocks), dim3(TM_FU_THREADS), shm, s, p->Q, lx, nrows,
                                  (int64_t)N2s, RL, B, (const double*)p->n_obs, (const float*)nullptr, rot_index_dev, scale, eps,
                                  max_scores, max_rot));
    else
      TM_TIMED("fused_pass_c", bytes, s,
               hipLaunchKernelGGL(k_pass_c<1>, dim3((unsigned)blocks), dim3(TM_FU_THREADS), shm, s, p->Q, lx, nrows,
                                  (int64_t)N2s, RL, B, (const double*)nullptr, p->inv_den, rot_index_dev, scale, eps, max_scores,
                                  max_rot));
    TM_CHECK_LAUNCH();
  }
  return TM_OK;
}
