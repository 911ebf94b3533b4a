// Batched real<->complex n-D FFTs (n <= 3) for C-ordered float32 volumes.
//
// Forward (numpy.fft.rfftn semantics, unnormalised):
//   pass 0: R2C along the last (contiguous) axis, real rows -> n_last/2+1 bins
//   pass k: C2C along axis ndim-1-k of the half spectrum, in place
// Inverse (numpy.fft.irfftn semantics, 1/N normalisation folded into the
// final pass; the input half spectrum is used as scratch and overwritten):
//   C2C along axes 0..ndim-2 in place, then C2R along the last axis.
// Each pass is one kernel: a workgroup loads a tile of lines into LDS with
// coalesced global accesses, runs the whole mixed-radix transform in LDS
// (fft_core.h) and writes the tile back, so every pass streams the volume
// through HBM exactly once (read + write).
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "fft_core.h"
#include "fft_ct.h"
#include "fft_plan.h"
#include "tm_common.h"


// --------------------------------------------------------------------------
// kernels

// C2C along the middle axis: in is [outer][lin][ld], out is [outer][n][ld]
// and the first `inner` of the ld inner indices are transformed (lin <= n;
// the missing lin..n-1 inputs are zeros, the pruned transform of a
// zero-padded line).  in == out (lin == n) is the in-place case.  A block owns TI consecutive inner indices of one outer
// index; the tile is element-major in LDS (element e of line l at e*ti + l)
// so both the global accesses and the butterflies are conflict-free.
template <int DIR>
__global__ void __launch_bounds__(TM_FFT_THREADS)
k_c2c_strided(const float2* in, float2* out, LinePlan p, int lin, int64_t inner, int64_t ld, int TI,
              int64_t tiles_per_outer, float scale) {
  extern __shared__ float2 lds[];
  const int n = p.n;
  const int64_t o = blockIdx.x / tiles_per_outer;
  const int64_t i0 = (blockIdx.x - o * tiles_per_outer) * (int64_t)TI;
  const int ti = (int)((inner - i0) < TI ? (inner - i0) : TI);
  float2* A = lds;
  float2* B = lds + n * ti;
  const float2* src = in + o * (int64_t)lin * ld + i0;
  float2* dst = out + o * (int64_t)n * ld + i0;
  const int tid = threadIdx.x;
  for (int f = tid; f < n * ti; f += TM_FFT_THREADS) {
    const int e = f / ti, l = f - e * ti;
    A[f] = e < lin ? src[(int64_t)e * ld + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* R = lds_fft_lines<DIR, true>(A, B, ti, 1, ti, p, tid, TM_FFT_THREADS);
  for (int f = tid; f < n * ti; f += TM_FFT_THREADS) {
    const int e = f / ti, l = f - e * ti;
    dst[(int64_t)e * ld + l] = c_scale(R[f], scale);
  }
}

// Compile-time specialised variant of k_c2c_strided for the common lengths:
// the tile (16 lines of 128 B for the long axes) goes straight from global
// memory into registers, LDS only carries the inter-stage exchanges.
// FULL: lin == N and every tile lies inside the row stride ld (lines past
// `inner` are padding that may be transformed), so no load or store is
// predicated; the compiler can then wait for the prefetched tile without
// draining the previous tile's stores.
//
// PRU: pruned input (lin <= N / R_IN, the zero-padded template lines of
// the fused scan): only the first input of each first-stage butterfly can
// be non-zero, so only those are loaded (rows past lin clamped to a valid
// row and replaced by zero) and the others are compile-time zeros; stores
// as FULL (tiles inside ld).
//
// PH > 0 (inverse transforms of mirror-pair rows, tm_pair_bin with h = PH
// and tiles of L = 2 PH positions): before the store each output is replaced
// by the C2R input tm_c2r_pre of its bin k from the pair (X[k], X[M-k]),
// whose partner sits in lane l ^ PH; twx holds the 2M twiddles of the row
// transform.  The row pass then runs its M-point inverse directly.
//
// PH == 1 (radix-4 units of tm_pair_bin, L = 16): the pair partner sits in
// lane c ^ 7 of the 8-lane unit; after the pre-processing each lane quad
// (one first-stage group j + G r, r = 0..3, of the row's M-point inverse,
// G = M / 4) runs that radix-4 stage across its lanes, lane m keeping the
// output of index (m >> 1) | ((m & 1) << 1).  The row pass then starts at
// its second stage.
template <int PH>
__device__ __forceinline__ float tm_lane_xor(float v) {
  if constexpr (PH == 1)  // row_half_mirror: lane c <-> 7 - c within 8 lanes
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xf, 0xf, false));
  else if constexpr (PH == 8)  // row_ror:8 within 16 lanes is lane ^ 8
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128, 0xf, 0xf, false));
  else  // ds_swizzle bit mode: and 0x1f, or 0, xor PH
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1f | (PH << 10)));
}

// 1: the 512-point compile-time C2C passes over two LDS exchange buffers
#ifndef TM_C2C_TWO_BUFFERS
#define TM_C2C_TWO_BUFFERS 1
#endif
#ifndef TM_PB_TWIDDLE_TABLE
#define TM_PB_TWIDDLE_TABLE 1
#endif
template <int N, int L, int NT, int DIR, bool FULL, int V = 0, int PH = 0, bool PRU = false>
__global__ void __launch_bounds__(NT)
k_c2c_ct(const float2* in, float2* out, const float2* __restrict__ tw, int lin, int64_t inner, int ld,
         int64_t tiles_per_outer, int64_t ntiles, float scale, const float2* __restrict__ twx, int M) {
  // radix-4 units (PH 1) of the 512-point axis: twiddles from the table,
  // leaving VALU time to the longer epilogue (one block per CU either way;
  // 15.34 -> 15.15 ms per 64 rotations of 512x512x200)
  using PL = CtPlan<N, L, NT, true, false, V, PH == 1 && N == 512 && TM_PB_TWIDDLE_TABLE>;
  constexpr int EI = PL::E_IN, EO = PL::E_OUT, EW = EI > EO ? EI : EO;
  constexpr int BI = EI / PL::R_IN;  // first-stage butterflies per thread
  static_assert(!PRU || (FULL && PH == 0), "pruned lines are stored unpredicated");
  constexpr bool WIDE = L * sizeof(float2) >= 128;  // whole-line row segments: streaming hints
  static_assert(PH == 0 || (L == (PH == 1 ? 16 : 2 * PH) && NT % L == 0 && FULL && (PH == 8 || PH == 4 || PH == 1)),
                "mirror pairs need whole tiles of 2 PH lines (16 for radix-4 units) and unpredicated tiles");
  extern __shared__ float2 lds[];
  __shared__ float2 tws[N];  // twiddles in LDS: no vector-memory waits inside the transforms
  __shared__ float2 twxs[PH > 0 ? 256 : 1];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tws[i] = tw[i];
  if constexpr (PH > 0)
    for (int i = tid; i < 2 * M; i += NT) twxs[i] = twx[i];
  // persistent blocks: tile t = blockIdx.x + k * gridDim.x; the next tile's
  // loads are issued before the current tile's transform
  float2 v[EW], nx[PRU ? BI : EI];
  auto load = [&](int64_t t, float2* dstv) {
    const int64_t o = t / tiles_per_outer;
    const int64_t i0 = (t - o * tiles_per_outer) * (int64_t)L;
    const int ti = (int)((inner - i0) < L ? (inner - i0) : L);
    const float2* src = in + o * (int64_t)lin * ld + i0;
    if constexpr (PRU) {
#pragma unroll
      for (int i = 0; i < BI; ++i) {
        int l, j;
        PL::template slot<PL::R_IN>(tid, i, l, j);
        dstv[i] = tm_ld_stream<WIDE>(src + (j < lin ? j : lin - 1) * ld + l);
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < EI / PL::R_IN; ++i)
#pragma unroll
      for (int r = 0; r < PL::R_IN; ++r) {
        int l, e;
        const bool ok = PL::in_pos(tid, i, r, l, e);
        if (FULL)
          dstv[i * PL::R_IN + r] = tm_ld_stream<WIDE>(src + e * ld + l);
        else
          dstv[i * PL::R_IN + r] = (ok && l < ti && e < lin) ? src[e * ld + l] : make_float2(0.f, 0.f);
      }
  };
  // the blocks of one XCD take consecutive tiles (neighbouring column tiles
  // of the same rows share cache lines and DRAM pages in that XCD's L2)
  // PRU: the loaded first inputs x into the transform registers w
  auto expand = [&](const float2* x, float2* w) {
#pragma unroll
    for (int i = 0; i < BI; ++i) {
      int l, j;
      PL::template slot<PL::R_IN>(tid, i, l, j);
#pragma unroll
      for (int r = 0; r < PL::R_IN; ++r) w[i * PL::R_IN + r] = r == 0 && j < lin ? x[i] : make_float2(0.f, 0.f);
    }
  };
  int64_t t = tm_xcd_block();
  if (t >= ntiles) return;
  if constexpr (PRU) {
    load(t, nx);
    ct_ready<BI>(nx);
  } else {
    load(t, v);
    ct_ready<EI>(v);
  }
  for (; t < ntiles; t += gridDim.x) {
    if constexpr (PRU) expand(nx, v);
    // on the last tile the current one is reloaded and discarded: the loop
    // carries no predicated loads, so the wait for the prefetch at the loop
    // end does not also wait for this tile's stores
    load(t + gridDim.x < ntiles ? t + gridDim.x : t, nx);
    if constexpr (TM_C2C_TWO_BUFFERS && N == 512 && !PRU && !ct2_last_in_x<PL>()) {
      // two exchange buffers: one barrier per stage, none before the tile
      // (the previous tile's last stage read the second buffer)
      ct_fft2<PL, DIR, L>(lds, lds + N * L, tws, v, v, tid);
    } else {
      ct_fft<PL, DIR, L>(lds, tws, v, v, tid);
    }
    const int64_t o = t / tiles_per_outer;
    const int64_t i0 = (t - o * tiles_per_outer) * (int64_t)L;
    const int ti = (int)((inner - i0) < L ? (inner - i0) : L);
    float2* dst = out + o * (int64_t)N * ld + i0;
    if constexpr (PH == 1) {
      // every output slot of this thread is on line tid % L of tile jt:
      // position q = L jt + lt, unit lane c, quad lane m
      const int q = (int)i0 + tid % L, c = q & 7, m = c & 3, j = (q >> 3) + 1, G = M / 4;
      const int k = c < 4 ? j + G * c : (G - j) + G * (c - 4);
      const float2 w = tw_load<1>(twxs, k);
#pragma unroll
      for (int s2 = 0; s2 < EO; ++s2) {
        const float2 a = v[s2];
        v[s2] = tm_quad_idft4(tm_c2r_pre(a, make_float2(tm_lane_xor<1>(a.x), tm_lane_xor<1>(a.y)), w, false), m);
      }
    } else if constexpr (PH > 0) {
      // every output slot of this thread is on line tid % L of tile jt
      const int jt = (int)(i0 / L), lt = tid % L;
      const int k = lt < PH ? PH * jt + lt : M - PH * jt - (lt - PH);
      const float2 w = tw_load<1>(twxs, k);
#pragma unroll
      for (int s2 = 0; s2 < EO; ++s2) {
        const float2 a = v[s2];
        const float2 b = make_float2(tm_lane_xor<PH>(a.x), tm_lane_xor<PH>(a.y));
        v[s2] = tm_c2r_pre(a, b, w, k == 0);
      }
    }
#pragma unroll
    for (int i = 0; i < EO / PL::R_OUT; ++i)
#pragma unroll
      for (int r = 0; r < PL::R_OUT; ++r) {
        int l, e;
        const bool ok = PL::out_pos(tid, i, r, l, e);
        if (FULL)
          tm_st_stream<WIDE>(dst + e * ld + l, c_scale(v[i * PL::R_OUT + r], scale));
        else if (ok && l < ti)
          dst[e * ld + l] = c_scale(v[i * PL::R_OUT + r], scale);
      }
    if constexpr (PRU) {
      ct_ready<BI>(nx);
    } else {
#pragma unroll
      for (int s = 0; s < EI; ++s) v[s] = nx[s];
      ct_ready<EI>(v);
    }
  }
}

int tm_device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

template <int N, int L, int NT, int V = 0, int PH = 0>
static int launch_c2c_ct(const LinePlan& lp, const float2* in, int lin, float2* out, int64_t outer, int64_t inner,
                         int64_t ld, int dir, float scale, hipStream_t s, const char* name,
                         const float2* twx = nullptr, int M = 0) {
  const int64_t tpo = tm_ceil_div(inner, L);
  const int64_t ntiles = outer * tpo;
  // LDS: one exchange buffer, two for the 512-point unpruned kernels
  const size_t shm_pr = (size_t)N * L * sizeof(float2);
  const size_t shm = (TM_C2C_TWO_BUFFERS && N == 512 && !ct2_last_in_x<CtPlan<N, L, NT, true, false, V>>() ? 2 : 1) *
                     shm_pr;
  // persistent grid: exactly the blocks that are resident at once (LDS and
  // register limits), so every block runs from the start
  auto occupancy = [&](const void* kernel, size_t bytes) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, NT, bytes) != hipSuccess || nb < 1) nb = 1;
    return nb;
  };
  static int per_cu = 0, per_cu_pruned = 0;
  if (per_cu == 0) {
    per_cu = occupancy((const void*)k_c2c_ct<N, L, NT, 1, true, V, PH>, shm);
    per_cu_pruned = occupancy((const void*)k_c2c_ct<N, L, NT, -1, true, V, 0, true>, shm_pr);
  }
  const double bytes = 8.0 * ((double)lin + (double)N) * (double)inner * (double)outer;
  using PL = CtPlan<N, L, NT, true, false, V>;
  // unpredicated form: all slots valid, no pruning, tiles inside the stride
  const bool full = (N / PL::R_IN * L) % NT == 0 && (N / PL::R_OUT * L) % NT == 0 && lin == N && tpo * L <= ld;
#define TM_C2C_LAUNCH(D, F)                                                                                      \
  TM_TIMED(name, bytes, s,                                                                                      \
           hipLaunchKernelGGL((k_c2c_ct<N, L, NT, D, F, V>), dim3((unsigned)blocks), dim3(NT), shm, s, in, out, lp.tw, \
                              lin, inner, (int)ld, tpo, ntiles, scale, (const float2*)nullptr, 0))
  // pruned form: zero-padded lines whose first-stage butterflies have one
  // non-zero input, all slots valid, tiles inside the stride
  const bool pruned = PH == 0 && lin <= N / PL::R_IN && (N / PL::R_IN * L) % NT == 0 &&
                      (N / PL::R_OUT * L) % NT == 0 && tpo * L <= ld;
  int64_t blocks = (int64_t)tm_device_cus() * (pruned ? per_cu_pruned : per_cu);
  if (blocks > ntiles) blocks = ntiles;
  if (pruned) {
    if (dir < 0)
      TM_TIMED(name, bytes, s,
               hipLaunchKernelGGL((k_c2c_ct<N, L, NT, -1, true, V, 0, true>), dim3((unsigned)blocks), dim3(NT), shm_pr, s,
                                  in, out, lp.tw, lin, inner, (int)ld, tpo, ntiles, scale, (const float2*)nullptr, 0));
    else
      TM_TIMED(name, bytes, s,
               hipLaunchKernelGGL((k_c2c_ct<N, L, NT, 1, true, V, 0, true>), dim3((unsigned)blocks), dim3(NT), shm_pr, s,
                                  in, out, lp.tw, lin, inner, (int)ld, tpo, ntiles, scale, (const float2*)nullptr, 0));
    TM_CHECK_LAUNCH();
    return TM_OK;
  }
  if constexpr (PH > 0) {
    TM_REQUIRE(full && dir > 0 && in == out && twx && 2 * M <= 256 && inner % L == 0,
               "mirror-pair C2C: in-place inverse over whole unpredicated tiles, rows of at most 129 bins");
    TM_REQUIRE(PH != 1 || (M % 4 == 0 && (M / 4) % 2 == 1 && inner == M - 4),
               "radix-4 units: M = 4 G with G odd and the M - 4 paired positions of a row");
    TM_TIMED(name, bytes, s,
             hipLaunchKernelGGL((k_c2c_ct<N, L, NT, 1, true, V, PH>), dim3((unsigned)blocks), dim3(NT), shm, s, in, out,
                                lp.tw, lin, inner, (int)ld, tpo, ntiles, scale, twx, M));
    TM_CHECK_LAUNCH();
    return TM_OK;
  }
  if (dir < 0) {
    if (full) TM_C2C_LAUNCH(-1, true);
    else TM_C2C_LAUNCH(-1, false);
  } else {
    if (full) TM_C2C_LAUNCH(1, true);
    else TM_C2C_LAUNCH(1, false);
  }
#undef TM_C2C_LAUNCH
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// R2C along contiguous rows: in rows of lin <= n reals (stride lin,
// zero-padded to n), out rows of n/2+1 complex (stride ldo).  pair_h > 0:
// the bins are stored in the mirror-pair order of tm_pair_bin (pair_nm
// paired positions, pairs in lanes l and l ^ pair_h of a column tile).
__global__ void __launch_bounds__(TM_FFT_THREADS)
k_r2c_rows(const float* __restrict__ in, float2* __restrict__ out, LinePlan p, int lin, int64_t nrows, int RL,
           int64_t ldo, int pair_nm, int pair_h) {
  extern __shared__ float2 lds[];
  const int n = p.n, nc = n / 2 + 1;
  const int ld = n + 1;
  float2* A = lds;
  float2* B = lds + RL * ld;
  const int64_t r0 = (int64_t)blockIdx.x * RL;
  const int rl = (int)((nrows - r0) < RL ? (nrows - r0) : RL);
  const int tid = threadIdx.x;
  const float* src = in + r0 * lin;
  for (int f = tid; f < n * rl; f += TM_FFT_THREADS) {
    const int l = f / n, e = f - l * n;
    A[l * ld + e] = make_float2(e < lin ? src[(int64_t)l * lin + e] : 0.f, 0.f);
  }
  __syncthreads();
  float2* R = lds_fft_lines<-1, false>(A, B, rl, ld, 1, p, tid, TM_FFT_THREADS);
  float2* dst = out + r0 * ldo;
  for (int f = tid; f < nc * rl; f += TM_FFT_THREADS) {
    const int l = f / nc, e = f - l * nc;
    dst[l * ldo + (pair_h ? tm_pair_pos(e, pair_nm, nc - 1, pair_h) : e)] = R[l * ld + e];
  }
}

// R2C of rows of lin <= 2M reals through one M-point compile-time transform:
// z[n] = x[2n] + i x[2n+1] (zero past lin), Z = FFT_M(z), then bin k of the
// row is X[k] = (Z[k] + Z*[M-k]) / 2 - i W^k (Z[k] - Z*[M-k]) / 2 with
// W = exp(-2 pi i / 2M) and Z[M] = Z[0] (tw2: the 2M-entry table of W^k).
// RL rows per block; bins stored as k_r2c_rows stores them.
template <int M, int RL, int NT, int V>
__global__ void __launch_bounds__(NT)
k_r2c_rows_ct(const float* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw2, int lin,
              int64_t nrows, int64_t ldo, int pair_nm, int pair_h) {
  using P = CtPlan<M, RL, NT, false, false, V>;
  constexpr int LD = M + 1;
  __shared__ float2 lds[P::LDS_LINES * LD];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * RL;
  const int rl = (int)((nrows - r0) < RL ? (nrows - r0) : RL);
  const float* src = in + r0 * lin;
  float2 v[P::E_IN], o[P::E_OUT];
#pragma unroll
  for (int i = 0; i < P::bpt(P::R_IN); ++i)
#pragma unroll
    for (int r = 0; r < P::R_IN; ++r) {
      int l, e;
      const bool ok = P::in_pos(tid, i, r, l, e);
      float2 z = make_float2(0.f, 0.f);
      if (ok && l < rl && 2 * e < lin) {
        const float* p = src + (int64_t)l * lin + 2 * e;
        z.x = p[0];
        if (2 * e + 1 < lin) z.y = p[1];
      }
      v[i * P::R_IN + r] = z;
    }
  ct_fft<P, -1, LD, 2>(lds, tw2, v, o, tid);
  lds_barrier();  // the last stage has read the buffer
  constexpr int RO = P::R_OUT, NB = M / RO;
#pragma unroll
  for (int i = 0; i < P::bpt(RO); ++i) {
    int l, j;
    if (ct_slot<P, RO>(tid, i, l, j)) {
#pragma unroll
      for (int r = 0; r < RO; ++r) lds[ct_lds<P, LD>(l, j + r * NB)] = o[i * RO + r];
    }
  }
  lds_barrier();
  float2* dst = out + r0 * ldo;
  for (int f = tid; f < rl * (M + 1); f += NT) {
    const int l = f / (M + 1), k = f - l * (M + 1);
    const float2 a = lds[l * LD + (k == M ? 0 : k)];
    const float2 b = lds[l * LD + (k == 0 ? 0 : M - k)];
    float2 x;
    if (k == 0 || k == M) {
      x = make_float2(k == 0 ? a.x + a.y : a.x - a.y, 0.f);
    } else {
      const float2 e = make_float2(a.x + b.x, a.y - b.y);
      const float2 d = c_mul(tw_load<-1>(tw2, k), make_float2(a.x - b.x, a.y + b.y));
      x = make_float2(0.5f * (e.x + d.y), 0.5f * (e.y - d.x));
    }
    dst[l * ldo + (pair_h ? tm_pair_pos(k, pair_nm, M, pair_h) : k)] = x;
  }
}

// C2R along contiguous rows of n/2+1 complex -> n reals, times scale.  The
// missing half is the Hermitian mirror; taking the real part of the complex
// inverse discards the imaginary parts of the DC and Nyquist bins exactly as
// numpy.fft.irfft does.
__global__ void __launch_bounds__(TM_FFT_THREADS)
k_c2r_rows(const float2* __restrict__ in, float* __restrict__ out, LinePlan p, int64_t nrows, int RL,
           float scale) {
  extern __shared__ float2 lds[];
  const int n = p.n, nc = n / 2 + 1;
  const int ld = n + 1;
  float2* A = lds;
  float2* B = lds + RL * ld;
  const int64_t r0 = (int64_t)blockIdx.x * RL;
  const int rl = (int)((nrows - r0) < RL ? (nrows - r0) : RL);
  const int tid = threadIdx.x;
  const float2* src = in + r0 * nc;
  for (int f = tid; f < nc * rl; f += TM_FFT_THREADS) {
    const int l = f / nc, e = f - l * nc;
    const float2 v = src[f];
    A[l * ld + e] = v;
    if (e > 0 && n - e >= nc) A[l * ld + (n - e)] = c_conj(v);
  }
  __syncthreads();
  float2* R = lds_fft_lines<1, false>(A, B, rl, ld, 1, p, tid, TM_FFT_THREADS);
  float* dst = out + r0 * n;
  for (int f = tid; f < n * rl; f += TM_FFT_THREADS) {
    const int l = f / n, e = f - l * n;
    dst[f] = R[l * ld + e].x * scale;
  }
}

// --------------------------------------------------------------------------
// host-side plans

static void factor_length(int n, std::vector<int>& out) {
  int p2 = 0;
  while (n % 2 == 0) { n /= 2; ++p2; }
  // radix-8 stages, finishing with 4 or 4*4 / 2 so no radix-2 stage is used
  // unless the length is 2 * odd.
  int n8 = p2 / 3, rem = p2 % 3;
  if (rem == 1 && n8 > 0) { n8 -= 1; rem = 4; }
  for (int i = 0; i < n8; ++i) out.push_back(8);
  if (rem == 2) out.push_back(4);
  else if (rem == 4) { out.push_back(4); out.push_back(4); }
  else if (rem == 1) out.push_back(2);
  const int odd[] = {5, 3, 7, 11, 13};
  for (int r : odd)
    while (n % r == 0) { out.push_back(r); n /= r; }
  if (n != 1) out.push_back(-n);  // unsupported prime factor
}


static int make_line(int n, tm_line* out) {
  std::vector<int> rad;
  factor_length(n, rad);
  TM_REQUIRE(n >= 1 && n <= TM_FFT_MAX_LEN, "FFT length %d outside [1, %d]", n, TM_FFT_MAX_LEN);
  TM_REQUIRE((int)rad.size() <= TM_MAX_STAGES, "FFT length %d needs too many stages", n);
  for (int r : rad) TM_REQUIRE(r > 0, "FFT length %d has prime factor %d > %d", n, -r, TM_MAX_GENERIC_RADIX);
  memset(out, 0, sizeof(*out));
  out->lp.n = n;
  out->lp.nst = (int)rad.size();
  for (size_t i = 0; i < rad.size(); ++i) out->lp.radix[i] = rad[i];
  std::vector<float2> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    tw[k] = make_float2((float)cos(a), (float)sin(a));
  }
  TM_CHECK_HIP(hipMalloc(&out->tw_dev, sizeof(float2) * n));
  TM_CHECK_HIP(hipMemcpy(out->tw_dev, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  out->lp.tw = out->tw_dev;
  return TM_OK;
}

static int tile_lines(int n, int64_t avail) {
  int t = 4096 / n;
  if (t > 16) t = 16;
  if (t < 1) t = 1;
  if (avail < t) t = (int)avail;
  return t;
}

template <int V>
static int launch_c2c_ct_len(const LinePlan& lp, const float2* in, int lin, float2* out, int64_t outer, int64_t inner,
                             int64_t ld, int dir, float scale, hipStream_t s, const char* name, bool& done) {
  done = true;
  switch (lp.n) {
    case 128: return launch_c2c_ct<128, 16, 256, V>(lp, in, lin, out, outer, inner, ld, dir, scale, s, name);
    case 256: return launch_c2c_ct<256, 16, 256, V>(lp, in, lin, out, outer, inner, ld, dir, scale, s, name);
    case 512:
      // unpredicated 8-line tiles when 16-line tiles would not fit the stride
      if (lin == 512 && tm_ceil_div(inner, 16) * 16 > ld && tm_ceil_div(inner, 8) * 8 <= ld)
        return launch_c2c_ct<512, 8, 512, V>(lp, in, lin, out, outer, inner, ld, dir, scale, s, name);
      return launch_c2c_ct<512, 16, 512, V>(lp, in, lin, out, outer, inner, ld, dir, scale, s, name);
    case 1024: return launch_c2c_ct<1024, 8, 512, V>(lp, in, lin, out, outer, inner, ld, dir, scale, s, name);
    default: done = false; return TM_OK;
  }
}

int tm_launch_c2c(const LinePlan& lp, const float2* in, int lin, float2* out, int64_t outer, int64_t inner,
                  int64_t ld, int dir, float scale, hipStream_t s, const char* name, int variant) {
  TM_REQUIRE(inner >= 1 && inner <= ld, "C2C: %lld inner lines at stride %lld", (long long)inner, (long long)ld);
  TM_REQUIRE(variant == 0 || variant == 1, "radix plan variant %d is not 0 or 1", variant);
  // the specialised kernels address a tile with 32-bit offsets
  if (inner >= 12 && (int64_t)lp.n * ld < (1LL << 31)) {
    bool done = false;
    const int rc = variant ? launch_c2c_ct_len<1>(lp, in, lin, out, outer, inner, ld, dir, scale, s, name, done)
                           : launch_c2c_ct_len<0>(lp, in, lin, out, outer, inner, ld, dir, scale, s, name, done);
    if (done) return rc;
  }
  const int TI = tile_lines(lp.n, inner);
  const int64_t tpo = tm_ceil_div(inner, TI);
  const int64_t blocks = outer * tpo;
  TM_REQUIRE(blocks <= 0x7fffffff, "FFT grid too large");
  const size_t shm = 2 * (size_t)TI * lp.n * sizeof(float2);
  const double bytes = 8.0 * ((double)lin + (double)lp.n) * (double)inner * (double)outer;
  if (dir < 0)
    TM_TIMED(name, bytes, s,
             hipLaunchKernelGGL(k_c2c_strided<-1>, dim3((unsigned)blocks), dim3(TM_FFT_THREADS), shm, s, in, out,
                                lp, lin, inner, ld, TI, tpo, scale));
  else
    TM_TIMED(name, bytes, s,
             hipLaunchKernelGGL(k_c2c_strided<1>, dim3((unsigned)blocks), dim3(TM_FFT_THREADS), shm, s, in, out,
                                lp, lin, inner, ld, TI, tpo, scale));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// In-place inverse 1024-point C2C over tiles of 16 lines (128-byte row
// segments, where the 16-line tile of the single 1024-point plan needs
// 16 elements per thread and spills): the first radix-2 step splits each
// line into its even and odd outputs, a[n] = x[n] + x[n+512] and
// b[n] = (x[n] - x[n+512]) W^n (W = exp(2 pi i / 1024)), whose 512-point
// inverses are outputs 2k and 2k + 1.  A 1024-thread block holds 8
// elements of each half per thread and runs the two 512-point transforms
// one after the other (twiddles: every second entry of the 1024 table);
// persistent blocks prefetch the next tile.  PH: the mirror-pair
// pre-processing of k_c2c_ct before the stores.
#ifndef TM_C2C_1024_HALVES
#define TM_C2C_1024_HALVES 1
#endif
template <int V, int PH>
__global__ void __launch_bounds__(1024)
k_c2c_1024_halves(float2* buf, const float2* __restrict__ tw, int64_t inner, int ld, int64_t tiles_per_outer,
                  int64_t ntiles, float scale, const float2* __restrict__ twx, int M) {
  constexpr int L = 16, NT = 1024;
  using PL = CtPlan<512, L, NT, true, false, V>;
  constexpr int EI = PL::E_IN, EO = PL::E_OUT;
  static_assert(EI == 8 && EO == 8 && (512 / PL::R_IN * L) % NT == 0 && (512 / PL::R_OUT * L) % NT == 0,
                "8 elements of every slot per thread");
  static_assert(PH == 0 || PH == 8, "mirror pairs at lane distance 8 in 16-line tiles");
  extern __shared__ float2 lds[];
  __shared__ float2 tws[1024];
  __shared__ float2 twxs[PH > 0 ? 256 : 1];
  const int tid = threadIdx.x;
  for (int i = tid; i < 1024; i += NT) tws[i] = tw[i];
  if constexpr (PH > 0)
    for (int i = tid; i < 2 * M; i += NT) twxs[i] = twx[i];
  float2 x0[EI], x1[EI];
  // rows half * 512 + e of tile t
  auto load = [&](int64_t t, int half, float2* d) {
    const int64_t o = t / tiles_per_outer;
    const int64_t i0 = (t - o * tiles_per_outer) * (int64_t)L;
    const float2* src = buf + o * 1024 * (int64_t)ld + i0 + (int64_t)half * 512 * ld;
#pragma unroll
    for (int i = 0; i < EI / PL::R_IN; ++i)
#pragma unroll
      for (int r = 0; r < PL::R_IN; ++r) {
        int l, e;
        PL::in_pos(tid, i, r, l, e);
        d[i * PL::R_IN + r] = tm_ld_stream(src + e * ld + l);
      }
  };
  int64_t t = tm_xcd_block();
  if (t >= ntiles) return;
  load(t, 0, x0);
  load(t, 1, x1);
  ct_ready<EI>(x0);
  ct_ready<EI>(x1);
  for (; t < ntiles; t += gridDim.x) {
    // the next tile's halves are loaded into x0 / x1 once this tile no longer
    // needs them: the first before the even outputs' transform, the second
    // before the odd outputs' (the last tile reloads itself)
    const int64_t tn = t + gridDim.x < ntiles ? t + gridDim.x : t;
    float2 a[EI], b[EI];
#pragma unroll
    for (int i = 0; i < EI / PL::R_IN; ++i)
#pragma unroll
      for (int r = 0; r < PL::R_IN; ++r) {
        int l, e;
        PL::in_pos(tid, i, r, l, e);
        const int s2 = i * PL::R_IN + r;
        a[s2] = c_add(x0[s2], x1[s2]);
        b[s2] = c_mul(c_sub(x0[s2], x1[s2]), tw_load<1>(tws, e));
      }
    const int64_t o = t / tiles_per_outer;
    const int64_t i0 = (t - o * tiles_per_outer) * (int64_t)L;
    float2* dst = buf + o * 1024 * (int64_t)ld + i0;
    int kb = 0;
    float2 wb = make_float2(1.f, 0.f);
    if constexpr (PH > 0) {
      const int jt = (int)(i0 / L), lt = tid % L;
      kb = lt < PH ? PH * jt + lt : M - PH * jt - (lt - PH);
      wb = tw_load<1>(twxs, kb);
    }
    auto finish = [&](float2* v, int half) {
      if constexpr (PH > 0) {
#pragma unroll
        for (int s2 = 0; s2 < EO; ++s2) {
          const float2 p = v[s2];
          const float2 q = make_float2(tm_lane_xor<PH>(p.x), tm_lane_xor<PH>(p.y));
          v[s2] = tm_c2r_pre(p, q, wb, kb == 0);
        }
      }
#pragma unroll
      for (int i = 0; i < EO / PL::R_OUT; ++i)
#pragma unroll
        for (int r = 0; r < PL::R_OUT; ++r) {
          int l, k;
          PL::out_pos(tid, i, r, l, k);
          tm_st_stream(dst + (2 * k + half) * ld + l, c_scale(v[i * PL::R_OUT + r], scale));
        }
    };
    // b waits in the second LDS region (each thread reads back only its own
    // slots, so no barrier) while a is transformed
    float2* park = lds + 512 * L;
#pragma unroll
    for (int s2 = 0; s2 < EI; ++s2) park[s2 * NT + tid] = b[s2];
    load(tn, 0, x0);
    ct_fft<PL, 1, L, 2>(lds, tws, a, a, tid);
    finish(a, 0);
    load(tn, 1, x1);
#pragma unroll
    for (int s2 = 0; s2 < EI; ++s2) b[s2] = park[s2 * NT + tid];
    ct_fft<PL, 1, L, 2>(lds, tws, b, b, tid);
    finish(b, 1);
    ct_ready<EI>(x0);
    ct_ready<EI>(x1);
  }
}

template <int V, int PH>
static int launch_c2c_1024_halves(const LinePlan& lp, float2* buf, int64_t outer, int64_t inner, int64_t ld,
                                  float scale, hipStream_t s, const char* name, const float2* twx, int M) {
  TM_REQUIRE(lp.n == 1024 && inner % 16 == 0 && inner <= ld && (PH == 0 || (twx && 2 * M <= 256)),
             "1024-point halves: whole 16-line tiles inside the stride");
  const int64_t tpo = inner / 16, ntiles = outer * tpo;
  const size_t shm = 2 * (size_t)512 * 16 * sizeof(float2);  // the transform's exchange buffer + the parked half
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_c2c_1024_halves<V, PH>, 1024, shm) !=
            hipSuccess ||
        nb < 1)
      nb = 1;
    per_cu = nb;
  }
  int64_t blocks = (int64_t)tm_device_cus() * per_cu;
  if (blocks > ntiles) blocks = ntiles;
  const double bytes = 16.0 * 1024 * (double)inner * (double)outer;
  TM_TIMED(name, bytes, s,
           hipLaunchKernelGGL((k_c2c_1024_halves<V, PH>), dim3((unsigned)blocks), dim3(1024), shm, s, buf, lp.tw, inner,
                              (int)ld, tpo, ntiles, scale, twx, M));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_launch_c2c_pairs(const LinePlan& lp, const float2* twx, int M, int ph, float2* buf, int64_t outer,
                        int64_t inner, int64_t ld, hipStream_t s, const char* name, int variant) {
  TM_REQUIRE(variant == 0 || variant == 1, "radix plan variant %d is not 0 or 1", variant);
  TM_REQUIRE(inner >= 1 && inner <= ld && (int64_t)lp.n * ld < (1LL << 31), "mirror-pair C2C: %lld lines at stride %lld",
             (long long)inner, (long long)ld);
  auto go = [&](auto vtag) -> int {
    constexpr int V = decltype(vtag)::value;
    if (ph == 1 && lp.n == 128)
      return launch_c2c_ct<128, 16, 256, V, 1>(lp, buf, lp.n, buf, outer, inner, ld, 1, 1.0f, s, name, twx, M);
    if (ph == 1 && lp.n == 256)
      return launch_c2c_ct<256, 16, 256, V, 1>(lp, buf, lp.n, buf, outer, inner, ld, 1, 1.0f, s, name, twx, M);
    if (ph == 1 && lp.n == 512)
      return launch_c2c_ct<512, 16, 512, V, 1>(lp, buf, lp.n, buf, outer, inner, ld, 1, 1.0f, s, name, twx, M);
    if (ph == 8 && lp.n == 128)
      return launch_c2c_ct<128, 16, 256, V, 8>(lp, buf, lp.n, buf, outer, inner, ld, 1, 1.0f, s, name, twx, M);
    if (ph == 8 && lp.n == 256)
      return launch_c2c_ct<256, 16, 256, V, 8>(lp, buf, lp.n, buf, outer, inner, ld, 1, 1.0f, s, name, twx, M);
    if (ph == 8 && lp.n == 512)
      return launch_c2c_ct<512, 16, 512, V, 8>(lp, buf, lp.n, buf, outer, inner, ld, 1, 1.0f, s, name, twx, M);
    if (ph == 8 && lp.n == 1024)
      return launch_c2c_1024_halves<V, 8>(lp, buf, outer, inner, ld, 1.0f, s, name, twx, M);
    if (ph == 4 && lp.n == 1024)
      return launch_c2c_ct<1024, 8, 512, V, 4>(lp, buf, lp.n, buf, outer, inner, ld, 1, 1.0f, s, name, twx, M);
    tm_set_error("mirror-pair C2C: no kernel for length %d with pairs at lane distance %d", lp.n, ph);
    return TM_ERR_VALUE;
  };
  return variant ? go(std::integral_constant<int, 1>{}) : go(std::integral_constant<int, 0>{});
}

// Lane distance of the mirror pairs for an in-place inverse C2C of length
// n whose compile-time kernel takes tiles of 2h lines (0: none).
int tm_c2c_pair_h(int n) {
  return (n == 128 || n == 256 || n == 512 || (n == 1024 && TM_C2C_1024_HALVES)) ? 8 : n == 1024 ? 4 : 0;
}

static int rows_per_block(int n, int64_t nrows) {
  int r = 2048 / (n + 1);
  if (r < 1) r = 1;
  if (r > 64) r = 64;
  if (nrows < r) r = (int)nrows;
  return r;
}

int tm_launch_r2c_rows(const LinePlan& lp, const float* in, int lin, float2* out, int64_t nrows, int64_t ldo,
                       hipStream_t s, const char* name, int pair_nm, int pair_h) {
  TM_REQUIRE(ldo >= lp.n / 2 + 1, "R2C: output row stride %lld below %d", (long long)ldo, lp.n / 2 + 1);
  TM_REQUIRE(pair_h == 0 || (pair_h > 0 && pair_nm % (pair_h == 1 ? 8 : 2 * pair_h) == 0 && pair_nm <= lp.n / 2 + 1),
             "R2C: mirror-pair order needs whole tiles of 2 * %d positions", pair_h);
  const int RL = rows_per_block(lp.n, nrows);
  const size_t shm = 2 * (size_t)RL * (lp.n + 1) * sizeof(float2);
  const int64_t blocks = tm_ceil_div(nrows, RL);
  TM_REQUIRE(blocks <= 0x7fffffff, "FFT grid too large");
  TM_TIMED(name, (4.0 * lin + 8.0 * (lp.n / 2 + 1)) * (double)nrows, s,
           hipLaunchKernelGGL(k_r2c_rows, dim3((unsigned)blocks), dim3(TM_FFT_THREADS), shm, s, in, out, lp, lin,
                              nrows, RL, ldo, pair_nm, pair_h));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// 0: the fused scan's x pass runs the generic LDS kernel k_r2c_rows.
#ifndef TM_X_PASS_CT
#define TM_X_PASS_CT 1
#endif

template <int M, int RL, int NT, int V>
static int launch_r2c_rows_ct(const LinePlan& lp, const float* in, int lin, float2* out, int64_t nrows,
                              int64_t ldo, hipStream_t s, const char* name, int pair_nm, int pair_h) {
  const int64_t blocks = tm_ceil_div(nrows, RL);
  TM_REQUIRE(blocks <= 0x7fffffff, "FFT grid too large");
  TM_TIMED(name, (4.0 * lin + 8.0 * (M + 1)) * (double)nrows, s,
           hipLaunchKernelGGL((k_r2c_rows_ct<M, RL, NT, V>), dim3((unsigned)blocks), dim3(NT), 0, s, in, out, lp.tw,
                              lin, nrows, ldo, pair_nm, pair_h));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_launch_r2c_rows_ct(const LinePlan& lp, const float* in, int lin, float2* out, int64_t nrows, int64_t ldo,
                          hipStream_t s, const char* name, int pair_nm, int pair_h, int variant) {
  TM_REQUIRE(ldo >= lp.n / 2 + 1, "R2C: output row stride %lld below %d", (long long)ldo, lp.n / 2 + 1);
  TM_REQUIRE(pair_h == 0 || (pair_h > 0 && pair_nm % (pair_h == 1 ? 8 : 2 * pair_h) == 0 && pair_nm <= lp.n / 2 + 1),
             "R2C: mirror-pair order needs whole tiles of 2 * %d positions", pair_h);
  TM_REQUIRE(lin >= 0 && lin <= lp.n, "R2C: row of %d reals longer than %d", lin, lp.n);
  TM_REQUIRE(variant == 0 || variant == 1, "radix plan variant %d is not 0 or 1", variant);
#define TM_R2C_CT(M_, RL_)                                                                                      \
  return variant ? launch_r2c_rows_ct<M_, RL_, 256, 1>(lp, in, lin, out, nrows, ldo, s, name, pair_nm, pair_h) \
                 : launch_r2c_rows_ct<M_, RL_, 256, 0>(lp, in, lin, out, nrows, ldo, s, name, pair_nm, pair_h)
  switch (TM_X_PASS_CT ? lp.n : 0) {
    case 128: TM_R2C_CT(64, 32);
    case 200: TM_R2C_CT(100, 10);
    case 256: TM_R2C_CT(128, 16);
    default: return tm_launch_r2c_rows(lp, in, lin, out, nrows, ldo, s, name, pair_nm, pair_h);
  }
#undef TM_R2C_CT
}

int tm_launch_c2r_rows(const LinePlan& lp, const float2* in, float* out, int64_t nrows, float scale, hipStream_t s,
                       const char* name) {
  const int RL = rows_per_block(lp.n, nrows);
  const size_t shm = 2 * (size_t)RL * (lp.n + 1) * sizeof(float2);
  const int64_t blocks = tm_ceil_div(nrows, RL);
  TM_REQUIRE(blocks <= 0x7fffffff, "FFT grid too large");
  TM_TIMED(name, (4.0 * lp.n + 8.0 * (lp.n / 2 + 1)) * (double)nrows, s,
           hipLaunchKernelGGL(k_c2r_rows, dim3((unsigned)blocks), dim3(TM_FFT_THREADS), shm, s, in, out, lp, nrows,
                              RL, scale));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

int tm_fft_r2c_batch(const tm_fft_plan* p, int64_t batch, const float* in, float2* out, hipStream_t s) {
  const int nd = p->ndim;
  int64_t nrows = batch;
  for (int i = 0; i < nd - 1; ++i) nrows *= p->shape[i];
  const LinePlan& last = p->lines[nd - 1].lp;
  int rc = tm_launch_r2c_rows(last, in, last.n, out, nrows, last.n / 2 + 1, s, "fft_r2c_rows");
  if (rc != TM_OK) return rc;
  for (int ax = nd - 2; ax >= 0; --ax) {
    int64_t outer = batch, inner = 1;
    for (int i = 0; i < ax; ++i) outer *= p->cshape[i];
    for (int i = ax + 1; i < nd; ++i) inner *= p->cshape[i];
    rc = tm_launch_c2c(p->lines[ax].lp, out, p->lines[ax].lp.n, out, outer, inner, inner, -1, 1.0f, s,
                       "fft_c2c_axis");
    if (rc != TM_OK) return rc;
  }
  return TM_OK;
}

int tm_fft_c2r_batch(const tm_fft_plan* p, int64_t batch, float2* in, float* out, hipStream_t s) {
  const int nd = p->ndim;
  double N = 1.0;
  for (int i = 0; i < nd; ++i) N *= (double)p->shape[i];
  for (int ax = 0; ax <= nd - 2; ++ax) {
    int64_t outer = batch, inner = 1;
    for (int i = 0; i < ax; ++i) outer *= p->cshape[i];
    for (int i = ax + 1; i < nd; ++i) inner *= p->cshape[i];
    int rc = tm_launch_c2c(p->lines[ax].lp, in, p->lines[ax].lp.n, in, outer, inner, inner, 1, 1.0f, s,
                           "fft_c2c_axis");
    if (rc != TM_OK) return rc;
  }
  int64_t nrows = batch;
  for (int i = 0; i < nd - 1; ++i) nrows *= p->shape[i];
  return tm_launch_c2r_rows(p->lines[nd - 1].lp, in, out, nrows, (float)(1.0 / N), s, "fft_c2r_rows");
}

extern "C" {

int tm_fft_plan_create(int ndim, const int64_t* shape, int64_t batch, tm_fft_plan** out) {
  TM_REQUIRE(out != nullptr, "null plan output");
  TM_REQUIRE(ndim >= 1 && ndim <= 3, "FFT rank %d not in [1, 3]", ndim);
  TM_REQUIRE(batch >= 1, "FFT batch must be >= 1");
  tm_fft_plan* p = (tm_fft_plan*)calloc(1, sizeof(tm_fft_plan));
  TM_REQUIRE(p != nullptr, "host allocation failed");
  p->ndim = ndim;
  p->batch = batch;
  for (int i = 0; i < ndim; ++i) {
    if (shape[i] < 1 || shape[i] > TM_FFT_MAX_LEN) {
      free(p);
      tm_set_error("FFT axis length %lld outside [1, %d]", (long long)shape[i], TM_FFT_MAX_LEN);
      return TM_ERR_VALUE;
    }
    p->shape[i] = shape[i];
    p->cshape[i] = shape[i];
  }
  p->cshape[ndim - 1] = shape[ndim - 1] / 2 + 1;
  for (int i = 0; i < ndim; ++i) {
    int rc = make_line((int)shape[i], &p->lines[i]);
    if (rc != TM_OK) {
      for (int j = 0; j < i; ++j) hipFree(p->lines[j].tw_dev);
      free(p);
      return rc;
    }
  }
  *out = p;
  return TM_OK;
}

int tm_fft_plan_destroy(tm_fft_plan* p) {
  if (!p) return TM_OK;
  for (int i = 0; i < p->ndim; ++i)
    if (p->lines[i].tw_dev) hipFree(p->lines[i].tw_dev);
  free(p);
  return TM_OK;
}

int tm_fft_r2c(const tm_fft_plan* p, const float* in, void* out, void* stream) {
  TM_REQUIRE(p && in && out, "null argument to tm_fft_r2c");
  return tm_fft_r2c_batch(p, p->batch, in, (float2*)out, (hipStream_t)stream);
}

int tm_fft_c2r(const tm_fft_plan* p, void* in, float* out, void* stream) {
  TM_REQUIRE(p && in && out, "null argument to tm_fft_c2r");
  return tm_fft_c2r_batch(p, p->batch, (float2*)in, out, (hipStream_t)stream);
}

// Expose the radix plan for a length (host-only helper used by tests).
int tm_fft_factor(int n, int* radices, int max_out, int* n_out) {
  std::vector<int> rad;
  factor_length(n, rad);
  TM_REQUIRE((int)rad.size() <= max_out, "output too small");
  for (size_t i = 0; i < rad.size(); ++i) radices[i] = rad[i];
  *n_out = (int)rad.size();
  return TM_OK;
}

}  // extern "C"
