// Peak extraction on the max-score volume (PeakCallerMaximumFilter,
// PeakCallerSort and the distance filter of PeakCaller._update):
//
//   * box maximum filter of edge d with constant-zero borders and the peak
//     test score == filtered score, written as C-ordered coordinates
//     (tme/backends/npfftw_backend.py:200-206: scipy maximum_filter with
//     size=(d,)*3, mode="constant", then nonzero);
//   * top-k indices of a float array, largest first, ties by lower index
//     (tme/backends/npfftw_backend.py:176-181, argpartition + argsort; numpy
//     orders NaN above every number and so does this);
//   * the greedy minimum-distance filter over coordinates already ordered by
//     score (tme/analyzer/peaks.py:68-106 with the int64 overload of
//     find_candidate_indices, tme/external/bindings.cpp:84-124: a point is
//     dropped when trunc(sqrt(d2)) <= min_distance for a kept point, i.e.
//     d2 < (min_distance + 1)^2).
//
// All three are HBM-streaming or latency-bound integer/compare work; none of
// it is shaped for MFMA.  Order-preserving compaction uses per-block counts,
// one single-block exclusive scan and a ballot-ranked write, so every output
// is deterministic.
#include <algorithm>
#include <mutex>

#include "tm_common.h"

#define PK_THREADS 256                 // threads per compaction block (4 waves)
#define PK_MF_TX 64                    // max-filter plane tile: up to 64 x 16
#define PK_MF_TY 16
#define PK_MF_TZ 8                     // z planes per block of the xy pass
#define PK_XY_PF 7                     // prefetched tile values per thread (the xy pass's small tiles)
#define PK_MAX_SIZE 64                 // largest filter edge
#define PK_SORT_LDS 8192               // bitonic sort in LDS up to this many keys
#define PK_SORT_MAX 65536              // larger sorts run on a global buffer
#define PK_KEEP_MAX 4096               // kept points of the distance filter
#define PK_TOPK_ITEMS 16               // elements per thread in the top-k scans
#define PK_MF_ITEMS 16                 // voxels per thread in the peak test and write
#define PK_MF_CHUNK (PK_THREADS * PK_MF_ITEMS)
#define PK_WRITE_BLOCKS 2048           // blocks of the coordinate write (8 per CU)
#define PK_CF_MAX 8                    // earlier conflicts listed per candidate
#define PK_PAR_MAX 16384               // candidates the parallel distance filter takes
#define PK_TOPK_CHUNK (PK_THREADS * PK_TOPK_ITEMS)

// ---------------------------------------------------------------------------
// shared helpers

// exclusive scan of n int64 counts into off[0..n], off[n] = total; one block
// of 1024 threads walking the counts in coalesced tiles of 1024
__global__ __launch_bounds__(1024) void k_pk_scan(const int64_t* __restrict__ cnt, int64_t* __restrict__ off,
                                                  int64_t n) {
  __shared__ int64_t wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int64_t run = 0;
  for (int64_t base = 0; base < n; base += 1024) {
    const int64_t i = base + t;
    const int64_t v = i < n ? cnt[i] : 0;
    int64_t x = v;  // inclusive scan within the wave
    for (int d = 1; d < 64; d <<= 1) {
      const int64_t y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int64_t wb = 0, tot = 0;
    for (int k = 0; k < 16; ++k) {
      const int64_t q = wsum[k];
      if (k < w) wb += q;
      tot += q;
    }
    if (i < n) off[i] = run + wb + x - v;
    run += tot;
    __syncthreads();
  }
  if (t == 0) off[n] = run;
}

// rank of this lane's flag among the set flags of the block (in thread
// order) and the block's total; 256 threads
__device__ __forceinline__ int pk_block_rank(bool flag, int* wsum, int* total) {
  const unsigned long long b = __ballot(flag);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = __popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[w] = __popcll(b);
  __syncthreads();
  int base = 0;
  for (int i = 0; i < w; ++i) base += wsum[i];
  *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();
  return base + r;
}

// ---------------------------------------------------------------------------
// maximum filter

// x and y maxima over a tx x 16 plane tile with a (d-1) halo, tx <= 64
// chosen so the x tiles split the rows evenly; outside the volume the input
// is the constant 0
__global__ __launch_bounds__(256) void k_mf_xy(const float* __restrict__ s, float* __restrict__ out, int n0, int n1,
                                               int n2, int d, int tx) {
  extern __shared__ float sm[];
  const int lo = d / 2;
  const int hx = tx + d - 1, hy = PK_MF_TY + d - 1;
  float* in = sm;             // hy x hx
  float* mx = sm + hy * hx;   // hy x tx
  const int x0 = blockIdx.x * tx, y0 = blockIdx.y * PK_MF_TY;
  const int zb = blockIdx.z * PK_MF_TZ, ze = min(zb + PK_MF_TZ, n0);
  // tiles of at most PK_XY_PF values per thread (narrow volumes): the next
  // plane's tile is loaded into registers while the current one is filtered
  const bool pf = hy * hx <= PK_XY_PF * 256;
  float pre[PK_XY_PF];
  auto fetch = [&](int64_t z) {
    const float* plane = s + z * (int64_t)n1 * n2;
#pragma unroll
    for (int k = 0; k < PK_XY_PF; ++k) {
      const int i = threadIdx.x + k * 256;
      const int r = i / hx, c = i - r * hx;
      const int y = y0 - lo + r, x = x0 - lo + c;
      pre[k] = (i < hy * hx && y >= 0 && y < n1 && x >= 0 && x < n2) ? plane[(int64_t)y * n2 + x] : 0.f;
    }
  };
  if (pf) fetch(zb);
  for (int64_t z = zb; z < ze; ++z) {
  __syncthreads();  // the previous plane's tiles are consumed
  if (pf) {
#pragma unroll
    for (int k = 0; k < PK_XY_PF; ++k) {
      const int i = threadIdx.x + k * 256;
      if (i < hy * hx) in[i] = pre[k];
    }
    if (z + 1 < ze) fetch(z + 1);
  } else {
    const float* plane = s + z * (int64_t)n1 * n2;
    for (int i = threadIdx.x; i < hy * hx; i += blockDim.x) {
      const int r = i / hx, c = i - r * hx;
      const int y = y0 - lo + r, x = x0 - lo + c;
      in[i] = (y >= 0 && y < n1 && x >= 0 && x < n2) ? plane[(int64_t)y * n2 + x] : 0.f;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < hy * tx; i += blockDim.x) {
    const int r = i / tx, c = i - r * tx;
    const float* row = in + r * hx + c;
    float m = row[0];
    for (int k = 1; k < d; ++k) m = fmaxf(m, row[k]);
    mx[i] = m;
  }
  __syncthreads();
  float* oplane = out + z * (int64_t)n1 * n2;
  for (int i = threadIdx.x; i < PK_MF_TY * tx; i += blockDim.x) {
    const int r = i / tx, c = i - r * tx;
    const int y = y0 + r, x = x0 + c;
    if (y >= n1 || x >= n2) continue;
    const float* col = mx + r * tx + c;
    float m = col[0];
    for (int k = 1; k < d; ++k) m = fmaxf(m, col[k * tx]);
    oplane[(int64_t)y * n2 + x] = m;
  }
  }
}

// z maximum of the xy maxima and the peak test over a block of 4096
// consecutive voxels: one ballot word per 64 voxels (word index = voxel / 64)
// and the block's peak count
__global__ __launch_bounds__(PK_THREADS) void k_mf_z(const float* __restrict__ s, const float* __restrict__ mxy,
                                                     unsigned long long* __restrict__ mask,
                                                     int64_t* __restrict__ cnt, int n0, int n1, int n2, int d) {
  __shared__ int wsum[4];
  const int64_t F = (int64_t)n0 * n1 * n2, plane = (int64_t)n1 * n2;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lo = d / 2;
  int total = 0;
  for (int it = 0; it < PK_MF_ITEMS; ++it) {
    const int64_t L = (int64_t)blockIdx.x * PK_MF_CHUNK + it * PK_THREADS + threadIdx.x;
    bool peak = false;
    if (L < F) {
      const int64_t z = L / plane;
      bool any_out = false;
      float mi = -INFINITY;
      for (int k = 0; k < d; ++k) {
        const int64_t zz = z - lo + k;
        if (zz < 0 || zz >= n0) {
          any_out = true;
          continue;
        }
        mi = fmaxf(mi, mxy[L + (zz - z) * plane]);
      }
      const float m = any_out ? fmaxf(mi, 0.f) : mi;
      peak = (s[L] == m);
    }
    const unsigned long long b = __ballot(peak);
    if (lane == 0) {
      mask[((int64_t)blockIdx.x * PK_MF_CHUNK + it * PK_THREADS) / 64 + w] = b;
      total += __popcll(b);
    }
  }
  if (lane == 0) wsum[w] = total;
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = (int64_t)wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// The whole filter and the peak test for edges up to 32 in one pass over
// the volume: a block owns a tile of tx columns (tx <= 64, splitting the
// rows evenly) x 16 rows and a chunk of PK_ZC z planes, and walks the planes
// z0 - LO .. z0 + PK_ZC - 1 + HI.  Each plane's tile with its (D - 1) halo
// goes into LDS (two buffers, the next plane loaded into registers while
// the current one is filtered), its x maxima into a second LDS array and
// its y maxima into a window of the last D planes held in registers, one
// per row of a thread: of the block's NW waves, wave w owns rows w, w + NW,
// ... (16 / NW rows), lane c column x0 + c.  The window's maximum is the box
// maximum of plane zi - HI, tested against that plane's tile centre (kept
// in registers since).  A wave's ballot covers 64 consecutive voxels of a
// row and is OR-ed into the linear-order words (zeroed beforehand), split
// across a word boundary when needed.  Outside the volume the input is the
// constant 0.  Blocks take the tiles in the order (z chunk, x tile, y tile)
// with the y tile fastest, so neighbouring y tiles, which share halo rows,
// run together on one XCD.  Edges up to 8 run 4 waves per block, up to 16
// 8 waves and larger ones 16 (the windows of 4 or 2 rows per thread would
// not fit in registers).
#ifndef PK_ZC
#define PK_ZC 32
#endif
#define PK_ZCOL_MAX 32

template <int D, int NW>
__global__ __launch_bounds__(64 * NW) void k_mf_fused(const float* __restrict__ s,
                                                      unsigned long long* __restrict__ mask, int n0, int n1, int n2,
                                                      int tx, int ntx, int nty, int64_t nb) {
  constexpr int NT = 64 * NW, RPT = PK_MF_TY / NW;  // threads, rows per thread
  static_assert(PK_MF_TY % NW == 0, "whole rows per wave");
  constexpr int LO = D / 2, HI = D - 1 - LO;
  constexpr int HY = PK_MF_TY + D - 1, HXM = PK_MF_TX + D - 1;  // tile rows, widest tile row
  constexpr int NPF = (HY * HXM + NT - 1) / NT;                 // tile values per thread
  __shared__ float tin[2][HY * HXM];
  __shared__ float tmx[HY * PK_MF_TX];
  const int64_t lb = tm_xcd_block();
  if (lb >= nb) return;  // padding blocks of the XCD-grouped grid
  const int yt = (int)(lb % nty), rest = (int)(lb / nty);
  const int xt = rest % ntx, zc = rest / ntx;
  const int x0 = xt * tx, y0 = yt * PK_MF_TY, z0 = zc * PK_ZC, zend = min(z0 + PK_ZC, n0);
  const int hx = tx + D - 1, nin = HY * hx;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int cl = lane < tx ? lane : tx - 1;  // lanes past the tile read a valid column and test nothing
  const int64_t plane = (int64_t)n1 * n2;
  // offsets in the plane of this thread's tile values (-1 outside the volume)
  int goff[NPF];
#pragma unroll
  for (int k = 0; k < NPF; ++k) {
    const int i = tid + k * NT, r = i / hx, c = i - r * hx;
    const int y = y0 - LO + r, x = x0 - LO + c;
    goff[k] = (i < nin && y >= 0 && y < n1 && x >= 0 && x < n2) ? y * n2 + x : -1;
  }
  float pre[NPF];
  auto fetch = [&](int z) {
    const bool zin = z >= 0 && z < n0;
    const float* p = s + (int64_t)(zin ? z : 0) * plane;
#pragma unroll
    for (int k = 0; k < NPF; ++k) pre[k] = (zin && goff[k] >= 0) ? p[goff[k]] : 0.f;
  };
  bool valid[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) valid[j] = lane < tx && x0 + lane < n2 && y0 + w + NW * j < n1;
  float win[RPT][D], ctr[RPT][HI + 1];
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
#pragma unroll
    for (int k = 0; k < D; ++k) win[j][k] = 0.f;
#pragma unroll
    for (int k = 0; k <= HI; ++k) ctr[j][k] = 0.f;
  }
  const int nst = zend - z0 + D - 1;
  fetch(z0 - LO);
  for (int t = 0; t < nst; ++t) {
    const int zi = z0 - LO + t;
    float* cur = tin[t & 1];
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int i = tid + k * NT;
      if (i < nin) cur[i] = pre[k];
    }
    if (t + 1 < nst) fetch(zi + 1);
    __syncthreads();  // this plane's tile is in LDS; the last plane's y maxima are read
    for (int r = w; r < HY; r += NW) {
      const float* row = cur + r * hx + cl;
      float m = row[0];
#pragma unroll
      for (int k = 1; k < D; ++k) m = fmaxf(m, row[k]);
      tmx[r * PK_MF_TX + lane] = m;
    }
    __syncthreads();  // x maxima in LDS
    const int zo = zi - HI;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int r = w + NW * j;
      float m = tmx[r * PK_MF_TX + lane];
#pragma unroll
      for (int k = 1; k < D; ++k) m = fmaxf(m, tmx[(r + k) * PK_MF_TX + lane]);
#pragma unroll
      for (int k = 0; k < D - 1; ++k) win[j][k] = win[j][k + 1];
      win[j][D - 1] = m;
#pragma unroll
      for (int k = 0; k < HI; ++k) ctr[j][k] = ctr[j][k + 1];
      ctr[j][HI] = cur[(r + LO) * hx + cl + LO];
      if (zo < z0) continue;  // uniform: the window is not full yet
      float mz = win[j][0];
#pragma unroll
      for (int k = 1; k < D; ++k) mz = fmaxf(mz, win[j][k]);
      const bool peak = valid[j] && ctr[j][0] == mz;
      const unsigned long long b = __ballot(peak);
      if (lane == 0 && b) {
        const int64_t L0 = (int64_t)zo * plane + (int64_t)(y0 + r) * n2 + x0;
        const int64_t wd = L0 >> 6;
        const int sh = (int)(L0 & 63);
        atomicOr(&mask[wd], b << sh);
        if (sh && (b >> (64 - sh))) atomicOr(&mask[wd + 1], b >> (64 - sh));
      }
    }
  }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// peak count of each 4096-voxel block of the bit words (64 words per block)
__global__ __launch_bounds__(64) void k_mask_count(const unsigned long long* __restrict__ mask,
                                                   int64_t* __restrict__ cnt) {
  int c = __popcll(mask[(int64_t)blockIdx.x * 64 + threadIdx.x]);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if (threadIdx.x == 0) cnt[blockIdx.x] = c;
}

// C-ordered coordinates of the set bits.  Blocks stride over the 64-word
// groups (4096 voxels; group g's first output is off[g]), loading the next
// group's words and offset before writing the current group's
// coordinates; a wave visits only its words that hold set bits.
__global__ __launch_bounds__(PK_THREADS) void k_mf_write(const unsigned long long* __restrict__ mask,
                                                         const int64_t* __restrict__ off, int64_t* __restrict__ xyz,
                                                         int n1, int n2, int64_t F) {
  // wave w writes the 16 words from 16 w on of a group; every wave loads
  // all 64 words (one per lane) for its prefix, so no barrier is needed
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t plane = (int64_t)n1 * n2;
  const int64_t nw = (F + 63) / 64, ng = (nw + 63) / 64;
  auto word = [&](int64_t g) { return g < ng && g * 64 + lane < nw ? mask[g * 64 + lane] : 0ull; };
  int64_t g = blockIdx.x;
  unsigned long long wv = word(g);
  int64_t base = g < ng ? off[g] : 0;
  for (; g < ng; g += gridDim.x) {
    const int64_t gn = g + gridDim.x;
    const unsigned long long wvn = word(gn);
    const int64_t basen = gn < ng ? off[gn] : 0;
    const int64_t wd0 = g * 64;
    const int c = __popcll(wv);
    // inclusive prefix of the 64 counts over the lanes
    int pre = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(pre, o, 64);
      if (lane >= o) pre += v;
    }
    // this wave's words with set bits (words past the volume are zero)
    unsigned long long todo = (__ballot(wv != 0ull) >> (16 * w)) & 0xffffull;
    while (todo) {  // uniform over the wave
      const int q = 16 * w + __ffsll((long long)todo) - 1;  // word of the group
      todo &= todo - 1ull;
      const int64_t wd = wd0 + q;
      const unsigned long long b = __shfl(wv, q, 64);
      const int before = __shfl(pre, q, 64) - __shfl(c, q, 64);  // set bits in the group's words before q
      const int64_t L = wd * 64 + lane;
      if (((b >> lane) & 1ull) && L < F) {
        const int64_t o = base + before + __popcll(b & ((1ull << lane) - 1ull));
        int64_t z, y, x;
        if (F <= 0xffffffffll) {  // uniform: 32-bit divisions
          const unsigned Lu = (unsigned)L, pu = (unsigned)plane, zu = Lu / pu, ru = Lu - zu * pu,
                         yu = ru / (unsigned)n2;
          z = zu;
          y = yu;
          x = ru - yu * (unsigned)n2;
        } else {
          z = L / plane;
          const int64_t r = L - z * plane;
          y = r / n2;
          x = r - y * n2;
        }
        xyz[3 * o + 0] = z;
        xyz[3 * o + 1] = y;
        xyz[3 * o + 2] = x;
      }
    }
    wv = wvn;
    base = basen;
  }
}

// ---------------------------------------------------------------------------
// top-k

struct pk_sel_state {
  unsigned prefix, pmask;
  int64_t k_rem, n_gt;
};

// float -> uint32 with the float order (NaN above +inf, as numpy sorts it)
__device__ __forceinline__ unsigned pk_key(float f) {
  if (f != f) return 0xffffffffu;
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_topk_init(pk_sel_state* st, unsigned* hist, int64_t k) {
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) hist[i] = 0u;
  if (threadIdx.x == 0) {
    st->prefix = 0u;
    st->pmask = 0u;
    st->k_rem = k;
    st->n_gt = 0;
  }
}

__global__ __launch_bounds__(PK_THREADS) void k_topk_hist(const float* __restrict__ v, int64_t n,
                                                          const pk_sel_state* __restrict__ st,
                                                          unsigned* __restrict__ hist, int shift) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const unsigned prefix = st->prefix, pmask = st->pmask;
  for (int64_t i = (int64_t)blockIdx.x * PK_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * PK_THREADS) {
    const unsigned key = pk_key(v[i]);
    if ((key & pmask) == prefix) atomicAdd(&h[(key >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// picks the digit bin holding the k_rem-th largest remaining key: the one
// bin b with (keys in bins above b) < k_rem <= (keys in bins b and above)
__global__ __launch_bounds__(256) void k_topk_pick(pk_sel_state* st, const unsigned* __restrict__ hist, int shift) {
  __shared__ int64_t suf[256];
  const int b = threadIdx.x;
  const int64_t k_rem = st->k_rem;
  const int64_t h = hist[b];
  suf[b] = h;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const int64_t v = b + d < 256 ? suf[b + d] : 0;
    __syncthreads();
    suf[b] += v;
    __syncthreads();
  }
  const int64_t above = suf[b] - h;
  if (above < k_rem && k_rem <= suf[b]) {
    st->n_gt += above;
    st->k_rem = k_rem - above;
    st->prefix |= (unsigned)b << shift;
    st->pmask |= 255u << shift;
  }
}

// per-chunk counts of keys above the threshold and equal to it
__global__ __launch_bounds__(PK_THREADS) void k_topk_count(const float* __restrict__ v, int64_t n,
                                                           const pk_sel_state* __restrict__ st,
                                                           int64_t* __restrict__ cgt, int64_t* __restrict__ ceq) {
  __shared__ int wsum[4];
  const unsigned T = st->prefix;
  const int64_t c0 = (int64_t)blockIdx.x * PK_TOPK_CHUNK;
  int ngt = 0, neq = 0;
  for (int it = 0; it < PK_TOPK_ITEMS; ++it) {
    const int64_t i = c0 + it * PK_THREADS + threadIdx.x;
    if (i < n) {
      const unsigned key = pk_key(v[i]);
      ngt += key > T;
      neq += key == T;
    }
  }
  // block sums through wave reductions
  for (int o = 32; o > 0; o >>= 1) {
    ngt += __shfl_down(ngt, o);
    neq += __shfl_down(neq, o);
  }
  __shared__ int wsum2[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    wsum[w] = ngt;
    wsum2[w] = neq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    cgt[blockIdx.x] = (int64_t)wsum[0] + wsum[1] + wsum[2] + wsum[3];
    ceq[blockIdx.x] = (int64_t)wsum2[0] + wsum2[1] + wsum2[2] + wsum2[3];
  }
}

// writes the selected indices: every key above the threshold, and the
// k_rem equal keys of lowest index, each class in index order
__global__ __launch_bounds__(PK_THREADS) void k_topk_write(const float* __restrict__ v, int64_t n,
                                                           const pk_sel_state* __restrict__ st,
                                                           const int64_t* __restrict__ ogt,
                                                           const int64_t* __restrict__ oeq,
                                                           unsigned long long* __restrict__ sel) {
  __shared__ int wsum[4];
  const unsigned T = st->prefix;
  const int64_t n_gt = st->n_gt, need = st->k_rem;
  int64_t bgt = ogt[blockIdx.x], beq = oeq[blockIdx.x];
  const int64_t c0 = (int64_t)blockIdx.x * PK_TOPK_CHUNK;
  for (int it = 0; it < PK_TOPK_ITEMS; ++it) {
    const int64_t i = c0 + it * PK_THREADS + threadIdx.x;
    unsigned key = 0u;
    if (i < n) key = pk_key(v[i]);
    const bool gt = i < n && key > T, eq = i < n && key == T;
    int tgt, teq;
    const int rgt = pk_block_rank(gt, wsum, &tgt);
    const int req = pk_block_rank(eq, wsum, &teq);
    const unsigned long long comp = ((unsigned long long)(~key) << 32) | (unsigned long long)(unsigned)i;
    if (gt) sel[bgt + rgt] = comp;
    if (eq && beq + req < need) sel[n_gt + beq + req] = comp;
    bgt += tgt;
    beq += teq;
  }
}

// compare-exchange of element e with element e ^ j (j < 64, so both lie in
// the wave's 64-element segment), ascending where (e & k) == 0
__device__ __forceinline__ unsigned long long pk_cas(unsigned long long v, int e, int j, int k) {
  const unsigned long long o = __shfl_xor(v, j, 64);
  const unsigned long long mn = v < o ? v : o, mx = v < o ? o : v;
  return (((e & j) == 0) == ((e & k) == 0)) ? mn : mx;
}

// one bitonic stage (k, j) over the block through a[] (LDS or global)
__device__ __forceinline__ void pk_stage(unsigned long long* a, int np2, int k, int j) {
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    const int p = i ^ j;
    if (p > i) {
      const unsigned long long x = a[i], y = a[p];
      const bool up = (i & k) == 0;
      if ((x > y) == up) {
        a[i] = y;
        a[p] = x;
      }
    }
  }
  __syncthreads();
}

// ascending bitonic sort of np2 64-bit keys (np2 a power of two) by the
// whole block (a multiple of 64 threads), in place: the stages with a
// stride below 64 run in registers within each wave's 64-element segment
// (shuffles, no barrier), the others through a[]
__device__ __forceinline__ void pk_bitonic(unsigned long long* a, int np2) {
  if (np2 < 64) {
    for (int k = 2; k <= np2; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) pk_stage(a, np2, k, j);
    return;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int seg = w; seg < np2 / 64; seg += nw) {
    const int e = seg * 64 + lane;
    unsigned long long v = a[e];
    for (int k = 2; k <= 64; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) v = pk_cas(v, e, j, k);
    a[e] = v;
  }
  __syncthreads();
  for (int k = 128; k <= np2; k <<= 1) {
    for (int j = k >> 1; j >= 64; j >>= 1) pk_stage(a, np2, k, j);
    for (int seg = w; seg < np2 / 64; seg += nw) {
      const int e = seg * 64 + lane;
      unsigned long long v = a[e];
      for (int j = 32; j > 0; j >>= 1) v = pk_cas(v, e, j, k);
      a[e] = v;
    }
    __syncthreads();
  }
}

// ascending bitonic sort of np2 64-bit keys (np2 a power of two) by one
// block; keys beyond n are the sentinel ~0
template <bool LDS>
__global__ __launch_bounds__(1024) void k_topk_sort(unsigned long long* __restrict__ sel, int64_t n, int np2,
                                                    int64_t* __restrict__ out) {
  extern __shared__ unsigned long long sk[];
  unsigned long long* a = LDS ? sk : sel;
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    if (LDS) a[i] = i < n ? sel[i] : ~0ull;
    else if (i >= n) a[i] = ~0ull;
  }
  __syncthreads();
  pk_bitonic(a, np2);
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = (int64_t)(a[i] & 0xffffffffull);
}

// top-k of few values (n <= PK_SORT_LDS) in one block: every value's
// composite (~key, index) sorted in LDS, the first k indices written; the
// same order as the selection path (key descending, ties by lower index)
__global__ __launch_bounds__(1024) void k_topk_small(const float* __restrict__ v, int n, int np2, int k,
                                                     int64_t* __restrict__ out) {
  extern __shared__ unsigned long long sk[];
  for (int i = threadIdx.x; i < np2; i += blockDim.x)
    sk[i] = i < n ? ((unsigned long long)(~pk_key(v[i])) << 32) | (unsigned long long)(unsigned)i : ~0ull;
  __syncthreads();
  pk_bitonic(sk, np2);
  for (int i = threadIdx.x; i < k; i += blockDim.x) out[i] = (int64_t)(sk[i] & 0xffffffffull);
}

// ---------------------------------------------------------------------------
// distance filter, parallel form.  Greedy selection in order is the fixed
// point "i is kept iff no kept j < i lies within the distance": a candidate
// without earlier neighbours is kept; the others are decided in rounds once
// a neighbour is known kept (drop) or all are known dropped (keep).  Each
// round decides at least the lowest undecided candidate, so at most n rounds.

__device__ __forceinline__ bool pk_near(const int64_t* __restrict__ a, const int64_t* __restrict__ b, int64_t md,
                                        int64_t r2) {
  const int64_t dz = llabs(a[0] - b[0]), dy = llabs(a[1] - b[1]), dx = llabs(a[2] - b[2]);
  return dz <= md && dy <= md && dx <= md && dz * dz + dy * dy + dx * dx < r2;
}

// earlier candidates within the distance of each candidate: ncf[i] counts
// them (zeroed before), cf[i][0..PK_CF_MAX) lists up to PK_CF_MAX of them in
// any order (the resolve step reads them as a set, and scans all earlier
// rows itself when there are more).  Block (x, y) pairs rows 256 x .. 256 x
// + 255 with the earlier rows 64 y .. 64 y + 63 (held in LDS), so the
// triangle of pairs spreads over many blocks
__global__ __launch_bounds__(256) void k_pf_conflicts(const int64_t* __restrict__ xyz, int64_t n_host, int64_t md,
                                                      int* __restrict__ ncf, int* __restrict__ cf,
                                                      const int64_t* __restrict__ n_dev) {
  __shared__ int tz[64], ty[64], tx[64];
  const int64_t n = n_dev ? *n_dev : n_host;  // rows: the host count or one on the device (grid sized for the most)
  const int64_t i0 = (int64_t)blockIdx.x * 256, j0 = (int64_t)blockIdx.y * 64;
  if (j0 >= i0 + 255 || i0 >= n) return;  // no earlier row in this tile for any row of the block (uniform)
  const int t = threadIdx.x;
  if (t < 64 && j0 + t < n) {
    tz[t] = (int)xyz[3 * (j0 + t) + 0];
    ty[t] = (int)xyz[3 * (j0 + t) + 1];
    tx[t] = (int)xyz[3 * (j0 + t) + 2];
  }
  __syncthreads();
  const int64_t i = i0 + t;
  if (i >= n || i <= j0) return;
  const int cz = (int)xyz[3 * i + 0], cy = (int)xyz[3 * i + 1], cx = (int)xyz[3 * i + 2];
  const int64_t r2 = (md + 1) * (md + 1);
  const int jn = (int)(i - j0 < 64 ? i - j0 : 64);  // rows j0 .. i-1 of this tile
  for (int q = 0; q < jn; ++q) {
    const int64_t dz = llabs((int64_t)cz - tz[q]), dy = llabs((int64_t)cy - ty[q]), dx = llabs((int64_t)cx - tx[q]);
    if (dz <= md && dy <= md && dx <= md && dz * dz + dy * dy + dx * dx < r2) {
      const int k = atomicAdd(&ncf[i], 1);
      if (k < PK_CF_MAX) cf[i * PK_CF_MAX + k] = (int)(j0 + q);
    }
  }
}

__global__ void k_zero_i32(int* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// rounds of decisions in LDS (0 open, 1 kept, 2 dropped), then the kept rows
// in order, at most max_keep
__global__ __launch_bounds__(1024) void k_pf_resolve(const int64_t* __restrict__ xyz, int64_t n_host, int64_t md,
                                                     const int* __restrict__ ncf, const int* __restrict__ cf,
                                                     int64_t max_keep, int64_t* __restrict__ keep,
                                                     int64_t* __restrict__ nkeep, const int64_t* __restrict__ n_dev) {
  __shared__ unsigned char st[PK_PAR_MAX];
  const int64_t n = n_dev ? *n_dev : n_host;
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t r2 = (md + 1) * (md + 1);
  for (int64_t i = t; i < n; i += 1024) st[i] = ncf[i] == 0 ? 1 : 0;
  __syncthreads();
  for (int64_t round = 0; round <= n; ++round) {
    bool open = false;
    for (int64_t i = t; i < n; i += 1024) {
      if (st[i] != 0) continue;
      const int k = ncf[i];
      bool kept_nb = false, all_dropped = true;
      if (k <= PK_CF_MAX) {
        for (int q = 0; q < k; ++q) {
          const unsigned char v = st[cf[i * PK_CF_MAX + q]];
          kept_nb |= v == 1;
          all_dropped &= v == 2;
        }
      } else {
        for (int64_t j = 0; j < i && !kept_nb; ++j) {
          if (!pk_near(xyz + 3 * i, xyz + 3 * j, md, r2)) continue;
          const unsigned char v = st[j];
          kept_nb |= v == 1;
          all_dropped &= v == 2;
        }
      }
      if (kept_nb) st[i] = 2;
      else if (all_dropped) st[i] = 1;
      else open = true;
    }
    if (!__syncthreads_or(open)) break;
  }
  int64_t base = 0;
  for (int64_t b0 = 0; b0 < n && base < max_keep; b0 += 1024) {
    const int64_t i = b0 + t;
    const bool kp = i < n && st[i] == 1;
    const unsigned long long b = __ballot(kp);
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    int wb = 0, tot = 0;
    for (int q = 0; q < 16; ++q) {
      if (q < w) wb += wsum[q];
      tot += wsum[q];
    }
    const int64_t o = base + wb + __popcll(b & ((1ull << lane) - 1ull));
    if (kp && o < max_keep) keep[o] = i;
    base += tot;
    __syncthreads();
  }
  if (t == 0) *nkeep = base < max_keep ? base : max_keep;
}

// ---------------------------------------------------------------------------
// greedy distance filter: candidates in the given order; the kept list is
// spread over the threads (slot j belongs to thread j % 256), so the only
// cross-thread exchange per candidate is one barrier-or
__global__ __launch_bounds__(PK_THREADS) void k_peak_filter(const int64_t* __restrict__ xyz, int64_t n,
                                                            int64_t min_distance, int64_t max_keep,
                                                            int64_t* __restrict__ keep, int64_t* __restrict__ nkeep) {
  __shared__ int kept[PK_KEEP_MAX * 3];
  const int t = threadIdx.x;
  const int64_t md = min_distance, r2 = (md + 1) * (md + 1);
  int64_t K = 0;
  int64_t i = 0;
  int64_t pz = n > 0 ? xyz[0] : 0, py = n > 0 ? xyz[1] : 0, px = n > 0 ? xyz[2] : 0;
  for (; i < n && K < max_keep; ++i) {
    const int64_t cz = pz, cy = py, cx = px;
    if (i + 1 < n) {
      pz = xyz[3 * (i + 1) + 0];
      py = xyz[3 * (i + 1) + 1];
      px = xyz[3 * (i + 1) + 2];
    }
    bool hit = false;
    for (int64_t j = t; j < K; j += PK_THREADS) {
      const int64_t dz = llabs(cz - kept[3 * j + 0]), dy = llabs(cy - kept[3 * j + 1]),
                    dx = llabs(cx - kept[3 * j + 2]);
      if (dz <= md && dy <= md && dx <= md && dz * dz + dy * dy + dx * dx < r2) hit = true;
    }
    if (__syncthreads_or(hit)) continue;
    if (K % PK_THREADS == t) {
      kept[3 * K + 0] = (int)cz;
      kept[3 * K + 1] = (int)cy;
      kept[3 * K + 2] = (int)cx;
      keep[K] = i;
    }
    ++K;
  }
  if (t == 0) *nkeep = K;
}

// ---------------------------------------------------------------------------
// C-ABI

// ---------------------------------------------------------------------------
// one step of the peak caller on the device: the candidates of one score
// volume folded into the caller state (PeakCaller.__call__ with
// _get_peak_mask and _update, tme/analyzer/peaks.py:241-290, 327-384,
// 531-583), sized by counts that stay on the device

struct PkFold {
  int64_t n0, n1, n2, mbd;
  int has_min, has_max;
  float lo, hi;
};

__device__ __forceinline__ bool pk_window(float s, const PkFold& f) {
  return s == s && (!f.has_min || s >= f.lo) && (!f.has_max || s <= f.hi);
}

__device__ __forceinline__ bool pk_inside(int64_t z, int64_t y, int64_t x, const PkFold& f) {
  return z >= f.mbd && y >= f.mbd && x >= f.mbd && z < f.n0 - f.mbd && y < f.n1 - f.mbd && x < f.n2 - f.mbd;
}

__global__ void k_fold_init(int64_t* __restrict__ counts, int64_t* __restrict__ m) {
  if (threadIdx.x < 3) counts[threadIdx.x] = 0;
  if (threadIdx.x == 0) *m = 0;
}

// candidate validity (boundary distance, NaN, score window), counted into
// counts[0]; with key, the preselection key (score where valid, else 0)
__global__ __launch_bounds__(256) void k_fold_valid(const int64_t* __restrict__ cand, int64_t n,
                                                    const float* __restrict__ scores, PkFold f,
                                                    float* __restrict__ key, int64_t* __restrict__ counts) {
  __shared__ int wsum[4];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool v = false;
  if (i < n) {
    const int64_t z = cand[3 * i], y = cand[3 * i + 1], x = cand[3 * i + 2];
    const float sc = scores[(z * f.n1 + y) * f.n2 + x];
    v = pk_inside(z, y, x, f) && pk_window(sc, f);
    if (key) key[i] = v ? sc : 0.f;
  }
  const unsigned long long b = __ballot(v);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (tot) atomicAdd((unsigned long long*)counts, (unsigned long long)tot);
  }
}

// the new rows in candidate order: all valid candidates, or (presel) the
// preselected ones (sel, any order) that pass NaN and the score window -
// the reference re-applies only the score limits after preselecting
// (peaks.py:275-288); positions and scores compacted, count in counts[1]
__global__ __launch_bounds__(1024) void k_fold_new(const int64_t* __restrict__ cand, const int64_t* __restrict__ sel,
                                                   int L, int np, int presel, const float* __restrict__ scores,
                                                   PkFold f, int64_t* __restrict__ new_pos,
                                                   float* __restrict__ new_score, int64_t* __restrict__ counts) {
  extern __shared__ unsigned long long sk[];
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < np; i += 1024)
    sk[i] = i < L ? (unsigned long long)(presel ? sel[i] : (int64_t)i) : ~0ull;
  __syncthreads();
  if (presel) pk_bitonic(sk, np);
  int64_t base = 0;
  for (int b0 = 0; b0 < L; b0 += 1024) {
    const int r = b0 + t;
    bool keep = false;
    int64_t z = 0, y = 0, x = 0;
    float sc = 0.f;
    if (r < L) {
      const int64_t c = (int64_t)sk[r];
      z = cand[3 * c];
      y = cand[3 * c + 1];
      x = cand[3 * c + 2];
      sc = scores[(z * f.n1 + y) * f.n2 + x];
      keep = pk_window(sc, f) && (presel || pk_inside(z, y, x, f));
    }
    const unsigned long long b = __ballot(keep);
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    int wb = 0, tot = 0;
    for (int q = 0; q < 16; ++q) {
      if (q < w) wb += wsum[q];
      tot += wsum[q];
    }
    const int64_t o = base + wb + __popcll(b & ((1ull << lane) - 1ull));
    if (keep) {
      new_pos[3 * o] = z;
      new_pos[3 * o + 1] = y;
      new_pos[3 * o + 2] = x;
      new_score[o] = sc;
    }
    base += tot;
    __syncthreads();
  }
  if (t == 0) counts[1] = base;
}

__device__ __forceinline__ bool pk_filler(const int64_t* __restrict__ pos, const float* __restrict__ sc, int64_t c) {
  return pos[3 * c] == -1 && pos[3 * c + 1] == -1 && pos[3 * c + 2] == -1 && sc[c] == -1.f;
}

// state rows then new rows, sorted by score (descending, ties by lower row,
// NaN first); with dedupe every placeholder row after the first is left out
// (it lies at distance 0 from the first and shares its fate or is dropped).
// The live rows are compacted first, so the sort covers the next power of
// two of their count only.  ord[r] = row of rank r, ordpos its position, *m
// the row count
__global__ __launch_bounds__(1024) void k_fold_sort(const int64_t* __restrict__ st_pos,
                                                    const float* __restrict__ st_score, int64_t n_state, int dedupe,
                                                    const int64_t* __restrict__ new_pos,
                                                    const float* __restrict__ new_score,
                                                    const int64_t* __restrict__ counts, int np2, int* __restrict__ ord,
                                                    int64_t* __restrict__ ordpos, int64_t* __restrict__ m) {
  extern __shared__ unsigned long long sk[];
  __shared__ int first;
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t P = counts[1];
  const int rows = (int)(n_state + P);
  if (t == 0) first = 0x7fffffff;
  __syncthreads();
  if (dedupe)
    for (int64_t c = t; c < n_state; c += 1024)
      if (pk_filler(st_pos, st_score, c)) atomicMin(&first, (int)c);
  __syncthreads();
  int live = 0;
  for (int b0 = 0; b0 < rows; b0 += 1024) {
    const int i = b0 + t;
    bool in = false;
    unsigned long long v = 0ull;
    if (i < n_state) {
      in = !(dedupe && i != first && pk_filler(st_pos, st_score, i));
      v = ((unsigned long long)(~pk_key(st_score[i])) << 32) | (unsigned)i;
    } else if (i < rows) {
      in = true;
      v = ((unsigned long long)(~pk_key(new_score[i - n_state])) << 32) | (unsigned)i;
    }
    const unsigned long long b = __ballot(in);
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    int wb = 0, tot = 0;
    for (int q = 0; q < 16; ++q) {
      if (q < w) wb += wsum[q];
      tot += wsum[q];
    }
    if (in) sk[live + wb + __popcll(b & ((1ull << lane) - 1ull))] = v;
    live += tot;
    __syncthreads();
  }
  int np = 1;
  while (np < live) np <<= 1;
  for (int i = live + t; i < np; i += 1024) sk[i] = ~0ull;
  __syncthreads();
  pk_bitonic(sk, np);
  for (int r = t; r < live; r += 1024) {
    const int c = (int)(sk[r] & 0xffffffffull);
    ord[r] = c;
    const int64_t* src = c < n_state ? st_pos + 3 * c : new_pos + 3 * (c - n_state);
    ordpos[3 * r] = src[0];
    ordpos[3 * r + 1] = src[1];
    ordpos[3 * r + 2] = src[2];
  }
  if (t == 0) *m = live;
}

// output rows: rank keep[r] (or r without a distance filter) of the sorted
// rows, gathered from the state or the new rows (rotation = this volume's,
// details -1); counts[2] = output rows
__global__ __launch_bounds__(256) void k_fold_out(const int* __restrict__ ord, const int64_t* __restrict__ keep,
                                                  const int64_t* __restrict__ nkeep, const int64_t* __restrict__ m,
                                                  int64_t num_peaks, const int64_t* __restrict__ st_pos,
                                                  const float* __restrict__ st_rot,
                                                  const float* __restrict__ st_score,
                                                  const float* __restrict__ st_det, int64_t n_state,
                                                  const float* __restrict__ rot, const int64_t* __restrict__ new_pos,
                                                  const float* __restrict__ new_score, int64_t* __restrict__ out_pos,
                                                  float* __restrict__ out_rot, float* __restrict__ out_score,
                                                  float* __restrict__ out_det, int64_t* __restrict__ counts) {
  const int64_t nout = keep ? *nkeep : (*m < num_peaks ? *m : num_peaks);
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r == 0) counts[2] = nout;
  if (r >= nout) return;
  const int c = ord[keep ? keep[r] : r];
  if (c < n_state) {
    for (int k = 0; k < 3; ++k) out_pos[3 * r + k] = st_pos[3 * c + k];
    for (int k = 0; k < 9; ++k) out_rot[9 * r + k] = st_rot[9 * c + k];
    out_score[r] = st_score[c];
    out_det[r] = st_det[c];
  } else {
    const int j = c - (int)n_state;
    for (int k = 0; k < 3; ++k) out_pos[3 * r + k] = new_pos[3 * j + k];
    for (int k = 0; k < 9; ++k) out_rot[9 * r + k] = rot[k];
    out_score[r] = new_score[j];
    out_det[r] = -1.f;
  }
}

extern "C" {

// Peaks of a 3-D float32 volume under a box maximum filter of edge `size`
// (constant-zero borders), first half: filters, marks and counts.  Edges up
// to 32 (planes below 2^31 voxels, below 2^30 tiles): one pass with the x
// and y maxima in LDS and the z window in registers; larger edges: x and y
// maxima in one pass into tmp, then z maxima and the peak test.  Buffers: tmp F floats (may be
// null for edges up to 32), mask 64*nblk words, cnt nblk and off nblk+1
// int64 with nblk = ceil(F/4096).  off[nblk] receives the number of peaks.
int tm_peaks_max_filter(const float* scores, const int64_t* shape, int size, float* tmp,
                        unsigned long long* mask, int64_t* cnt, int64_t* off, void* stream) {
  TM_REQUIRE(scores && shape && mask && cnt && off, "tm_peaks_max_filter: null pointer");
  TM_REQUIRE(size >= 1 && size <= PK_MAX_SIZE, "tm_peaks_max_filter: size must be in [1, %d], got %d", PK_MAX_SIZE,
             size);
  const int64_t n0 = shape[0], n1 = shape[1], n2 = shape[2];
  // grid: x tiles of 64, y tiles of 16 (gridDim.y <= 65535), one z plane per
  // gridDim.z (<= 65535); coordinates stay far inside int32
  TM_REQUIRE(n0 > 0 && n1 > 0 && n2 > 0 && n0 < 65536 && n1 < (1 << 20) && n2 < (1 << 20),
             "tm_peaks_max_filter: shape must be positive with z < 65536 and y, x < 2^20");
  const int64_t F = n0 * n1 * n2, nblk = tm_ceil_div(F, PK_MF_CHUNK), plane = n1 * n2;
  hipStream_t s = (hipStream_t)stream;
  // x tiles of at most 64 columns that split the rows evenly (200 -> 4 x 50)
  const int ntx = (int)tm_ceil_div(n2, PK_MF_TX);
  const int tx = (int)tm_ceil_div(n2, ntx);
  const int nty = (int)tm_ceil_div(n1, PK_MF_TY);
  const int64_t nb = (int64_t)ntx * nty * tm_ceil_div(n0, PK_ZC);  // blocks of the one-pass filter
  if (size <= PK_ZCOL_MAX && plane < (1ll << 31) && nb < (1ll << 30)) {
    const int64_t words = 64 * nblk;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<int64_t>(tm_ceil_div(words, 256), 4096)), dim3(256), 0, s,
                       mask, words);
    TM_CHECK_LAUNCH();
    const dim3 g((unsigned)(tm_ceil_div(nb, 8) * 8));  // a multiple of 8 for the XCD grouping
#define PK_FUSED(D_)                                                                                       \
  case D_:                                                                                                 \
    TM_TIMED("peaks_max_filter", 4.0 * F + F / 8.0, s,                                                     \
             hipLaunchKernelGGL((k_mf_fused<D_, (D_ <= 8 ? 4 : D_ <= 16 ? 8 : 16)>), g,                    \
                                dim3(D_ <= 8 ? 256 : D_ <= 16 ? 512 : 1024), 0, s, scores, mask, (int)n0,  \
                                (int)n1, (int)n2, tx, ntx, nty, nb));                                      \
    break;
    switch (size) {
      PK_FUSED(1) PK_FUSED(2) PK_FUSED(3) PK_FUSED(4) PK_FUSED(5) PK_FUSED(6) PK_FUSED(7) PK_FUSED(8)
      PK_FUSED(9) PK_FUSED(10) PK_FUSED(11) PK_FUSED(12) PK_FUSED(13) PK_FUSED(14) PK_FUSED(15) PK_FUSED(16)
      PK_FUSED(17) PK_FUSED(18) PK_FUSED(19) PK_FUSED(20) PK_FUSED(21) PK_FUSED(22) PK_FUSED(23) PK_FUSED(24)
      PK_FUSED(25) PK_FUSED(26) PK_FUSED(27) PK_FUSED(28) PK_FUSED(29) PK_FUSED(30) PK_FUSED(31) PK_FUSED(32)
    }
#undef PK_FUSED
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_mask_count, dim3((unsigned)nblk), dim3(64), 0, s, mask, cnt);
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pk_scan, dim3(1), dim3(1024), 0, s, cnt, off, nblk);
    TM_CHECK_LAUNCH();
    return TM_OK;
  }
  TM_REQUIRE(tmp, "tm_peaks_max_filter: null tmp (needed for edges above %d)", PK_ZCOL_MAX);
  const int hx = tx + size - 1, hy = PK_MF_TY + size - 1;
  const size_t shm = sizeof(float) * (size_t)hy * (hx + tx);
  dim3 g1((unsigned)ntx, (unsigned)tm_ceil_div(n1, PK_MF_TY), (unsigned)tm_ceil_div(n0, PK_MF_TZ));
  TM_TIMED("peaks_max_filter_xy", 8.0 * F, s,
           hipLaunchKernelGGL(k_mf_xy, g1, dim3(256), shm, s, scores, tmp, (int)n0, (int)n1, (int)n2, size, tx));
  TM_CHECK_LAUNCH();
  TM_TIMED("peaks_max_filter_z", 8.0 * F + F / 8.0, s,
           hipLaunchKernelGGL(k_mf_z, dim3((unsigned)nblk), dim3(PK_THREADS), 0, s, scores, tmp, mask, cnt, (int)n0,
                              (int)n1, (int)n2, size));
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_pk_scan, dim3(1), dim3(1024), 0, s, cnt, off, nblk);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// Second half: writes the off[nblk] peak coordinates (int64 z, y, x rows) in
// C order.
int tm_peaks_write(const unsigned long long* mask, const int64_t* off, const int64_t* shape, int64_t* coords,
                   void* stream) {
  TM_REQUIRE(mask && off && shape && coords, "tm_peaks_write: null pointer");
  const int64_t F = shape[0] * shape[1] * shape[2], nblk = tm_ceil_div(F, PK_MF_CHUNK);
  hipStream_t s = (hipStream_t)stream;
  TM_TIMED("peaks_write", F / 8.0, s,
           hipLaunchKernelGGL(k_mf_write, dim3((unsigned)std::min<int64_t>(nblk, PK_WRITE_BLOCKS)),
                              dim3(PK_THREADS), 0, s, mask, off, coords,
                              (int)shape[1], (int)shape[2], F));
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// Workspace (bytes) of tm_topk_f32 for n values.
int64_t tm_topk_workspace_bytes(int64_t n, int64_t k) {
  const int64_t nb = tm_ceil_div(n > 0 ? n : 1, PK_TOPK_CHUNK);
  int64_t np2 = 1;
  while (np2 < k) np2 <<= 1;
  return 256 + 4 * 256 * 4 + 4 * 8 * (nb + 1) + 8 * np2;
}

// Indices of the k largest of n float32 values, largest first, ties by lower
// index, NaN above every number.  k <= 65536, n < 2^32.
int tm_topk_f32(const float* v, int64_t n, int64_t k, int64_t* out, void* work, int64_t work_bytes, void* stream) {
  TM_REQUIRE(v && out && work, "tm_topk_f32: null pointer");
  TM_REQUIRE(n > 0 && n < (1ll << 32), "tm_topk_f32: n must be in [1, 2^32), got %lld", (long long)n);
  TM_REQUIRE(k >= 1 && k <= n && k <= PK_SORT_MAX, "tm_topk_f32: k must be in [1, min(n, %d)], got %lld",
             PK_SORT_MAX, (long long)k);
  TM_REQUIRE(work_bytes >= tm_topk_workspace_bytes(n, k), "tm_topk_f32: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (n <= PK_SORT_LDS) {
    int np = 1;
    while (np < n) np <<= 1;
    hipLaunchKernelGGL(k_topk_small, dim3(1), dim3(1024), sizeof(unsigned long long) * np, s, v, (int)n, np, (int)k,
                       out);
    TM_CHECK_LAUNCH();
    return TM_OK;
  }
  const int64_t nb = tm_ceil_div(n, PK_TOPK_CHUNK);
  int np2 = 1;
  while (np2 < k) np2 <<= 1;
  char* w = (char*)work;
  pk_sel_state* st = (pk_sel_state*)w;
  unsigned* hist = (unsigned*)(w + 256);
  int64_t* cgt = (int64_t*)(w + 256 + 4 * 256 * 4);
  int64_t* ceq = cgt + (nb + 1);
  int64_t* ogt = ceq + (nb + 1);
  int64_t* oeq = ogt + (nb + 1);
  unsigned long long* sel = (unsigned long long*)(oeq + (nb + 1));
  hipLaunchKernelGGL(k_topk_init, dim3(1), dim3(256), 0, s, st, hist, k);
  TM_CHECK_LAUNCH();
  unsigned hb = (unsigned)std::min<int64_t>(tm_ceil_div(n, PK_THREADS * 8), 2048);
  for (int p = 0; p < 4; ++p) {
    const int shift = 24 - 8 * p;
    TM_TIMED("topk_hist", 4.0 * n, s,
             hipLaunchKernelGGL(k_topk_hist, dim3(hb), dim3(PK_THREADS), 0, s, v, n, st, hist + 256 * p, shift));
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_topk_pick, dim3(1), dim3(256), 0, s, st, hist + 256 * p, shift);
    TM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_topk_count, dim3((unsigned)nb), dim3(PK_THREADS), 0, s, v, n, st, cgt, ceq);
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_pk_scan, dim3(1), dim3(1024), 0, s, cgt, ogt, nb);
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_pk_scan, dim3(1), dim3(1024), 0, s, ceq, oeq, nb);
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_topk_write, dim3((unsigned)nb), dim3(PK_THREADS), 0, s, v, n, st, ogt, oeq, sel);
  TM_CHECK_LAUNCH();
  if (np2 <= PK_SORT_LDS) {
    hipLaunchKernelGGL(k_topk_sort<true>, dim3(1), dim3(1024), sizeof(unsigned long long) * np2, s, sel, k, np2,
                       out);
  } else {
    hipLaunchKernelGGL(k_topk_sort<false>, dim3(1), dim3(1024), 0, s, sel, k, np2, out);
  }
  TM_CHECK_LAUNCH();
  return TM_OK;
}

// Workspace (bytes) of tm_peak_filter for n rows.
int64_t tm_peak_filter_workspace_bytes(int64_t n) { return (n > 0 ? n : 1) * 4 * (1 + PK_CF_MAX); }

// Greedy minimum-distance filter over n int64 (z, y, x) rows taken in the
// given order: writes the kept row indices to keep[0..*nkeep), at most
// max_keep of them.  Up to 16384 rows are resolved in parallel (any
// max_keep); beyond that a one-block sequential pass keeps at most 4096.
// nkeep is a device int64.
int tm_peak_filter(const int64_t* coords, int64_t n, int64_t min_distance, int64_t max_keep, int64_t* keep,
                   int64_t* nkeep, void* work, int64_t work_bytes, void* stream) {
  TM_REQUIRE(coords && keep && nkeep, "tm_peak_filter: null pointer");
  TM_REQUIRE(n >= 1, "tm_peak_filter: needs at least one row");
  TM_REQUIRE(min_distance >= 0 && min_distance < (1 << 20), "tm_peak_filter: min_distance out of range");
  TM_REQUIRE(max_keep >= 1, "tm_peak_filter: max_keep must be positive");
  hipStream_t s = (hipStream_t)stream;
  if (n <= PK_PAR_MAX) {
    TM_REQUIRE(work && work_bytes >= tm_peak_filter_workspace_bytes(n), "tm_peak_filter: workspace too small");
    int* ncf = (int*)work;
    int* cf = ncf + n;
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)std::min<int64_t>(tm_ceil_div(n, 256), 1024)), dim3(256), 0, s, ncf,
                       n);
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pf_conflicts, dim3((unsigned)tm_ceil_div(n, 256), (unsigned)tm_ceil_div(n, 64)), dim3(256), 0,
                       s, coords, n, min_distance, ncf, cf, (const int64_t*)nullptr);
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pf_resolve, dim3(1), dim3(1024), 0, s, coords, n, min_distance, ncf, cf, max_keep, keep,
                       nkeep, (const int64_t*)nullptr);
    TM_CHECK_LAUNCH();
    return TM_OK;
  }
  TM_REQUIRE(max_keep <= PK_KEEP_MAX, "tm_peak_filter: beyond %d rows at most %d can be kept", PK_PAR_MAX,
             PK_KEEP_MAX);
  hipLaunchKernelGGL(k_peak_filter, dim3(1), dim3(PK_THREADS), 0, s, coords, n, min_distance, max_keep, keep, nkeep);
  TM_CHECK_LAUNCH();
  return TM_OK;
}


// rows of the native caller step: state rows + new rows (<= 2 num_peaks)
#define PK_FOLD_MAX PK_SORT_LDS

static int64_t pk_align(int64_t b) { return (b + 255) / 256 * 256; }

// Workspace (bytes) of tm_peak_caller_fold.
int64_t tm_peak_caller_fold_workspace_bytes(int64_t n_cand, int64_t n_state, int64_t num_peaks) {
  if (n_cand < 1) n_cand = 1;
  const bool presel = n_cand > num_peaks;
  const int64_t kk = presel ? std::min<int64_t>(n_cand, 2 * num_peaks) : n_cand;
  const int64_t M = n_state + kk;
  int64_t b = 0;
  if (presel) b += pk_align(4 * n_cand) + pk_align(8 * kk) + pk_align(tm_topk_workspace_bytes(n_cand, kk));
  b += pk_align(24 * kk) + pk_align(4 * kk);              // new rows
  b += pk_align(4 * M) + pk_align(24 * M) + pk_align(8);  // sorted rows, count
  b += pk_align(4 * M * (1 + PK_CF_MAX)) + pk_align(8 * num_peaks) + pk_align(8);  // distance filter
  return b;
}

// One step of PeakCaller.__call__ after call_peaks (the reference's
// _get_peak_mask and _update, tme/analyzer/peaks.py:241-290, 531-583): the
// n_cand candidate positions (int64 [n][3], C order) of the score volume
// (float32, shape[3]) are checked against the boundary distance and the
// score window, preselected to the 2 num_peaks highest when there are more
// than num_peaks, merged with the state (n_state rows: int64 [n][3]
// positions, float32 [n][3][3] rotations, [n] scores and details), sorted
// by score and distance-filtered in that order (trunc(sqrt(d2)) <=
// min_distance drops a row); the first num_peaks rows go to out_* (row
// capacity num_peaks).  counts (device int64 [3]): valid candidates (0: the
// state is to be left unchanged), new rows, output rows.  Requires
// n_state + min(n_cand, 2 num_peaks when n_cand > num_peaks) <= 8192.
int tm_peak_caller_fold(const int64_t* cand, int64_t n_cand, const float* scores, const int64_t* shape,
                        const int64_t* st_pos, const float* st_rot, const float* st_score, const float* st_det,
                        int64_t n_state, const float* rotation, int64_t num_peaks, int64_t min_distance,
                        int64_t min_boundary_distance, int has_min, float min_score, int has_max, float max_score,
                        int64_t* out_pos, float* out_rot, float* out_score, float* out_det, int64_t* counts,
                        void* work, int64_t work_bytes, void* stream) {
  TM_REQUIRE(cand && scores && shape && rotation && out_pos && out_rot && out_score && out_det && counts && work,
             "tm_peak_caller_fold: null pointer");
  TM_REQUIRE(n_cand >= 1, "tm_peak_caller_fold: needs at least one candidate");
  TM_REQUIRE(num_peaks >= 1, "tm_peak_caller_fold: num_peaks must be positive");
  TM_REQUIRE(n_state >= 0 && (n_state == 0 || (st_pos && st_rot && st_score && st_det)),
             "tm_peak_caller_fold: bad state");
  TM_REQUIRE(min_distance < (1 << 20) && min_boundary_distance >= 0, "tm_peak_caller_fold: distance out of range");
  const bool presel = n_cand > num_peaks;
  const int64_t kk = presel ? std::min<int64_t>(n_cand, 2 * num_peaks) : n_cand;
  const int64_t M = n_state + kk;
  TM_REQUIRE(M <= PK_FOLD_MAX, "tm_peak_caller_fold: %lld state rows + %lld new rows exceed %d", (long long)n_state,
             (long long)kk, PK_FOLD_MAX);
  TM_REQUIRE(work_bytes >= tm_peak_caller_fold_workspace_bytes(n_cand, n_state, num_peaks),
             "tm_peak_caller_fold: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  PkFold f;
  f.n0 = shape[0];
  f.n1 = shape[1];
  f.n2 = shape[2];
  f.mbd = min_boundary_distance;
  f.has_min = has_min;
  f.has_max = has_max;
  f.lo = min_score;
  f.hi = max_score;
  char* w = (char*)work;
  auto take = [&](int64_t bytes) {
    char* p = w;
    w += pk_align(bytes);
    return p;
  };
  float* key = nullptr;
  int64_t* sel = nullptr;
  char* tws = nullptr;
  const int64_t tw_bytes = presel ? tm_topk_workspace_bytes(n_cand, kk) : 0;
  if (presel) {
    key = (float*)take(4 * n_cand);
    sel = (int64_t*)take(8 * kk);
    tws = take(tw_bytes);
  }
  int64_t* new_pos = (int64_t*)take(24 * kk);
  float* new_score = (float*)take(4 * kk);
  int* ord = (int*)take(4 * M);
  int64_t* ordpos = (int64_t*)take(24 * M);
  int64_t* m = (int64_t*)take(8);
  int* ncf = (int*)take(4 * M * (1 + PK_CF_MAX));
  int* cf = ncf + M;
  int64_t* keep = (int64_t*)take(8 * num_peaks);
  int64_t* nkeep = (int64_t*)take(8);

  hipLaunchKernelGGL(k_fold_init, dim3(1), dim3(64), 0, s, counts, m);
  TM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_fold_valid, dim3((unsigned)tm_ceil_div(n_cand, 256)), dim3(256), 0, s, cand, n_cand, scores, f,
                     key, counts);
  TM_CHECK_LAUNCH();
  if (presel) {
    const int rc = tm_topk_f32(key, n_cand, kk, sel, tws, tw_bytes, stream);
    if (rc != TM_OK) return rc;
  }
  int np = 1;
  while (np < kk) np <<= 1;
  hipLaunchKernelGGL(k_fold_new, dim3(1), dim3(1024), sizeof(unsigned long long) * np, s, cand, sel, (int)kk, np,
                     presel ? 1 : 0, scores, f, new_pos, new_score, counts);
  TM_CHECK_LAUNCH();
  int np2 = 1;
  while (np2 < M) np2 <<= 1;
  const int dedupe = min_distance > 0 && n_state > 1;
  TM_TIMED("peaks_fold_sort", 40.0 * M, s,
           hipLaunchKernelGGL(k_fold_sort, dim3(1), dim3(1024), sizeof(unsigned long long) * np2, s, st_pos, st_score,
                              n_state, dedupe, new_pos, new_score, counts, np2, ord, ordpos, m));
  TM_CHECK_LAUNCH();
  const bool filter = min_distance > 0;
  if (filter) {
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)tm_ceil_div(M, 256)), dim3(256), 0, s, ncf, M);
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pf_conflicts, dim3((unsigned)tm_ceil_div(M, 256), (unsigned)tm_ceil_div(M, 64)), dim3(256), 0,
                       s, ordpos, M, min_distance, ncf, cf, (const int64_t*)m);
    TM_CHECK_LAUNCH();
    TM_TIMED("peaks_fold_filter", 60.0 * M, s,
             hipLaunchKernelGGL(k_pf_resolve, dim3(1), dim3(1024), 0, s, ordpos, M, min_distance, ncf, cf, num_peaks,
                                keep, nkeep, (const int64_t*)m));
    TM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_fold_out, dim3((unsigned)tm_ceil_div(num_peaks, 256)), dim3(256), 0, s, ord,
                     filter ? keep : (const int64_t*)nullptr, nkeep, m, num_peaks, st_pos, st_rot, st_score, st_det,
                     n_state, rotation, new_pos, new_score, out_pos, out_rot, out_score, out_det, counts);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

}  // extern "C"

// Scratch of tm_peak_caller_fold_noalloc: a file-scope device array, one
// copy per GPU (each at its own address), held until the process ends.
// Every call on a GPU shares that GPU's copy and leaves its contents behind;
// the fold's own kernels clear what must start at zero (counts, conflict
// lists, the top-k selection state), so no call depends on an earlier one.
// Calls that overlap on one GPU race on it.
#define PK_NOALLOC_SCRATCH_BYTES (8ll << 20)
#define PK_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) char scratch[PK_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[PK_NOALLOC_MAX_DEVICES];

// Device address of this GPU's scratch copy, looked up once per device (with
// that device current): host code uses this address, never the array's name.
static int pk_scratch_address(void** out) {
  int dev = -1;
  TM_CHECK_HIP(hipGetDevice(&dev));
  TM_REQUIRE(dev >= 0 && dev < PK_NOALLOC_MAX_DEVICES, "device %d beyond the %d devices the scratch covers", dev,
             PK_NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TM_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TM_OK;
}

extern "C" {

int64_t tm_peak_caller_fold_noalloc_max_bytes(void) { return PK_NOALLOC_SCRATCH_BYTES; }

// tm_peak_caller_fold without a workspace argument and without any device
// allocation: the workspace is the library's static per-device scratch of
// PK_NOALLOC_SCRATCH_BYTES.  Arguments, results and errors are those of
// tm_peak_caller_fold; a step whose workspace
// (tm_peak_caller_fold_workspace_bytes) exceeds the scratch returns
// TM_ERR_VALUE before anything is launched.
int tm_peak_caller_fold_noalloc(const int64_t* cand, int64_t n_cand, const float* scores, const int64_t* shape,
                                const int64_t* st_pos, const float* st_rot, const float* st_score,
                                const float* st_det, int64_t n_state, const float* rotation, int64_t num_peaks,
                                int64_t min_distance, int64_t min_boundary_distance, int has_min, float min_score,
                                int has_max, float max_score, int64_t* out_pos, float* out_rot, float* out_score,
                                float* out_det, int64_t* counts, void* stream) {
  TM_REQUIRE(n_cand >= 1 && n_state >= 0 && num_peaks >= 1, "tm_peak_caller_fold_noalloc: bad sizes");
  const int64_t need = tm_peak_caller_fold_workspace_bytes(n_cand, n_state, num_peaks);
  TM_REQUIRE(need <= PK_NOALLOC_SCRATCH_BYTES,
             "tm_peak_caller_fold_noalloc: %lld candidates, %lld state rows and %lld peaks need %lld bytes of "
             "scratch, more than its %lld; use tm_peak_caller_fold with a workspace",
             (long long)n_cand, (long long)n_state, (long long)num_peaks, (long long)need,
             (long long)PK_NOALLOC_SCRATCH_BYTES);
  void* work = nullptr;
  const int rc = pk_scratch_address(&work);
  if (rc != TM_OK) return rc;
  return tm_peak_caller_fold(cand, n_cand, scores, shape, st_pos, st_rot, st_score, st_det, n_state, rotation,
                             num_peaks, min_distance, min_boundary_distance, has_min, min_score, has_max, max_score,
                             out_pos, out_rot, out_score, out_det, counts, work, PK_NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"
