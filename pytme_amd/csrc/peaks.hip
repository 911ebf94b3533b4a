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

#include "tm_common.h"

#define PK_THREADS 256                 // threads per compaction block (4 waves)
#define PK_MF_TX 64                    // max-filter plane tile: 64 x 16
#define PK_MF_TY 16
#define PK_MF_TZ 8                     // z planes per block of the xy pass
#define PK_MAX_SIZE 64                 // largest filter edge
#define PK_SORT_LDS 8192               // bitonic sort in LDS up to this many keys
#define PK_SORT_MAX 65536              // larger sorts run on a global buffer
#define PK_KEEP_MAX 4096               // kept points of the distance filter
#define PK_TOPK_ITEMS 16               // elements per thread in the top-k scans
#define PK_MF_ITEMS 16                 // voxels per thread in the peak test and write
#define PK_MF_CHUNK (PK_THREADS * PK_MF_ITEMS)
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

// x and y maxima over a 64 x 16 plane tile with a (d-1) halo; outside the
// volume the input is the constant 0
__global__ __launch_bounds__(256) void k_mf_xy(const float* __restrict__ s, float* __restrict__ out, int n0, int n1,
                                               int n2, int d) {
  extern __shared__ float sm[];
  const int lo = d / 2;
  const int hx = PK_MF_TX + d - 1, hy = PK_MF_TY + d - 1;
  float* in = sm;             // hy x hx
  float* mx = sm + hy * hx;   // hy x 64
  const int x0 = blockIdx.x * PK_MF_TX, y0 = blockIdx.y * PK_MF_TY;
  const int zb = blockIdx.z * PK_MF_TZ, ze = min(zb + PK_MF_TZ, n0);
  for (int64_t z = zb; z < ze; ++z) {
  __syncthreads();  // the previous plane's tiles are consumed
  const float* plane = s + z * (int64_t)n1 * n2;
  for (int i = threadIdx.x; i < hy * hx; i += blockDim.x) {
    const int r = i / hx, c = i - r * hx;
    const int y = y0 - lo + r, x = x0 - lo + c;
    in[i] = (y >= 0 && y < n1 && x >= 0 && x < n2) ? plane[(int64_t)y * n2 + x] : 0.f;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < hy * PK_MF_TX; i += blockDim.x) {
    const int r = i / PK_MF_TX, c = i - r * PK_MF_TX;
    const float* row = in + r * hx + c;
    float m = row[0];
    for (int k = 1; k < d; ++k) m = fmaxf(m, row[k]);
    mx[i] = m;
  }
  __syncthreads();
  float* oplane = out + z * (int64_t)n1 * n2;
  for (int i = threadIdx.x; i < PK_MF_TY * PK_MF_TX; i += blockDim.x) {
    const int r = i / PK_MF_TX, c = i - r * PK_MF_TX;
    const int y = y0 + r, x = x0 + c;
    if (y >= n1 || x >= n2) continue;
    const float* col = mx + r * PK_MF_TX + c;
    float m = col[0];
    for (int k = 1; k < d; ++k) m = fmaxf(m, col[k * PK_MF_TX]);
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

// z maxima and the peak test for edges up to 8: one thread per (y, x)
// column walks a chunk of 64 z planes (plus the halo planes) holding the
// last D xy maxima in registers, so every xy maximum is read about once and
// consecutive threads read consecutive voxels.  A wave's 64 columns are 64
// consecutive voxels of a plane; its ballot is OR-ed into the linear-order
// words (zeroed beforehand), split across a word boundary when needed.
#define PK_ZC 64
#define PK_ZCOL_MAX 8

template <int D>
__global__ __launch_bounds__(256) void k_mf_zcol(const float* __restrict__ s, const float* __restrict__ mxy,
                                                 unsigned long long* __restrict__ mask, int n0, int64_t plane) {
  constexpr int LO = D / 2, HI = D - 1 - LO;
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool col = q < plane;
  const int z0 = blockIdx.y * PK_ZC, zend = min(z0 + PK_ZC, n0);
  float w[D];
#pragma unroll
  for (int k = 0; k < D; ++k) w[k] = 0.f;
  // window slots hold planes zi - D + 1 .. zi; planes outside the volume are 0
  for (int zi = z0 - LO; zi <= zend - 1 + HI; ++zi) {
#pragma unroll
    for (int k = 0; k < D - 1; ++k) w[k] = w[k + 1];
    w[D - 1] = (col && zi >= 0 && zi < n0) ? mxy[(int64_t)zi * plane + q] : 0.f;
    const int zo = zi - HI;
    if (zo < z0) continue;  // uniform
    float m = w[0];
#pragma unroll
    for (int k = 1; k < D; ++k) m = fmaxf(m, w[k]);
    const bool peak = col && s[(int64_t)zo * plane + q] == m;
    const unsigned long long b = __ballot(peak);
    if (lane == 0 && b) {
      const int64_t L0 = (int64_t)zo * plane + (q - lane);
      const int64_t wd = L0 >> 6;
      const int sh = (int)(L0 & 63);
      atomicOr(&mask[wd], b << sh);
      if (sh && (b >> (64 - sh))) atomicOr(&mask[wd + 1], b >> (64 - sh));
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

// C-ordered coordinates of the set bits
__global__ __launch_bounds__(PK_THREADS) void k_mf_write(const unsigned long long* __restrict__ mask,
                                                         const int64_t* __restrict__ off, int64_t* __restrict__ xyz,
                                                         int n1, int n2, int64_t F) {
  __shared__ int wsum[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t plane = (int64_t)n1 * n2;
  int64_t base = off[blockIdx.x];
  for (int it = 0; it < PK_MF_ITEMS; ++it) {
    const int64_t L0 = (int64_t)blockIdx.x * PK_MF_CHUNK + it * PK_THREADS;
    const unsigned long long b = L0 < F ? mask[L0 / 64 + w] : 0ull;
    if (lane == 0) wsum[w] = __popcll(b);
    __syncthreads();
    int wb = 0;
    for (int i = 0; i < w; ++i) wb += wsum[i];
    const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    const int64_t L = L0 + threadIdx.x;
    if (((b >> lane) & 1ull) && L < F) {
      const int64_t o = base + wb + __popcll(b & ((1ull << lane) - 1ull));
      const int64_t z = L / plane, r = L - z * plane, y = r / n2, x = r - y * n2;
      xyz[3 * o + 0] = z;
      xyz[3 * o + 1] = y;
      xyz[3 * o + 2] = x;
    }
    base += tot;
    __syncthreads();
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
  for (int k = 2; k <= np2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
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
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = (int64_t)(a[i] & 0xffffffffull);
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

// earlier candidates within the distance of each candidate (first PK_CF_MAX
// listed, all counted); earlier rows stream through LDS in tiles of 256
__global__ __launch_bounds__(256) void k_pf_conflicts(const int64_t* __restrict__ xyz, int64_t n, int64_t md,
                                                      int* __restrict__ ncf, int* __restrict__ cf) {
  __shared__ int tz[256], ty[256], tx[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool valid = i < n;
  const int64_t r2 = (md + 1) * (md + 1);
  const int cz = valid ? (int)xyz[3 * i + 0] : 0, cy = valid ? (int)xyz[3 * i + 1] : 0,
            cx = valid ? (int)xyz[3 * i + 2] : 0;
  int k = 0;
  const int64_t last = (int64_t)blockIdx.x * 256;  // the block's first row: earlier tiles are whole
  for (int64_t j0 = 0; j0 <= last; j0 += 256) {
    const int64_t j = j0 + threadIdx.x;
    if (j < n) {
      tz[threadIdx.x] = (int)xyz[3 * j + 0];
      ty[threadIdx.x] = (int)xyz[3 * j + 1];
      tx[threadIdx.x] = (int)xyz[3 * j + 2];
    }
    __syncthreads();
    if (valid) {
      const int64_t jn = i - j0 < 256 ? i - j0 : 256;  // rows j0 .. i-1 of this tile
      for (int64_t q = 0; q < jn; ++q) {
        const int64_t dz = llabs((int64_t)cz - tz[q]), dy = llabs((int64_t)cy - ty[q]),
                      dx = llabs((int64_t)cx - tx[q]);
        if (dz <= md && dy <= md && dx <= md && dz * dz + dy * dy + dx * dx < r2) {
          if (k < PK_CF_MAX) cf[i * PK_CF_MAX + k] = (int)(j0 + q);
          ++k;
        }
      }
    }
    __syncthreads();
  }
  if (valid) ncf[i] = k;
}

// rounds of decisions in LDS (0 open, 1 kept, 2 dropped), then the kept rows
// in order, at most max_keep
__global__ __launch_bounds__(1024) void k_pf_resolve(const int64_t* __restrict__ xyz, int64_t n, int64_t md,
                                                     const int* __restrict__ ncf, const int* __restrict__ cf,
                                                     int64_t max_keep, int64_t* __restrict__ keep,
                                                     int64_t* __restrict__ nkeep) {
  __shared__ unsigned char st[PK_PAR_MAX];
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

extern "C" {

// Peaks of a 3-D float32 volume under a box maximum filter of edge `size`
// (constant-zero borders), first half: filters, marks and counts (separable:
// x and y maxima in one pass, then z maxima and the peak test; for edges up
// to 8 the z pass walks columns with the window in registers).  Buffers:
// tmp F floats, mask 64*nblk words, cnt nblk and off nblk+1 int64 with
// nblk = ceil(F/4096).  off[nblk] receives the number of peaks.
int tm_peaks_max_filter(const float* scores, const int64_t* shape, int size, float* tmp,
                        unsigned long long* mask, int64_t* cnt, int64_t* off, void* stream) {
  TM_REQUIRE(scores && shape && tmp && mask && cnt && off, "tm_peaks_max_filter: null pointer");
  TM_REQUIRE(size >= 1 && size <= PK_MAX_SIZE, "tm_peaks_max_filter: size must be in [1, %d], got %d", PK_MAX_SIZE,
             size);
  const int64_t n0 = shape[0], n1 = shape[1], n2 = shape[2];
  // grid: x tiles of 64, y tiles of 16 (gridDim.y <= 65535), one z plane per
  // gridDim.z (<= 65535); coordinates stay far inside int32
  TM_REQUIRE(n0 > 0 && n1 > 0 && n2 > 0 && n0 < 65536 && n1 < (1 << 20) && n2 < (1 << 20),
             "tm_peaks_max_filter: shape must be positive with z < 65536 and y, x < 2^20");
  const int64_t F = n0 * n1 * n2, nblk = tm_ceil_div(F, PK_MF_CHUNK);
  hipStream_t s = (hipStream_t)stream;
  const int hx = PK_MF_TX + size - 1, hy = PK_MF_TY + size - 1;
  const size_t shm = sizeof(float) * (size_t)hy * (hx + PK_MF_TX);
  dim3 g1((unsigned)tm_ceil_div(n2, PK_MF_TX), (unsigned)tm_ceil_div(n1, PK_MF_TY), (unsigned)tm_ceil_div(n0, PK_MF_TZ));
  TM_TIMED("peaks_max_filter_xy", 8.0 * F, s,
           hipLaunchKernelGGL(k_mf_xy, g1, dim3(256), shm, s, scores, tmp, (int)n0, (int)n1, (int)n2, size));
  TM_CHECK_LAUNCH();
  if (size <= PK_ZCOL_MAX) {
    const int64_t words = 64 * nblk, plane = n1 * n2;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)std::min<int64_t>(tm_ceil_div(words, 256), 4096)), dim3(256), 0, s,
                       mask, words);
    TM_CHECK_LAUNCH();
    dim3 gz((unsigned)tm_ceil_div(plane, 256), (unsigned)tm_ceil_div(n0, PK_ZC));
#define PK_ZCOL(D_)                                                                                       \
  case D_:                                                                                                \
    TM_TIMED("peaks_max_filter_z", 8.0 * F + F / 8.0, s,                                                  \
             hipLaunchKernelGGL(k_mf_zcol<D_>, gz, dim3(256), 0, s, scores, tmp, mask, (int)n0, plane)); \
    break;
    switch (size) {
      PK_ZCOL(1) PK_ZCOL(2) PK_ZCOL(3) PK_ZCOL(4) PK_ZCOL(5) PK_ZCOL(6) PK_ZCOL(7) PK_ZCOL(8)
    }
#undef PK_ZCOL
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_mask_count, dim3((unsigned)nblk), dim3(64), 0, s, mask, cnt);
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pk_scan, dim3(1), dim3(1024), 0, s, cnt, off, nblk);
    TM_CHECK_LAUNCH();
    return TM_OK;
  }
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
           hipLaunchKernelGGL(k_mf_write, dim3((unsigned)nblk), dim3(PK_THREADS), 0, s, mask, off, coords,
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
    hipLaunchKernelGGL(k_pf_conflicts, dim3((unsigned)tm_ceil_div(n, 256)), dim3(256), 0, s, coords, n,
                       min_distance, ncf, cf);
    TM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_pf_resolve, dim3(1), dim3(1024), 0, s, coords, n, min_distance, ncf, cf, max_keep, keep,
                       nkeep);
    TM_CHECK_LAUNCH();
    return TM_OK;
  }
  TM_REQUIRE(max_keep <= PK_KEEP_MAX, "tm_peak_filter: beyond %d rows at most %d can be kept", PK_PAR_MAX,
             PK_KEEP_MAX);
  hipLaunchKernelGGL(k_peak_filter, dim3(1), dim3(PK_THREADS), 0, s, coords, n, min_distance, max_keep, keep, nkeep);
  TM_CHECK_LAUNCH();
  return TM_OK;
}

}  // extern "C"
