// Compile-time specialised Stockham FFT for a group of L lines of length N.
//
// The line length, line count, thread count and radix sequence are template
// constants, so every index, stride and loop bound folds at compile time.
// Data enter and leave in registers; LDS is used only to exchange values
// between radix stages (one in-place buffer: each stage reads all of its
// inputs into registers, synchronises, then writes its outputs).  An FFT
// with S radix stages therefore touches LDS 2(S-1) times per element and
// global memory only where the caller loads and stores.
//
// Ownership: butterfly t = tid + i*NT of a stage with radix R maps to line l
// and butterfly j (LFAST: l = t % L, j = t / L; otherwise j = t % (N/R),
// l = t / (N/R)).  The first stage reads elements j + r*N/R and the last
// stage produces elements j + r*N/R of its own radix, so a forward plan's
// outputs and the inverse plan with the reversed radix order (REV) have the
// same per-thread ownership: spectra can be multiplied in registers between
// a forward and an inverse transform.
#pragma once
#include <type_traits>

#include "fft_core.h"

constexpr int ct_factor(int n, int idx) {
  int m = n, p2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++p2;
  }
  int n8 = p2 / 3, rem = p2 % 3;
  if (rem == 1 && n8 > 0) {
    n8 -= 1;
    rem = 4;
  }
  int k = 0;
  for (int i = 0; i < n8; ++i, ++k)
    if (k == idx) return 8;
  if (rem == 2) {
    if (k == idx) return 4;
    ++k;
  } else if (rem == 4) {
    if (k == idx) return 4;
    ++k;
    if (k == idx) return 4;
    ++k;
  } else if (rem == 1) {
    if (k == idx) return 2;
    ++k;
  }
  const int odd[3] = {5, 3, 7};
  for (int q = 0; q < 3; ++q) {
    while (m % odd[q] == 0) {
      if (k == idx) return odd[q];
      ++k;
      m /= odd[q];
    }
  }
  return 0;
}

constexpr int ct_nstages0(int n) {
  int s = 0;
  while (ct_factor(n, s) != 0) ++s;
  return s;
}

// Radix plan variants of a length.  Variant 0 is ct_factor's order (radix 8
// stages first, then 4 or 2, then the odd radices).  Variant 1 takes the
// stages in the reverse order; where that order is the same (all stages of
// one radix: 64 = 8*8, 512 = 8*8*8) it splits every pair of radix-8 stages
// into three radix-4 stages instead (512 = 4*4*4*8).  Both variants have
// the same radices per thread (8 elements for radix 8 and 4), so a kernel
// instantiated for either keeps its register ownership.
constexpr bool ct_palindrome(int n) {
  const int ns = ct_nstages0(n);
  for (int s = 0; s < ns; ++s)
    if (ct_factor(n, s) != ct_factor(n, ns - 1 - s)) return false;
  return true;
}

constexpr int ct_eights(int n) {
  int c = 0;
  for (int s = 0; s < ct_nstages0(n); ++s) c += ct_factor(n, s) == 8;
  return c;
}

constexpr int ct_factor_v(int n, int idx, int variant) {
  const int ns = ct_nstages0(n);
  if (variant == 0) return ct_factor(n, idx);
  if (!ct_palindrome(n) || ct_eights(n) < 2) return idx < ns ? ct_factor(n, ns - 1 - idx) : 0;
  // pairs of radix-8 stages become three radix-4 stages; the rest follows
  const int pairs = ct_eights(n) / 2;
  if (idx < 3 * pairs) return 4;
  return ct_factor(n, 2 * pairs + (idx - 3 * pairs));
}

constexpr int ct_nstages(int n, int variant = 0) {
  int s = 0;
  while (ct_factor_v(n, s, variant) != 0) ++s;
  return s;
}

constexpr bool ct_supported(int n, int variant = 0) {
  int p = 1;
  for (int s = 0; s < ct_nstages(n, variant); ++s) p *= ct_factor_v(n, s, variant);
  return p == n;
}

// TT: every twiddle power read from the table (R - 1 table reads per
// butterfly) instead of one read and the powers by recurrence (R - 2 complex
// multiplies), for kernels with VALU work to spare less than LDS bandwidth.
template <int N_, int L_, int NT_, bool LFAST_, bool REV_, int V_ = 0, bool TT_ = false>
struct CtPlan {
  static constexpr int N = N_, L = L_, NT = NT_, V = V_;
  static constexpr bool LFAST = LFAST_, REV = REV_, TT = TT_;
  static constexpr int NST = ct_nstages(N_, V_);
  static_assert(ct_supported(N_, V_), "length not supported by the compile-time FFT");
  static constexpr int radix(int s) { return REV ? ct_factor_v(N_, NST - 1 - s, V_) : ct_factor_v(N_, s, V_); }
  static constexpr int ns(int s) {
    int v = 1;
    for (int i = 0; i < s; ++i) v *= radix(i);
    return v;
  }
  static constexpr int bpt(int R) { return (N_ / R * L_ + NT_ - 1) / NT_; }
  static constexpr int R_IN = radix(0), R_OUT = radix(NST - 1);
  static constexpr int E_IN = bpt(R_IN) * R_IN, E_OUT = bpt(R_OUT) * R_OUT;
  static constexpr bool all_valid() {
    for (int s = 0; s < NST; ++s)
      if ((N_ / radix(s) * L_) % NT_ != 0) return false;
    return true;
  }
  // Line-major plans with idle butterfly slots send those slots to a junk
  // line L (the LDS buffer then holds L + 1 lines), so no LDS access and no
  // butterfly is predicated; element-major plans with idle slots predicate.
  static constexpr bool JUNK = !LFAST_ && !all_valid();
  static constexpr int LDS_LINES = JUNK ? L_ + 1 : L_;

  // line / element of register slot (i, r) of a stage with radix R
  template <int R>
  __device__ static __forceinline__ bool slot(int tid, int i, int& l, int& j) {
    constexpr int NB = N_ / R;
    const int t = tid + i * NT_;
    if (LFAST) {
      l = t % L_;
      j = t / L_;
    } else {
      j = t % NB;
      l = t / NB;
    }
    return (NB * L_) % NT_ == 0 || t < NB * L_;
  }
  __device__ static __forceinline__ bool in_pos(int tid, int i, int r, int& l, int& e) {
    int j;
    const bool ok = slot<R_IN>(tid, i, l, j);
    e = j + r * (N_ / R_IN);
    return ok;
  }
  __device__ static __forceinline__ bool out_pos(int tid, int i, int r, int& l, int& e) {
    int j;
    const bool ok = slot<R_OUT>(tid, i, l, j);
    e = j + r * (N_ / R_OUT);
    return ok;
  }
};

template <int DIR, int R>
__device__ __forceinline__ void ct_dft(float2* v) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 3) dft3<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5) dft5<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else {
    // generic odd prime with compile-time twiddles
    float2 out[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int m = (r * q) % R;
        const double a = (double)DIR * 2.0 * 3.14159265358979323846 * (double)m / (double)R;
        acc = c_add(acc, c_mul(v[r], make_float2((float)__builtin_cos(a), (float)__builtin_sin(a))));
      }
      out[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
  }
}

// Consume registers here: the compiler waits for their loads at this point
// (after whatever was issued since) instead of at their first arithmetic use
// in the next loop iteration, where merging the loop-entry path would make it
// wait for every outstanding store as well.
template <int E>
__device__ __forceinline__ void ct_ready(const float2* v) {
#pragma unroll
  for (int s = 0; s < E; ++s) asm volatile("" ::"v"(v[s].x), "v"(v[s].y));
}

// Slot (i) of a radix-R stage: true if the thread works on it.  For JUNK
// plans every slot is worked on and idle ones address the junk line.
template <class P, int R>
__device__ __forceinline__ bool ct_slot(int tid, int i, int& l, int& j) {
  const bool ok = P::template slot<R>(tid, i, l, j);
  if constexpr (P::JUNK) {
    if (!ok) l = P::L;
    return true;
  } else {
    return ok;
  }
}

// LDS address of (line l, element e): element-major e*LD + l (LFAST) or
// line-major l*LD + e.
template <class P, int LD>
__device__ __forceinline__ int ct_lds(int l, int e) {
  return P::LFAST ? e * LD + l : l * LD + e;
}

// twiddle W_N^(idx) (DIR-signed) from a table of a multiple TWS*N length
template <int DIR, int TWS>
__device__ __forceinline__ float2 ct_tw(const float2* __restrict__ tw, int idx) {
  return tw_load<DIR>(tw, idx * TWS);
}

// v[r] *= w1^r for r = 1..R-1, powers by recurrence (at most R-2 = 6 extra
// roundings for radix 8, ~4e-7 relative).
template <int R>
__device__ __forceinline__ void ct_twiddle_w(float2* v, float2 w1) {
  float2 w = w1;
#pragma unroll
  for (int r = 1; r < R; ++r) {
    v[r] = c_mul(v[r], w);
    if (r + 1 < R) w = c_mul(w, w1);
  }
}

// v[r] *= W^(r*base) for r = 1..R-1 with W^base from the table.
template <int DIR, int TWS, int R>
__device__ __forceinline__ void ct_twiddle(float2* v, const float2* __restrict__ tw, int base) {
  if (base == 0) return;
  ct_twiddle_w<R>(v, ct_tw<DIR, TWS>(tw, base));
}

// No work between stages (the default hook of ct_fft).
struct ct_no_hook {
  template <class K>
  __device__ __forceinline__ void operator()(K) const {}
};

// middle stage s (reads LDS, writes LDS); hook(integral_constant<S>) runs
// after the stage's LDS writes, before the barrier
// wm: this stage's twiddle powers (ct_last_twiddles<P, DIR, TWS, S>), or
// null to form them here
template <class P, int DIR, int LD, int TWS, int S, class H>
__device__ __forceinline__ void ct_mid(float2* lds, const float2* __restrict__ tw, int tid, H& hook,
                                       const float2* wm = nullptr) {
  if constexpr (S < P::NST - 1) {
    constexpr int R = P::radix(S), NS = P::ns(S), NB = P::N / R, TS = P::N / (NS * R), BPT = P::bpt(R);
    // the twiddle is read with the data, before the barrier: the table is
    // never written during the transform
    float2 v[BPT][R], w1[BPT], wt[P::TT ? BPT : 1][R];
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      int l, j;
      if (ct_slot<P, R>(tid, i, l, j)) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[i][r] = lds[ct_lds<P, LD>(l, j + r * NB)];
        const int k = j % NS;
        if (wm) {
        } else if constexpr (P::TT) {
#pragma unroll
          for (int r = 1; r < R; ++r) wt[i][r] = ct_tw<DIR, TWS>(tw, r * k * TS);
        } else if (P::JUNK || k != 0) {
          w1[i] = ct_tw<DIR, TWS>(tw, k * TS);
        }
      }
    }
    lds_barrier();
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      int l, j;
      if (ct_slot<P, R>(tid, i, l, j)) {
        const int k = j % NS;
        if (wm) {
#pragma unroll
          for (int r = 1; r < R; ++r) v[i][r] = c_mul(v[i][r], wm[i * R + r]);
        } else if constexpr (P::TT) {
#pragma unroll
          for (int r = 1; r < R; ++r) v[i][r] = c_mul(v[i][r], wt[i][r]);  // k = 0: W^0 = 1, exact
        } else if (P::JUNK)
          ct_twiddle_w<R>(v[i], w1[i]);  // k = 0 reads W^0 = 1: exact, and no branch
        else if (k != 0)
          ct_twiddle_w<R>(v[i], w1[i]);
        ct_dft<DIR, R>(v[i]);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[ct_lds<P, LD>(l, base + r * NS)] = v[i][r];
      }
    }
    hook(std::integral_constant<int, S>{});
    lds_barrier();
    ct_mid<P, DIR, LD, TWS, S + 1>(lds, tw, tid, hook);
  }
}

// The last stage's twiddle powers of this thread (wl[i * R + r] for slot i,
// r = 1..R-1; r = 0 unused): they depend on the slot alone, so a kernel
// that runs many transforms can compute them once and keep them.
template <class P, int DIR, int TWS, int S = P::NST - 1>
__device__ __forceinline__ void ct_last_twiddles(const float2* __restrict__ tw, float2* wl, int tid) {
  constexpr int R = P::radix(S), NS = P::ns(S), TS = P::N / (NS * R), BPT = P::bpt(R);
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    int l, j;
    ct_slot<P, R>(tid, i, l, j);
    const int k = j % NS;
    wl[i * R] = make_float2(1.f, 0.f);
#pragma unroll
    for (int r = 1; r < R; ++r) wl[i * R + r] = ct_tw<DIR, TWS>(tw, r * k * TS);
  }
}

// last stage (reads LDS, results in registers out[E_OUT]); wl: the powers
// of ct_last_twiddles, or null to form them here
template <class P, int DIR, int LD, int TWS>
__device__ __forceinline__ void ct_last(float2* lds, const float2* __restrict__ tw, float2* out, int tid,
                                        const float2* wl = nullptr) {
  constexpr int S = P::NST - 1;
  constexpr int R = P::radix(S), NS = P::ns(S), NB = P::N / R, TS = P::N / (NS * R), BPT = P::bpt(R);
#pragma unroll
  for (int i = 0; i < BPT; ++i) {
    int l, j;
    if (ct_slot<P, R>(tid, i, l, j)) {
      float2* v = out + i * R;
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = lds[ct_lds<P, LD>(l, j + r * NB)];
      const int k = j % NS;
      if (wl) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], wl[i * R + r]);
      } else if constexpr (P::TT) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], ct_tw<DIR, TWS>(tw, r * k * TS));
      } else if (P::JUNK) {
        ct_twiddle_w<R>(v, ct_tw<DIR, TWS>(tw, k * TS));
      } else {
        ct_twiddle<DIR, TWS, R>(v, tw, k * TS);
      }
      ct_dft<DIR, R>(v);
    }
  }
}

// Transform: in[E_IN] (first-stage ownership) -> out[E_OUT] (last-stage
// ownership).  All threads of the block must call it.  `lds` holds at least
// the L lines at stride LD.  TWS: the table tw has TWS*N entries of
// exp(-2 pi i k / (TWS*N)).  hook(integral_constant<s>) is called after the
// LDS writes of stage s = 0 .. NST-2, before the barrier that ends it: a
// caller issues independent work there (stores of an earlier result) so
// that it spreads over the transform instead of coming in one burst.
template <class P, int DIR, int LD, int TWS = 1, class H = ct_no_hook>
__device__ __forceinline__ void ct_fft(float2* lds, const float2* __restrict__ tw, float2* in, float2* out,
                                       int tid, H hook = H{}) {
  constexpr int R0 = P::radix(0);
  constexpr int B0 = P::bpt(R0);
  if constexpr (P::NST == 1) {
#pragma unroll
    for (int i = 0; i < B0; ++i) {
      ct_dft<DIR, R0>(in + i * R0);
#pragma unroll
      for (int r = 0; r < R0; ++r) out[i * R0 + r] = in[i * R0 + r];
    }
    return;
  } else {
    lds_barrier();  // the buffer may still be read by a previous transform
#pragma unroll
    for (int i = 0; i < B0; ++i) {
      int l, j;
      if (ct_slot<P, R0>(tid, i, l, j)) {
        ct_dft<DIR, R0>(in + i * R0);
#pragma unroll
        for (int r = 0; r < R0; ++r) lds[ct_lds<P, LD>(l, j * R0 + r)] = in[i * R0 + r];
      }
    }
    hook(std::integral_constant<int, 0>{});
    lds_barrier();
    ct_mid<P, DIR, LD, TWS, 1>(lds, tw, tid, hook);
    ct_last<P, DIR, LD, TWS>(lds, tw, out, tid);
  }
}

// The transform from its second stage on: `lds` already holds the outputs
// of the first stage (element j * R0 + r of each line, as ct_fft writes
// them), computed elsewhere.  Out as ct_fft.  All threads of the block must
// call it, after a barrier that follows the writes of those outputs.
template <class P, int DIR, int LD, int TWS = 1>
__device__ __forceinline__ void ct_fft_from1(float2* lds, const float2* __restrict__ tw, float2* out, int tid,
                                             const float2* wl = nullptr, const float2* w1 = nullptr) {
  static_assert(P::NST >= 2, "a transform of one stage has nothing left");
  ct_no_hook hook;
  ct_mid<P, DIR, LD, TWS, 1>(lds, tw, tid, hook, w1);
  ct_last<P, DIR, LD, TWS>(lds, tw, out, tid, wl);
}

// ---- two-buffer forms ------------------------------------------------------
// The same transforms over two LDS buffers X and Y: stage 0 writes X and
// every later stage reads one buffer and writes the other, so a stage needs
// no barrier between its reads and its writes, only one after its writes.
// A transform begins by writing X without a barrier.  That is safe when the
// caller's previous transform on these buffers read X only before its own
// last barrier, i.e. its last stage read Y (ct2_last_in_x false: an odd
// stage count); otherwise the caller swaps X and Y for the next transform.

// middle stage S: reads src, writes dst.  BC (S = 1 only): the first
// stage's outputs are broadcasts, src holding one value per first-stage
// butterfly (element e of that stage's output at e / R0).
template <class P, int DIR, int LD, int TWS, int S, class H, bool BC = false>
__device__ __forceinline__ void ct_mid2(float2* src, float2* dst, const float2* __restrict__ tw, int tid, H& hook) {
  if constexpr (S < P::NST - 1) {
    constexpr int R = P::radix(S), NS = P::ns(S), NB = P::N / R, TS = P::N / (NS * R), BPT = P::bpt(R);
    constexpr int R0 = P::radix(0);
    static_assert(!BC || (S == 1 && NB % R0 == 0), "broadcast input only for the second stage");
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      int l, j;
      if (ct_slot<P, R>(tid, i, l, j)) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[r] = src[ct_lds<P, LD>(l, BC ? (int)((unsigned)j / R0) + r * (NB / R0) : j + r * NB)];
        const int k = j % NS;
        if constexpr (P::TT) {
#pragma unroll
          for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], ct_tw<DIR, TWS>(tw, r * k * TS));
        } else if (P::JUNK) {
          ct_twiddle_w<R>(v, ct_tw<DIR, TWS>(tw, k * TS));
        } else {
          ct_twiddle<DIR, TWS, R>(v, tw, k * TS);
        }
        ct_dft<DIR, R>(v);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[ct_lds<P, LD>(l, base + r * NS)] = v[r];
      }
    }
    hook(std::integral_constant<int, S>{});
    lds_barrier();
    ct_mid2<P, DIR, LD, TWS, S + 1>(dst, src, tw, tid, hook);
  }
}

// the buffer the last stage reads: X after an odd number of stages
// before it (stage 0 included), else Y
template <class P>
constexpr bool ct2_last_in_x() { return (P::NST - 1) % 2 == 1; }

template <class P, int DIR, int LD, int TWS = 1, class H = ct_no_hook>
__device__ __forceinline__ void ct_fft2(float2* X, float2* Y, const float2* __restrict__ tw, float2* in, float2* out,
                                        int tid, H hook = H{}) {
  static_assert(P::NST >= 2, "two-buffer transforms have at least two stages");
  constexpr int R0 = P::radix(0), B0 = P::bpt(R0);
#pragma unroll
  for (int i = 0; i < B0; ++i) {
    int l, j;
    if (ct_slot<P, R0>(tid, i, l, j)) {
      ct_dft<DIR, R0>(in + i * R0);
#pragma unroll
      for (int r = 0; r < R0; ++r) X[ct_lds<P, LD>(l, j * R0 + r)] = in[i * R0 + r];
    }
  }
  hook(std::integral_constant<int, 0>{});
  lds_barrier();
  ct_mid2<P, DIR, LD, TWS, 1>(X, Y, tw, tid, hook);
  ct_last<P, DIR, LD, TWS>(ct2_last_in_x<P>() ? X : Y, tw, out, tid);
}

// Transform of lines whose only non-zero inputs are the first input of each
// first-stage butterfly (zero-padded lines with at most N / R0 leading
// values): the first stage then only broadcasts, so each thread writes its
// butterflies' inputs x[i] (first-stage ownership, one per butterfly) to X
// once and the second stage reads them there (one LDS write per butterfly
// instead of R0).  Out and hooks as ct_fft2.
template <class P, int DIR, int LD, int TWS = 1, class H = ct_no_hook>
__device__ __forceinline__ void ct_fft2_bc0(float2* X, float2* Y, const float2* __restrict__ tw, const float2* x,
                                            float2* out, int tid, H hook = H{}) {
  static_assert(P::NST >= 3, "the broadcast is read by a middle stage");
  constexpr int R0 = P::radix(0), B0 = P::bpt(R0);
#pragma unroll
  for (int i = 0; i < B0; ++i) {
    int l, j;
    if (ct_slot<P, R0>(tid, i, l, j)) X[ct_lds<P, LD>(l, j)] = x[i];
  }
  hook(std::integral_constant<int, 0>{});
  lds_barrier();
  ct_mid2<P, DIR, LD, TWS, 1, H, true>(X, Y, tw, tid, hook);
  ct_last<P, DIR, LD, TWS>(ct2_last_in_x<P>() ? X : Y, tw, out, tid);
}
