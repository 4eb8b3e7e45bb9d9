// Scans of NT2 (cumsum, cumprod, cummax, cummin) over one dimension of a
// contiguous float or double table: cumulative_<tag, bop, neutral> with plus/Zero,
// multiplies/One, max/Minf, min/Inf (boost/simd/swar/functions/{cumsum,
// cumprod,cummax,cummin}.hpp) run by inner_scan / outer_scan over the middle
// index of the (lo, mid, hi) reshaping, with the scalar recurrence
// summary = bop(summary, x(i, m, h)); y(i, m, h) = summary
// (core/swar/include/nt2/core/functions/details/scan_step.hpp,
// scalar/{inner,outer}_scan.hpp).
//
// Every scan is an ordered monoid scan, so the association is fixed and the
// results are the same on every call:
//  * cummax / cummin: the monoid (contains NaN, NT2 recurrence value) with
//    combine(L, R) = R if R contains a NaN, else (L.nan, L.acc > R.acc ?
//    L.acc : R.acc) — associative, and each output equals NT2's recurrence
//    bit for bit (a NaN resets the running value, ties keep the later
//    element, as in the scalar a0 > a1 ? a0 : a1);
//  * cumsum: TwoSum-compensated pairs in the table's type inside a tile,
//    compensated FP64 pairs carried across tiles and segments (more accurate
//    than NT2's sequential recurrence, so not bit-identical to it);
//  * cumprod: an FP64 mantissa with a separate integer exponent, rounded to
//    the table's type per output (no intermediate overflow).
// Kernels:
//  * run kernel (lo == 1, mid >= 512): one workgroup per (run, segment)
//    walks the segment in tiles of 256 x W elements (W = 4 floats or 2
//    doubles per 16 bytes; lane l holds elements W*l..W*l+W-1), scans each
//    tile in the workgroup (wave shuffles, then the four wave totals
//    through LDS) and carries the tile total forward, loading the next tile
//    while it scans the current one;
//  * column kernel (lo > 1, or short runs): one lane per (output, segment)
//    running the recurrence — or, when lo % W == 0, per group of W adjacent
//    outputs with 16-byte loads and stores; neighbouring lanes touch
//    neighbouring addresses; narrow tables (at most 16 such groups) use the
//    tile kernel, whose waves read 4 or more adjacent rows per load,
//    transpose them through LDS so that each lane holds four consecutive
//    rows, scan those in the lane and the lane totals across the wave with
//    shuffles;
//  * long runs split into segments: pass 1 computes each segment's total,
//    pass 2 the exclusive prefix of the segment totals per run (one
//    workgroup per run; one lane per output on the column paths), pass 3
//    rescans each segment from its prefix;
//  * runs of 2^14 to 2^20 floats, 2^22 floats in all (16-byte vectors):
//    one launch with decoupled look-back between tiles, the prefix of each
//    tile folded in the order of the serial recurrence so that the results
//    are reproducible.
// (Kernels and launch plans shared by libnt2hip's nt2h_scan_f32/f64 and the
// C++ front end, which instantiates them over expressions.)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../nt2hip.h"
#include "fold.hpp"
#include "fold_kernels.hpp"

namespace nt2h {


// Scan geometry: element m of run o (the table's output o = i + lo*h) sits
// at i + lo*(m + mid*h).  Elements are read from Src (a table, or an
// expression evaluated per element: at(i), and raw(i)/value()/vec(i) for i
// a multiple of the 16-byte width W) and written to the table y.
template <class T, class Src> struct SGeo {
  Src src;
  T* y;
  int64_t lo, mid, hi;
  __device__ __forceinline__ int64_t at(int64_t o, int64_t m) const {
    if (lo == 1) return o * mid + m;
    const int64_t h = o / lo, i = o - h * lo;
    return i + lo * (m + mid * h);
  }
};
struct SSumm {
  double w[2];
};

template <class T> __device__ __forceinline__ void two_sum_s(T a, T b, T& s, T& e) {
  s = a + b;
  T bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}
__device__ __forceinline__ float shfl_up_t(float v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ double shfl_up_t(double v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ int shfl_up_t(int v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ float shfl_t(float v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ double shfl_t(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ int shfl_t(int v, int src) { return __shfl(v, src, 64); }

// ------------------------------------------------------------- cumsum
template <class T> struct MSum {
  struct V {
    T s, c;
  };
  struct C {
    double s, c;  // compensated FP64 carry
  };
  __device__ static V ident() { return V{T(0), T(0)}; }
  __device__ static V lift(T x) { return V{x, T(0)}; }
  __device__ static V comb(const V& a, const V& b) {
    T t, e;
    two_sum_s(a.s, b.s, t, e);
    return V{t, (a.c + b.c) + e};
  }
  // combine inside a tile (at most 1024 elements) and the form kept in LDS
  __device__ static V combs(const V& a, const V& b) { return comb(a, b); }
  __device__ static V tnorm(const V& v) { return v; }
  static constexpr bool kComm = true;  // the total may be accumulated in any fixed order
  __device__ static V shfl_up(const V& v, int d) { return V{shfl_up_t(v.s, d), shfl_up_t(v.c, d)}; }
  __device__ static V shfl(const V& v, int src) { return V{shfl_t(v.s, src), shfl_t(v.c, src)}; }
  __device__ static C cident() { return C{0.0, 0.0}; }
  __device__ static C cshfl_up(const C& v, int d) { return C{shfl_up_t(v.s, d), shfl_up_t(v.c, d)}; }
  __device__ static C ccomb(const C& a, const C& b) {
    double t, e;
    two_sum_s(a.s, b.s, t, e);
    return C{t, (a.c + b.c) + e};
  }
  __device__ static C cadd(const C& c, const V& v) { return ccomb(c, C{(double)v.s, (double)v.c}); }
  __device__ static T fin(const C& c, const V& v) {
    C r = cadd(c, v);
    return (T)comp_value(r.s, r.c);
  }
  __device__ static SSumm pack(const C& c) { return SSumm{{c.s, c.c}}; }
  __device__ static C unpack(const SSumm& s) { return C{s.w[0], s.w[1]}; }
};

// ------------------------------------------------------------ cumprod
template <class T> struct MProd {
  struct V {
    double m;
    int e;
  };
  using C = V;
  __device__ static V norm(V a) {
    if (a.m != 0.0 && __builtin_isfinite(a.m)) {
      int k;
      a.m = __builtin_frexp(a.m, &k);
      a.e += k;
      a.e = a.e > (1 << 28) ? (1 << 28) : (a.e < -(1 << 28) ? -(1 << 28) : a.e);
    }
    return a;
  }
  __device__ static V ident() { return V{1.0, 0}; }
  __device__ static V lift(T x) { return norm(V{(double)x, 0}); }
  __device__ static V comb(const V& a, const V& b) { return norm(V{a.m * b.m, a.e + b.e}); }
  // Inside a tile a chain multiplies at most 256 normalised mantissas per
  // wave (>= 2^-256) and three normalised wave totals, so the product stays
  // a normal double without renormalising; wave totals are renormalised
  // before they are combined across waves.
  __device__ static V combs(const V& a, const V& b) { return V{a.m * b.m, a.e + b.e}; }
  __device__ static V tnorm(const V& v) { return norm(v); }
  static constexpr bool kComm = true;
  __device__ static V shfl_up(const V& v, int d) { return V{shfl_up_t(v.m, d), shfl_up_t(v.e, d)}; }
  __device__ static V shfl(const V& v, int src) { return V{shfl_t(v.m, src), shfl_t(v.e, src)}; }
  __device__ static C cident() { return ident(); }
  __device__ static C cshfl_up(const C& v, int d) { return shfl_up(v, d); }
  __device__ static C cadd(const C& c, const V& v) { return comb(c, v); }
  __device__ static C ccomb(const C& a, const C& b) { return comb(a, b); }
  __device__ static T fin(const C& c, const V& v) {
    V r = comb(c, v);
    if (r.m == 0.0 || !__builtin_isfinite(r.m)) return (T)r.m;
    int e = r.e > 2000 ? 2000 : (r.e < -2000 ? -2000 : r.e);
    return (T)__builtin_ldexp(r.m, e);
  }
  __device__ static SSumm pack(const C& c) { return SSumm{{c.m, (double)c.e}}; }
  __device__ static C unpack(const SSumm& s) { return C{s.w[0], (int)s.w[1]}; }
};

// ---------------------------------------------------- cummax / cummin
template <bool MAX, class T> struct MExt {
  struct V {
    int nan;  // the covered elements contain a NaN
    T acc;    // NT2's recurrence over them, started from the neutral
  };
  using C = V;
  __device__ static T neutral() { return MAX ? -(T)__builtin_inf() : (T)__builtin_inf(); }
  __device__ static V ident() { return V{0, neutral()}; }
  __device__ static V lift(T x) { return V{x != x ? 1 : 0, x}; }
  __device__ static V comb(const V& a, const V& b) {
    if (b.nan) return b;
    const bool keep_a = MAX ? (a.acc > b.acc) : (a.acc < b.acc);
    return V{a.nan, keep_a ? a.acc : b.acc};
  }
  __device__ static V combs(const V& a, const V& b) { return comb(a, b); }
  __device__ static V tnorm(const V& v) { return v; }
  static constexpr bool kComm = false;  // a NaN resets: only ordered combines
  __device__ static V shfl_up(const V& v, int d) { return V{shfl_up_t(v.nan, d), shfl_up_t(v.acc, d)}; }
  __device__ static V shfl(const V& v, int src) { return V{shfl_t(v.nan, src), shfl_t(v.acc, src)}; }
  __device__ static C cident() { return ident(); }
  __device__ static C cshfl_up(const C& v, int d) { return shfl_up(v, d); }
  __device__ static C cadd(const C& c, const V& v) { return comb(c, v); }
  __device__ static C ccomb(const C& a, const C& b) { return comb(a, b); }
  __device__ static T fin(const C& c, const V& v) { return comb(c, v).acc; }
  __device__ static SSumm pack(const C& c) { return SSumm{{(double)c.nan, (double)c.acc}}; }
  __device__ static C unpack(const SSumm& s) { return C{(int)s.w[0], (T)s.w[1]}; }
};

template <int OP, class T> struct SFam;
template <class T> struct SFam<NT2H_SCAN_CUMSUM, T> { using type = MSum<T>; };
template <class T> struct SFam<NT2H_SCAN_CUMPROD, T> { using type = MProd<T>; };
template <class T> struct SFam<NT2H_SCAN_CUMMAX, T> { using type = MExt<true, T>; };
template <class T> struct SFam<NT2H_SCAN_CUMMIN, T> { using type = MExt<false, T>; };

// Ordered inclusive scan of one value per lane over the 64 lanes of a wave
// (Hillis-Steele: each step combines a lane with the lane d below it, the
// lower lane on the left).
template <class M> __device__ __forceinline__ typename M::V wave_inscan(typename M::V v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    typename M::V o = M::shfl_up(v, d);
    if (lane >= d) v = M::combs(o, v);
  }
  return v;
}

// Run kernel body, one workgroup: scan elements [m0, m1) of run o starting
// from carry, in tiles of kBlock x W elements (lane l holds elements
// W*l..W*l+W-1 of the tile; W = 4 floats or 2 doubles with 16-byte
// accesses, else 1).  Each wave scans its quarter of the tile with
// shuffles; the four wave totals pass through LDS (two slots used in turn,
// so one barrier per tile suffices), and the tile total is carried forward.
// The next tile is loaded before the current one is scanned, so the loads
// stay in flight across the scan and the barrier.  WRITE = false computes
// only the total (pass 1 of segmented scans); for monoids whose total may be
// taken in any fixed order (sum, product) each lane then folds its own
// elements and the lanes are combined once, in lane order, at the end.
template <class M, class T, bool VEC, bool WRITE, class Src>
__device__ __forceinline__ typename M::C run_segment(const SGeo<T, Src>& g, int64_t o, int64_t m0, int64_t m1,
                                                     typename M::C carry) {
  using V = typename M::V;
  using VT = typename Vec16<T>::type;
  constexpr int W = VEC ? Vec16<T>::n : 1;
  constexpr int64_t kTile = (int64_t)kBlock * W;
  constexpr bool kLaneTotals = !WRITE && M::kComm;
  __shared__ V lds[2][kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t run = o * g.mid;
  T* out = g.y + o * g.mid;
  auto load = [&](int64_t base, T(&e)[W]) {
    const int64_t m = base + W * (int64_t)threadIdx.x;
    if (VEC && m + W - 1 < m1) {
      VT q = g.src.vec(run + m);
#pragma unroll
      for (int k = 0; k < W; ++k) e[k] = q[k];
    } else {
#pragma unroll
      for (int k = 0; k < W; ++k) e[k] = (m + k < m1) ? g.src.at(run + m + k) : T(0);
    }
  };
  T cur[W] = {}, nxt[W] = {};
  if (m0 < m1) load(m0, cur);
  V lanes = M::ident();
  int par = 0;
  for (int64_t base = m0; base < m1; base += kTile) {
    if (base + kTile < m1) load(base + kTile, nxt);
    const int64_t m = base + W * (int64_t)threadIdx.x;
    V p[W];
    V acc = M::ident();
#pragma unroll
    for (int k = 0; k < W; ++k) {
      V xk = (m + k < m1) ? M::lift(cur[k]) : M::ident();
      acc = (k == 0) ? xk : M::combs(acc, xk);
      p[k] = acc;
    }
    if (kLaneTotals) {
      lanes = M::comb(lanes, acc);
    } else {
      V inc = wave_inscan<M>(acc);
      if (lane == 63) lds[par][wave] = M::tnorm(inc);
      __syncthreads();
      V wpre = M::ident(), tot = M::ident();
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        V t = lds[par][w];
        if (w < wave) wpre = M::combs(wpre, t);
        tot = M::combs(tot, t);
      }
      if (WRITE) {
        V ex = M::shfl_up(inc, 1);
        if (lane == 0) ex = M::ident();
        ex = M::combs(wpre, ex);
        T r[W];
#pragma unroll
        for (int k = 0; k < W; ++k) r[k] = M::fin(carry, M::combs(ex, p[k]));
        if (VEC && m + W - 1 < m1) {
          VT q;
#pragma unroll
          for (int k = 0; k < W; ++k) q[k] = r[k];
          __builtin_nontemporal_store(q, reinterpret_cast<VT*>(out + m));
        } else {
#pragma unroll
          for (int k = 0; k < W; ++k)
            if (m + k < m1) out[m + k] = r[k];
        }
      }
      carry = M::cadd(carry, tot);
      par ^= 1;
    }
#pragma unroll
    for (int k = 0; k < W; ++k) cur[k] = nxt[k];
  }
  if (kLaneTotals) {
    V inc = wave_inscan<M>(lanes);
    if (lane == 63) lds[par][wave] = M::tnorm(inc);
    __syncthreads();
    V tot = M::ident();
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot = M::comb(tot, lds[par][w]);
    carry = M::cadd(carry, tot);
  }
  return carry;
}

// lo == 1: block b -> (run o = b / segs, segment s).  PASS 0: single pass
// (segs == 1); PASS 1: segment totals; PASS 3: rescan from the prefixes.
template <int OP, class T, bool VEC, int PASS, class Src>
__global__ __launch_bounds__(kBlock) void scan_run_kernel(SGeo<T, Src> g, int64_t segs, int64_t seglen, SSumm* summ) {
  using M = typename SFam<OP, T>::type;
  const int64_t b = blockIdx.x;
  const int64_t o = b / segs, s = b - o * segs;
  const int64_t m0 = s * seglen;
  const int64_t m1 = (m0 + seglen < g.mid) ? m0 + seglen : g.mid;
  if (PASS == 1) {
    typename M::C t = run_segment<M, T, VEC, false, Src>(g, o, m0, m1, M::cident());
    if (threadIdx.x == 0) summ[o * segs + s] = M::pack(t);
  } else {
    typename M::C c = (PASS == 3) ? M::unpack(summ[o * segs + s]) : M::cident();
    run_segment<M, T, VEC, true, Src>(g, o, m0, m1, c);
  }
}

// Column kernel: lane per (output, segment), the recurrence run in order.
// PASS as above.
template <int OP, class T, int PASS, class Src>
__global__ __launch_bounds__(kBlock) void scan_col_kernel(SGeo<T, Src> g, int64_t segs, int64_t seglen, SSumm* summ,
                                                         int64_t so, int64_t ss) {
  using M = typename SFam<OP, T>::type;
  const int64_t nout = g.lo * g.hi;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nout * segs) return;
  const int64_t s = t / nout, o = t - s * nout;
  const int64_t m0 = s * seglen;
  const int64_t m1 = (m0 + seglen < g.mid) ? m0 + seglen : g.mid;
  typename M::C c = (PASS == 3) ? M::unpack(summ[o * so + s * ss]) : M::cident();
  typename M::V acc = M::ident();
  for (int64_t m = m0; m < m1; ++m) {
    const int64_t k = g.at(o, m);
    acc = M::comb(acc, M::lift(g.src.at(k)));
    if (PASS != 1) g.y[k] = M::fin(c, acc);
  }
  if (PASS == 1) summ[o * so + s * ss] = M::pack(M::cadd(c, acc));
}

// Column kernel, 16-byte form (lo % W == 0, x and y 16-byte aligned): one
// lane per (group of W adjacent outputs, segment), one 16-byte load and
// store per step along m, loads issued four steps ahead.
template <int OP, class T, int PASS, class Src>
__global__ __launch_bounds__(kBlock) void scan_colv_kernel(SGeo<T, Src> g, int64_t segs, int64_t seglen, SSumm* summ,
                                                          int64_t so, int64_t ss) {
  using M = typename SFam<OP, T>::type;
  using V = typename M::V;
  using C = typename M::C;
  using VT = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::n;
  const int64_t nq = g.lo * g.hi / W;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nq * segs) return;
  const int64_t s = t / nq, q = t - s * nq;
  const int64_t o = W * q;
  const int64_t h = o / g.lo, i = o - h * g.lo;
  const int64_t off = i + g.lo * g.mid * h;
  VT* py = reinterpret_cast<VT*>(g.y + off);
  const int64_t step = g.lo / W;
  const int64_t m0 = s * seglen;
  const int64_t m1 = (m0 + seglen < g.mid) ? m0 + seglen : g.mid;
  C c[W];
  V acc[W];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    c[k] = (PASS == 3) ? M::unpack(summ[(o + k) * so + s * ss]) : M::cident();
    acc[k] = M::ident();
  }
  int64_t m = m0;
  for (; m + 4 <= m1; m += 4) {
    typename Src::Raw w[4];  // every load in flight before the first value
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = g.src.raw(off + (m + u) * g.lo);
    VT v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = g.src.value(w[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      VT r;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        acc[k] = M::comb(acc[k], M::lift(v[u][k]));
        r[k] = M::fin(c[k], acc[k]);
      }
      if (PASS != 1) __builtin_nontemporal_store(r, py + (m + u) * step);
    }
  }
  for (; m < m1; ++m) {
    VT v = g.src.vec(off + m * g.lo);
    VT r;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      acc[k] = M::comb(acc[k], M::lift(v[k]));
      r[k] = M::fin(c[k], acc[k]);
    }
    if (PASS != 1) __builtin_nontemporal_store(r, py + m * step);
  }
  if (PASS == 1) {
#pragma unroll
    for (int k = 0; k < W; ++k) summ[(o + k) * so + s * ss] = M::pack(M::cadd(c[k], acc[k]));
  }
}

// Lanes of one wave exchanging data through LDS: make the writes visible
// to the other lanes of the wave (no workgroup barrier).
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Column kernel for narrow tables (lo % W == 0 and at most 16 groups of W
// outputs): a wave is R = 64 / P row phases of P lanes (P a power of two).  A step covers K R rows (K = 4): the wave loads them with lane
// (r, p) reading rows mb + r + R j (j < K) of group q = qb * P + p — one load
// instruction reads R adjacent rows, 1 KiB of consecutive addresses when the
// groups cover the rows — and transposes them through LDS (padded: one
// 16-byte slot every four rows, so neither the writes nor the reads of four
// consecutive rows per lane pile onto a few banks) so that lane (r, p) holds
// the K consecutive rows mb + K r .. mb + K r + K - 1.  Each lane scans its K
// rows in order, the lane totals are scanned across the phases with
// shuffles (lower rows on the left), each row's result combines the
// exclusive prefix of the lanes before with the lane's own prefix (the
// results go back through LDS, so that the stores are as contiguous as the
// loads), and the step total — the last phase's inclusive value — is
// carried to the next step.  A wave past the last
// (segment, group block) leaves at once: no workgroup barriers.  PASS as
// for the other column kernels.
template <int OP, class T, int PASS, class Src>
__global__ __launch_bounds__(kBlock) void scan_tile_kernel(SGeo<T, Src> g, int lgP, int64_t qblocks, int64_t segs,
                                                           int64_t seglen, SSumm* summ, int64_t so, int64_t ss) {
  using M = typename SFam<OP, T>::type;
  using V = typename M::V;
  using C = typename M::C;
  using VT = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::n;
  constexpr int K = 4;
  const int P = 1 << lgP, R = 64 >> lgP;
  const int lane = threadIdx.x & 63;
  const int p = lane & (P - 1), r = lane >> lgP;
  const int last = (R - 1) * P + p;  // this group's last phase
  const int64_t wv = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (wv >= qblocks * segs) return;
  const int64_t s = wv / qblocks, q = (wv - s * qblocks) * P + p;
  const int64_t nout = g.lo * g.hi;
  const bool live = q < nout / W;
  const int64_t o = W * q;
  const int64_t h = o / g.lo, i = o - h * g.lo;
  const int64_t off = i + g.lo * g.mid * h;
  VT* py = reinterpret_cast<VT*>(g.y + off);
  const int64_t step = g.lo / W;
  const int64_t m0 = s * seglen;
  const int64_t m1 = (m0 + seglen < g.mid) ? m0 + seglen : g.mid;
  __shared__ VT tl[kWaves][K * 64 + K * 16];  // K R rows per group, P groups, one pad per four rows
  VT* tw = tl[threadIdx.x >> 6];
  const int gb = p * K * R;  // this group's rows in the wave's tile
  if (PASS == 1 && M::kComm) {
    // segment totals of sum and product: each lane folds its rows in load
    // order (no transposition, no scan), the phases are combined once at
    // the end in a fixed tree
    V acc[W];
#pragma unroll
    for (int k = 0; k < W; ++k) acc[k] = M::ident();
    for (int64_t mb = m0; mb < m1; mb += (int64_t)K * R) {
      typename Src::Raw w[K];  // every load in flight before the first value
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int64_t m = mb + r + (int64_t)R * j;
        w[j] = (live && m < m1) ? g.src.raw(off + m * g.lo) : typename Src::Raw{};
      }
      VT v[K];  // rows past the segment are masked where they are combined
#pragma unroll
      for (int j = 0; j < K; ++j) v[j] = g.src.value(w[j]);
#pragma unroll
      for (int k = 0; k < W; ++k) {
        V a = M::ident();
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const bool ok = live && mb + r + (int64_t)R * j < m1;
          a = M::combs(a, ok ? M::lift(v[j][k]) : M::ident());
        }
        acc[k] = M::comb(acc[k], a);
      }
    }
    for (int d = R >> 1; d >= 1; d >>= 1) {
      const int src = (r + d < R) ? lane + d * P : lane;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        V other = M::shfl(acc[k], src);
        if (r < d) acc[k] = M::comb(acc[k], other);
      }
    }
    if (r == 0 && live) {
#pragma unroll
      for (int k = 0; k < W; ++k) summ[(o + k) * so + s * ss] = M::pack(M::cadd(M::cident(), M::tnorm(acc[k])));
    }
    return;
  }
  C c[W];
#pragma unroll
  for (int k = 0; k < W; ++k) c[k] = (PASS == 3 && live) ? M::unpack(summ[(o + k) * so + s * ss]) : M::cident();
  for (int64_t mb = m0; mb < m1; mb += (int64_t)K * R) {
    const int64_t mr = mb + (int64_t)K * r;  // this lane's first row after the transposition
    typename Src::Raw w[K];  // every load in flight before the first value
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int64_t m = mb + r + (int64_t)R * j;
      w[j] = (live && m < m1) ? g.src.raw(off + m * g.lo) : typename Src::Raw{};
    }
    VT v[K];  // rows past the segment are masked where they are scanned
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] = g.src.value(w[j]);
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int t = gb + r + R * j;
      tw[t + (t >> 2)] = v[j];
    }
    wave_sync_lds();
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int t = gb + K * r + j;
      v[j] = tw[t + (t >> 2)];
    }
    wave_sync_lds();  // the tile is rewritten by the next step
    VT res[K];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      V pre[K];
      V a = M::ident();
#pragma unroll
      for (int j = 0; j < K; ++j) {
        V xj = (live && mr + j < m1) ? M::lift(v[j][k]) : M::ident();
        a = (j == 0) ? xj : M::combs(a, xj);
        pre[j] = a;
      }
      V inc = a;
      for (int d = 1; d < R; d <<= 1) {
        V lower = M::shfl_up(inc, d * P);
        if (r >= d) inc = M::combs(lower, inc);
      }
      if (PASS != 1) {
        V ex = M::shfl_up(inc, P);
        if (r == 0) ex = M::ident();
#pragma unroll
        for (int j = 0; j < K; ++j) res[j][k] = M::fin(c[k], M::combs(ex, pre[j]));
      }
      c[k] = M::cadd(c[k], M::tnorm(M::shfl(inc, last)));
    }
    if (PASS != 1) {
      // back through LDS, so that each store instruction writes R adjacent rows
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int t = gb + K * r + j;
        tw[t + (t >> 2)] = res[j];
      }
      wave_sync_lds();
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int t = gb + r + R * j;
        const int64_t m = mb + r + (int64_t)R * j;
        const VT out = tw[t + (t >> 2)];
        if (live && m < m1) __builtin_nontemporal_store(out, py + m * step);
      }
      wave_sync_lds();
    }
  }
  if (PASS == 1 && r == 0 && live) {
#pragma unroll
    for (int k = 0; k < W; ++k) summ[(o + k) * so + s * ss] = M::pack(c[k]);
  }
}

// Replace base[s * st], s = s0..s1-1, by their exclusive prefixes starting
// from pre, in order (eight loads in flight); returns the running total.
template <class M>
__device__ __forceinline__ typename M::C exclusive_range(typename M::C pre, SSumm* base, int64_t s0, int64_t s1,
                                                         int64_t st) {
  int64_t s = s0;
  for (; s + 8 <= s1; s += 8) {
    SSumm b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) b[u] = base[(s + u) * st];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      base[(s + u) * st] = M::pack(pre);
      pre = M::ccomb(pre, M::unpack(b[u]));
    }
  }
  for (; s < s1; ++s) {
    typename M::C t = M::unpack(base[s * st]);
    base[s * st] = M::pack(pre);
    pre = M::ccomb(pre, t);
  }
  return pre;
}
// Total of base[s * st], s = s0..s1-1, in order (eight loads in flight).
template <class M>
__device__ __forceinline__ typename M::C total_range(const SSumm* base, int64_t s0, int64_t s1, int64_t st) {
  typename M::C a = M::cident();
  int64_t s = s0;
  for (; s + 8 <= s1; s += 8) {
    SSumm b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) b[u] = base[(s + u) * st];
#pragma unroll
    for (int u = 0; u < 8; ++u) a = M::ccomb(a, M::unpack(b[u]));
  }
  for (; s < s1; ++s) a = M::ccomb(a, M::unpack(base[s * st]));
  return a;
}

// Pass 2 of the column paths: one lane per output replaces its segment
// totals (segment-major) by their exclusive prefixes, in segment order.
template <int OP, class T>
__global__ __launch_bounds__(kBlock) void scan_col_prefix_kernel(int64_t nout, int64_t segs, SSumm* summ) {
  using M = typename SFam<OP, T>::type;
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= nout) return;
  exclusive_range<M>(M::cident(), summ + o, 0, segs, nout);
}

// Pass 2: per output (one workgroup of kPrefixBlock lanes), replace the
// segment totals by their exclusive prefixes, in segment order.  Lane l
// totals a contiguous range of segments (loads batched), the lane totals are
// scanned in lane order (shuffles within each wave, then the wave totals in
// wave order through LDS), and each lane rewrites its range from its
// exclusive prefix.  The totals of output o sit at o * ostride + s * sstride
// (run paths: o * segs + s; column paths with few outputs: s * nout + o).
constexpr int kPrefixBlock = 1024;
template <int OP, class T>
__global__ __launch_bounds__(kPrefixBlock) void scan_prefix_kernel(int64_t segs, SSumm* summ, int64_t ostride,
                                                                   int64_t sstride) {
  using M = typename SFam<OP, T>::type;
  using C = typename M::C;
  constexpr int kW = kPrefixBlock / 64;
  __shared__ SSumm sh[kW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t o = blockIdx.x;
  SSumm* base = summ + o * ostride;
  const int64_t per = (segs + kPrefixBlock - 1) / kPrefixBlock;
  const int64_t s0 = threadIdx.x * per < segs ? threadIdx.x * per : segs;
  const int64_t s1 = (s0 + per < segs) ? s0 + per : segs;
  C inc = total_range<M>(base, s0, s1, sstride);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    C lower = M::cshfl_up(inc, d);
    if (lane >= d) inc = M::ccomb(lower, inc);
  }
  if (lane == 63) sh[w] = M::pack(inc);
  __syncthreads();
  C wpre = M::cident();
  for (int ww = 0; ww < w; ++ww) wpre = M::ccomb(wpre, M::unpack(sh[ww]));
  C ex = M::cshfl_up(inc, 1);
  if (lane == 0) ex = M::cident();
  exclusive_range<M>(M::ccomb(wpre, ex), base, s0, s1, sstride);
}

// Single-pass scan of long runs (lo == 1, 16-byte vectors) with decoupled
// look-back.  Workgroups take tiles of kLbSub x kBlock x W elements in
// ticket order; each scans its tile in registers (as run_segment does, the
// tile aggregate A falling out), publishes A, and finds the exclusive
// prefix P of its tile from the published state of the tiles before it:
// wave 0 reads the states of up to 64 predecessors at once, takes the
// nearest one that has published its inclusive value I_j (every tile in
// between having published its aggregate) and folds
//   P = (((I_j + A_{j+1}) + A_{j+2}) + ... ) + A_{i-1}
// left to right — the grouping of the serial recurrence I_k = I_{k-1} + A_k,
// so the result does not depend on which j the timing offered: the output
// bits are the same on every run.  It then publishes I_i = P + A_i and
// writes its tile.  Every tile reads and writes its input and output once,
// in one launch (the three passes take three).
//
// Publication: the values are written with agent-scope atomic stores
// (through to the point of coherence of all XCDs), completed (s_waitcnt),
// then the tile's state word — (epoch << 32) | 1 for "aggregate", | 2 for
// "inclusive" — with another; readers load the state word and then the
// values with agent-scope atomic loads.  The epoch (the number of previous
// launches on this state) makes the state of earlier calls read as "not
// yet"; the workgroup holding the last ticket resets the ticket counter and
// advances the epoch, so no memset is ever needed (graph-safe).  The ticket
// is the kernel's only atomic on a shared address: such atomics are served
// one at a time (~35 ns each measured), which is what bounds this kernel to
// small tile counts.
// Forward progress: a tile waits only on tiles with smaller tickets, which
// are already running and publish their aggregate before waiting on
// anything.  Every wait is nevertheless bounded (spin_cap polls with a
// sleep between them): past it the workgroup publishes a poisoned state
// (which every later tile of the run passes on instead of a value), records
// an error in ctrl[3] and finishes, so the kernel always ends; nt2h_sync on
// the stream reports and clears the error.
// Sub-tiles of kBlock x W elements per look-back tile (the kernel's kLbSub,
// chosen per call): 4 (16 KiB) for short jobs, 8 (32 KiB: half the tickets
// and half the chain) once the 16 KiB
// tiles would exceed 64 per run or 256 in all (measured: 2^14 floats 7.6 us
// with 4 against 9.8 with 8; 2^19 floats 18.3 against 15.9; 64 runs of 2^16
// 26.7 us in three passes against 20.0 with 8)
constexpr unsigned kLbSpinCap = 1u << 20;
// Two limits, both measured against the three passes (three launches):
// the tickets cost ~35 ns each, one after the other (16384 tiles: 887 us
// against 154), and the inclusive values of a run advance by at most 64
// tiles per memory round trip.  Up to 512 tiles in all and 128 per run the
// single launch wins (2^14 floats: 18.5 -> 7.2 us, 2^18: 19.1 -> 13.0 us,
// 16 runs of 2^16: 22.8 -> 12.4 us); beyond, three passes.
constexpr int64_t kLbMaxTiles = 128;  // per run
constexpr int64_t kLbMaxTotal = 512;
__device__ __forceinline__ void lb_put(SSumm* p, const SSumm& v) {
  publish(&p->w[0], v.w[0]);
  publish(&p->w[1], v.w[1]);
}
__device__ __forceinline__ SSumm lb_get(const SSumm* p) { return SSumm{{fetch(&p->w[0]), fetch(&p->w[1])}}; }
__device__ __forceinline__ SSumm shfl_summ(const SSumm& v, int src) {
  return SSumm{{__shfl(v.w[0], src, 64), __shfl(v.w[1], src, 64)}};
}
__device__ __forceinline__ void lb_state(unsigned long long* f, unsigned long long v) {
  __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long lb_state_get(const unsigned long long* f) {
  return __hip_atomic_load(const_cast<unsigned long long*>(f), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int OP, class T, int kLbSub, class Src>
__global__ __launch_bounds__(kBlock) void scan_lookback_kernel(SGeo<T, Src> g, int64_t tpr, int64_t ntiles,
                                                               unsigned* ctrl, unsigned long long* state,
                                                               SSumm* agg, SSumm* incl, unsigned spin_cap,
                                                               unsigned epoch_mask) {
  using M = typename SFam<OP, T>::type;
  using V = typename M::V;
  using C = typename M::C;
  using VT = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::n;
  constexpr int64_t kSub = (int64_t)kBlock * W;
  __shared__ V lds[kLbSub][kWaves];
  __shared__ unsigned sh_t[2];
  __shared__ SSumm sh_p;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    // the epoch is read before the ticket is taken (and the load completed):
    // the tile holding the last ticket advances it only after every ticket
    // is gone, so every workgroup of this launch sees the same epoch
    sh_t[1] = __hip_atomic_load(&ctrl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    sh_t[0] = __hip_atomic_fetch_add(&ctrl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const int64_t t = sh_t[0];
  const unsigned epoch = sh_t[1];
  if (t >= ntiles) {
    // more tickets than tiles: another launch shares this state (a graph
    // replayed alongside calls on its capture stream).  Touch nothing, flag
    // the error for nt2h_sync.
    if (threadIdx.x == 0) __hip_atomic_store(&ctrl[3], 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const int64_t o = t / tpr, i = t - o * tpr;  // run, tile within the run
  const int64_t m0 = i * kLbSub * kSub;
  const int64_t m1 = (m0 + kLbSub * kSub < g.mid) ? m0 + kLbSub * kSub : g.mid;
  const int64_t run = o * g.mid;
  T* out = g.y + o * g.mid;
  // the whole tile in registers (mid % W == 0: a vector is all in or all out)
  typename Src::Raw w[kLbSub];  // every load in flight before the first value
#pragma unroll
  for (int u = 0; u < kLbSub; ++u) {
    const int64_t m = m0 + u * kSub + W * (int64_t)threadIdx.x;
    w[u] = (m < m1) ? g.src.raw(run + m) : typename Src::Raw{};
  }
  VT e[kLbSub];
#pragma unroll
  for (int u = 0; u < kLbSub; ++u) {
    const int64_t m = m0 + u * kSub + W * (int64_t)threadIdx.x;
    e[u] = (m < m1) ? g.src.value(w[u]) : VT{};
  }
  V p[kLbSub][W], inc[kLbSub];
#pragma unroll
  for (int u = 0; u < kLbSub; ++u) {
    const int64_t m = m0 + u * kSub + W * (int64_t)threadIdx.x;
    V acc = M::ident();
#pragma unroll
    for (int k = 0; k < W; ++k) {
      V xk = (m < m1) ? M::lift(e[u][k]) : M::ident();
      acc = (k == 0) ? xk : M::combs(acc, xk);
      p[u][k] = acc;
    }
    inc[u] = wave_inscan<M>(acc);
    if (lane == 63) lds[u][wave] = M::tnorm(inc[u]);
  }
  __syncthreads();
  C L[kLbSub + 1];  // local carries: the tile's own prefix before each sub-tile
  V wpre[kLbSub];
  L[0] = M::cident();
#pragma unroll
  for (int u = 0; u < kLbSub; ++u) {
    V tot = M::ident();
    wpre[u] = M::ident();
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      V tw = lds[u][w];
      if (w < wave) wpre[u] = M::combs(wpre[u], tw);
      tot = M::combs(tot, tw);
    }
    L[u + 1] = M::cadd(L[u], tot);
  }
  if (wave == 0) {
    const unsigned long long tag = (unsigned long long)epoch << 32;
    C P = M::cident();
    if (i == 0) {
      if (lane == 0) {
        lb_put(&incl[t], M::pack(L[kLbSub]));
        __builtin_amdgcn_s_waitcnt(0);
        lb_state(&state[t], tag | 2ull);
      }
    } else {
      if (lane == 0) {
        lb_put(&agg[t], M::pack(L[kLbSub]));
        __builtin_amdgcn_s_waitcnt(0);
        lb_state(&state[t], tag | 1ull);
      }
      // nearest predecessor in a final state: inclusive (2), or poisoned (3)
      // by a tile whose wait gave up — poison is passed on, so that no tile
      // after a failed wait publishes a value as if it were valid
      int jl = -1;     // lane (= distance - 1) of that predecessor
      bool poison = false;
      const unsigned want = epoch ^ epoch_mask;
      for (unsigned spins = 0;; ++spins) {
        const bool inrun = i - 1 - lane >= 0;
        const unsigned long long f = inrun ? lb_state_get(&state[t - 1 - lane]) : 0ull;
        const bool ready = inrun && (unsigned)(f >> 32) == want && (f & 3ull) != 0;
        const uint64_t mi = __ballot(ready && (f & 3ull) >= 2ull), mr = __ballot(ready);
        const uint64_t mp = __ballot(ready && (f & 3ull) == 3ull);
        if (mi) {
          const int j = __builtin_ctzll(mi);
          const uint64_t need = (j == 0) ? 0ull : (~0ull >> (64 - j));  // lanes 0..j-1
          if ((mr & need) == need) {
            jl = j;
            poison = ((mp >> j) & 1ull) != 0;
            break;
          }
        }
        if (spins >= spin_cap) {
          poison = true;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (!poison) {
        SSumm v{{0.0, 0.0}};
        if (lane == jl) v = lb_get(&incl[t - 1 - lane]);
        else if (lane < jl) v = lb_get(&agg[t - 1 - lane]);
        P = M::unpack(shfl_summ(v, jl));
        for (int l = jl - 1; l >= 0; --l) P = M::ccomb(P, M::unpack(shfl_summ(v, l)));
      }
      if (lane == 0) {
        if (poison) {
          // the tile's output is invalid: flag the error for nt2h_sync
          __hip_atomic_store(&ctrl[3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          lb_state(&state[t], tag | 3ull);
        } else {
          lb_put(&incl[t], M::pack(M::ccomb(P, L[kLbSub])));
          __builtin_amdgcn_s_waitcnt(0);
          lb_state(&state[t], tag | 2ull);
        }
      }
    }
    if (lane == 0) sh_p = M::pack(P);
  }
  __syncthreads();
  const C P = M::unpack(sh_p);
#pragma unroll
  for (int u = 0; u < kLbSub; ++u) {
    const int64_t m = m0 + u * kSub + W * (int64_t)threadIdx.x;
    V ex = M::shfl_up(inc[u], 1);
    if (lane == 0) ex = M::ident();
    ex = M::combs(wpre[u], ex);
    const C carry = M::ccomb(P, L[u]);
    VT r;
#pragma unroll
    for (int k = 0; k < W; ++k) r[k] = M::fin(carry, M::combs(ex, p[u][k]));
    if (m < m1) __builtin_nontemporal_store(r, reinterpret_cast<VT*>(out + m));
  }
  if (threadIdx.x == 0 && t == ntiles - 1) {  // every ticket is taken, every epoch read
    __hip_atomic_store(&ctrl[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ctrl[2], epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Launch plan of a scan of (lo, mid, hi) elements of src into the table y
// on stream s.  Env supplies the stream's scratch and look-back state and
// reports HIP errors: int env.scratch(size_t, void**), int env.lookback(
// size_t, void**), void env.lookback_wait(unsigned* spin_cap, unsigned*
// epoch_mask), int env.hip(hipError_t).
template <int OP, class T, class Src, class Env>
int scan_launch(const Src& src, T* y, int64_t lo, int64_t mid, int64_t hi, hipStream_t s, Env& env) {
  constexpr int W = Vec16<T>::n;
  const int64_t nout = lo * hi;
  SGeo<T, Src> g{src, y, lo, mid, hi};
  const bool y16 = (reinterpret_cast<uintptr_t>(y) & 15u) == 0;
  if (lo == 1 && mid >= 512) {
    // one workgroup per (run, segment): about 8192 of them, segments of at
    // least 8 tiles
    const bool vec = (mid % W == 0) && src.aligned() && y16;
    const int64_t tile = vec ? (int64_t)W * kBlock : kBlock;
    int64_t segs = (8192 + nout - 1) / nout;
    int64_t maxsegs = (mid + 8 * tile - 1) / (8 * tile);
    if (segs > maxsegs) segs = maxsegs;
    if (segs < 1) segs = 1;
    int64_t seglen = (mid + segs - 1) / segs;
    seglen = (seglen + tile - 1) / tile * tile;
    segs = (mid + seglen - 1) / seglen;
    const int64_t blocks = nout * segs;
    if (blocks > 0x7fffffffLL) return NT2H_ERR_TOO_LARGE;
    SSumm* summ = nullptr;
#define NT2H_SCAN_RUN(PASS)                                                                                      \
  do {                                                                                                            \
    if (vec) hipLaunchKernelGGL((scan_run_kernel<OP, T, true, PASS, Src>), dim3((unsigned)blocks), dim3(kBlock), \
                                0, s, g, segs, seglen, summ);                                                     \
    else hipLaunchKernelGGL((scan_run_kernel<OP, T, false, PASS, Src>), dim3((unsigned)blocks), dim3(kBlock), 0, \
                            s, g, segs, seglen, summ);                                                            \
    if (int st_ = env.hip(hipGetLastError())) return st_;                                                         \
  } while (0)
    if (segs == 1) {
      NT2H_SCAN_RUN(0);
      return NT2H_OK;
    }
    // up to kLbMaxTiles look-back tiles per run and kLbMaxTotal in all, with
    // 16-byte vectors: one pass with decoupled look-back (beyond: three
    // passes, see kLbMaxTiles)
    const int64_t t4 = (mid + 4 * kBlock * W - 1) / (4 * kBlock * W);  // 16 KiB tiles per run
    const int sub = (t4 > 64 || nout * t4 > 256) ? 8 : 4;
    const int64_t lbtile = (int64_t)sub * kBlock * W;
    const int64_t tpr = (mid + lbtile - 1) / lbtile;
    // cumprod's 32 KiB tiles hold 256 registers per lane (one wave per SIMD):
    // half the limits (measured: three passes win from 128 such tiles on)
    const bool heavy = OP == NT2H_SCAN_CUMPROD && sub == 8;
    const int64_t max_run = heavy ? kLbMaxTiles / 2 : kLbMaxTiles, max_all = heavy ? kLbMaxTotal / 2 : kLbMaxTotal;
    if (vec && tpr <= max_run && nout * tpr <= max_all) {
      const int64_t ntiles = nout * tpr;
      const size_t need = 256 + (size_t)ntiles * (sizeof(unsigned long long) + 2 * sizeof(SSumm));
      void* base = nullptr;
      if (int st = env.lookback(need, &base)) return st;
      unsigned* ctrl = reinterpret_cast<unsigned*>(base);
      unsigned long long* lbst = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(base) + 256);
      SSumm* agg = reinterpret_cast<SSumm*>(lbst + ntiles);
      SSumm* incl = agg + ntiles;
      unsigned cap = 0, emask = 0;
      env.lookback_wait(&cap, &emask);
      if (cap == 0) cap = kLbSpinCap;
      if (sub == 8)
        hipLaunchKernelGGL((scan_lookback_kernel<OP, T, 8, Src>), dim3((unsigned)ntiles), dim3(kBlock), 0, s, g,
                           tpr, ntiles, ctrl, lbst, agg, incl, cap, emask);
      else
        hipLaunchKernelGGL((scan_lookback_kernel<OP, T, 4, Src>), dim3((unsigned)ntiles), dim3(kBlock), 0, s, g,
                           tpr, ntiles, ctrl, lbst, agg, incl, cap, emask);
      return env.hip(hipGetLastError());
    }
    if (int st = env.scratch((size_t)blocks * sizeof(SSumm), (void**)&summ)) return st;
    NT2H_SCAN_RUN(1);
    hipLaunchKernelGGL((scan_prefix_kernel<OP, T>), dim3((unsigned)nout), dim3(kPrefixBlock), 0, s, segs, summ,
                       segs, (int64_t)1);
    if (int st = env.hip(hipGetLastError())) return st;
    NT2H_SCAN_RUN(3);
#undef NT2H_SCAN_RUN
    return NT2H_OK;
  }
  const bool colv = (lo % W == 0) && src.aligned() && y16;
  // narrow tables (at most 16 groups of W outputs: the column kernel runs
  // out of lanes below that, and from 32 groups on it is faster): the
  // row-phase tile kernel
  const bool tile = colv && nout / W <= 16;
  int lgP = 0;
  int64_t qblocks = 0, segs, seglen, blocks;
  int64_t so = 1, ss = 0;  // segment total of (o, s) at summ[o * so + s * ss], set below
  if (tile) {
    // about 16384 waves (the shuffle scan across the phases makes the waves
    // latency-bound: measured fastest between 8192 and 32768), at least
    // four steps of R rows per segment
    const int64_t nq = nout / W;
    while (((int64_t)1 << lgP) < nq) ++lgP;
    const int64_t R = 64 >> lgP;
    qblocks = (nq + ((int64_t)1 << lgP) - 1) >> lgP;  // 1: all groups in one wave
    segs = 16384 / qblocks;
    const int64_t maxsegs = (mid + 4 * R - 1) / (4 * R);
    if (segs > maxsegs) segs = maxsegs;
    if (segs < 1) segs = 1;
    seglen = (mid + segs - 1) / segs;
    segs = (mid + seglen - 1) / seglen;
    blocks = (qblocks * segs + kWaves - 1) / kWaves;
  } else {
    const int64_t units = colv ? nout / W : nout;  // lanes per segment
    // 16-byte lanes: about 2^17 of them, four loads each in flight (8 MiB),
    // measured to stream strided rows fastest; scalar lanes: 2^18
    segs = ((int64_t)1 << (colv ? 17 : 18)) / units;
    int64_t maxsegs = (mid + 63) / 64;
    if (segs > maxsegs) segs = maxsegs;
    // (up to 2^17 segments: tables with a few outputs per row need them for
    // the lanes — 3 x 2^24 floats: sum 1085 -> 187 us against a 4096 cap)
    if (segs > ((int64_t)1 << 17)) segs = (int64_t)1 << 17;
    if (segs < 1) segs = 1;
    seglen = (mid + segs - 1) / segs;
    segs = (mid + seglen - 1) / seglen;
    blocks = (units * segs + kBlock - 1) / kBlock;
  }
  if (blocks > 0x7fffffffLL || (segs > 1 && nout > 0x7fffffffLL)) return NT2H_ERR_TOO_LARGE;
  SSumm* summ = nullptr;
#define NT2H_SCAN_COL(PASS)                                                                                      \
  do {                                                                                                            \
    if (tile) hipLaunchKernelGGL((scan_tile_kernel<OP, T, PASS, Src>), dim3((unsigned)blocks), dim3(kBlock), 0, s, \
                                 g, lgP, qblocks, segs, seglen, summ, so, ss);                                    \
    else if (colv) hipLaunchKernelGGL((scan_colv_kernel<OP, T, PASS, Src>), dim3((unsigned)blocks), dim3(kBlock),  \
                                      0, s, g, segs, seglen, summ, so, ss);                                       \
    else hipLaunchKernelGGL((scan_col_kernel<OP, T, PASS, Src>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g,    \
                            segs, seglen, summ, so, ss);                                                          \
    if (int st_ = env.hip(hipGetLastError())) return st_;                                                         \
  } while (0)
  if (segs == 1) {
    NT2H_SCAN_COL(0);
    return NT2H_OK;
  }
  if (int st = env.scratch((size_t)(nout * segs) * sizeof(SSumm), (void**)&summ)) return st;
  // segment totals segment-major (s * nout + o) for one lane per output,
  // output-major (o * segs + s) for one workgroup per output, so that the
  // prefix pass reads them contiguously
  const bool lane_per_output = nout >= 8192 || segs <= 16;
  so = lane_per_output ? 1 : segs;
  ss = lane_per_output ? nout : 1;
  NT2H_SCAN_COL(1);
  if (lane_per_output)
    hipLaunchKernelGGL((scan_col_prefix_kernel<OP, T>), dim3((unsigned)((nout + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, nout, segs, summ);
  else
    hipLaunchKernelGGL((scan_prefix_kernel<OP, T>), dim3((unsigned)nout), dim3(kPrefixBlock), 0, s, segs, summ,
                       segs, (int64_t)1);
  if (int st = env.hip(hipGetLastError())) return st;
  NT2H_SCAN_COL(3);
#undef NT2H_SCAN_COL
  return NT2H_OK;
}

// A scan over (lo, mid, hi) with the argument checks of the C ABI: op is an
// nt2h_scan_op; have_src says whether the source is present.
template <class T, class Src, class Env>
int scan_entry(int op, const Src& src, bool have_src, int64_t lo, int64_t mid, int64_t hi, T* y, hipStream_t s,
               Env& env) {
  if (op < NT2H_SCAN_CUMSUM || op > NT2H_SCAN_CUMMIN) return NT2H_ERR_BAD_OP;
  if (lo < 0 || mid < 0 || hi < 0) return NT2H_ERR_INVALID_ARG;
  if (lo > 0 && hi > 0 && (mid > INT64_MAX / lo || lo * mid > INT64_MAX / hi)) return NT2H_ERR_TOO_LARGE;
  if (lo * mid * hi == 0) return NT2H_OK;
  if (!have_src || !y) return NT2H_ERR_NULL_POINTER;
  switch (op) {
    case NT2H_SCAN_CUMSUM: return scan_launch<NT2H_SCAN_CUMSUM, T>(src, y, lo, mid, hi, s, env);
    case NT2H_SCAN_CUMPROD: return scan_launch<NT2H_SCAN_CUMPROD, T>(src, y, lo, mid, hi, s, env);
    case NT2H_SCAN_CUMMAX: return scan_launch<NT2H_SCAN_CUMMAX, T>(src, y, lo, mid, hi, s, env);
    case NT2H_SCAN_CUMMIN: return scan_launch<NT2H_SCAN_CUMMIN, T>(src, y, lo, mid, hi, s, env);
  }
  return NT2H_ERR_BAD_OP;
}

}  // namespace nt2h
