// Kernels and launch plans of the reduction_ functors of NT2 (sum, prod,
// maximum, minimum, asum1, asum2, norm2, mean) over one dimension of a float
// or double table or of a table expression evaluated per element (shared by
// libnt2hip's nt2h_fold_f32/f64 and the C++ front end).
//
// NT2 reshapes a reduction along dimension k into (lo, mid, hi) = (product
// of the dimensions before k, dimension k, product of those after) and runs
// inner_fold (lo == 1) or outer_fold (lo > 1) over the middle index
// (core/reduction/include/nt2/dsl/functions/reduction/run.hpp:48-93;
// scalar/inner_fold.hpp:37-52, scalar/outer_fold.hpp:22-52): out(i, h) is
// the left-to-right fold of x(i, m, h), m = 0..mid-1, from the functor's
// neutral element (details/fold_step.hpp:31-43).
//
// Kernels (one or two launches per call):
//  * run kernel (lo == 1, mid >= 512): one workgroup per (run, segment); the
//    lanes stride over the segment (16-byte loads when aligned), and the
//    workgroup combines its lanes in a fixed LDS tree;
//  * column kernels (lo > 1, or short runs): one lane per (output, segment)
//    folding sequentially — or, when lo is a multiple of the 16-byte width
//    W (4 floats, 2 doubles), the tile kernel: lanes own groups of W
//    adjacent outputs with one 16-byte load per row, and the row phases of a
//    workgroup share a segment and combine in LDS;
//    neighbouring lanes read neighbouring addresses;
//  * combine kernels: when a run is split into segments, one workgroup per
//    output combines the segment summaries (each lane a contiguous range of
//    segments, then a fixed tree: the same order on every call); for the
//    column paths one lane per output combines them in segment order, or,
//    with few outputs and many segments, one workgroup per output.
// Arithmetic per functor family:
//  * sum, asum1, asum2, norm2, mean: TwoSum-compensated lanes in the table's
//    type (asum2 adds the exact error of each square), compensated FP64
//    pairs across segments; the total then takes NT2's last steps: sqrt for
//    norm2, rec(T(mid)) * total for mean (expr/mean.hpp:22-23);
//  * prod: an FP64 mantissa with a separate integer exponent (double
//    factors enter as their frexp mantissa), renormalised every four
//    factors, so that no partial product overflows or underflows; the
//    result is rounded to the table's type once;
//  * maximum, minimum: NT2's selection a > b ? a : b (a < b ? a : b) folded
//    from Maxinit = -FLT_MAX (Valmax = FLT_MAX) returns, for a fold with no
//    NaN, the latest element equal to the extreme value, or the neutral if
//    it is more extreme; a NaN resets the accumulator, so the result is the
//    fold of the elements after the last NaN (or that NaN when it is the
//    last element).  Each lane keeps (index of its last NaN, latest extreme
//    and its index) — a commutative summary — and the rare case where the
//    overall extreme lies before the last NaN is finished by rescanning
//    the tail after that NaN.  Results are bit-identical to NT2's.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../nt2hip.h"
#include "fold.hpp"

namespace nt2h {

// Element source of a contiguous table.  The kernels read elements as
// at(i), or as 16-byte vectors in two steps: raw(i) issues the loads of
// elements i..i+W-1 (i a multiple of W), value(raw) does any arithmetic on
// them — for a table none, for an expression tree evaluated per element its
// nodes — so that a kernel can have several vectors' loads in flight before
// the first one's arithmetic waits on memory; vec(i) is value(raw(i)).
template <class T> struct PtrSrc {
  using Raw = typename Vec16<T>::type;
  const T* x;
  __device__ __forceinline__ T at(int64_t i) const { return x[i]; }
  __device__ __forceinline__ Raw raw(int64_t i) const {
    return __builtin_nontemporal_load(reinterpret_cast<const Raw*>(x + i));
  }
  __device__ __forceinline__ typename Vec16<T>::type value(const Raw& r) const { return r; }
  __device__ __forceinline__ typename Vec16<T>::type vec(int64_t i) const { return value(raw(i)); }
  bool aligned() const { return (reinterpret_cast<uintptr_t>(x) & 15u) == 0; }
};


enum FoldOpId {
  F_SUM = NT2H_FOLD_SUM,
  F_PROD = NT2H_FOLD_PROD,
  F_MAX = NT2H_FOLD_MAXIMUM,
  F_MIN = NT2H_FOLD_MINIMUM,
  F_ASUM1 = NT2H_FOLD_ASUM1,
  F_ASUM2 = NT2H_FOLD_ASUM2,
  F_NORM2 = NT2H_FOLD_NORM2,
  F_MEAN = NT2H_FOLD_MEAN
};

// Table geometry: element m of output o = i + lo*h sits at i + lo*(m + mid*h);
// the elements come from Src (a table, or an expression evaluated per
// element: at(i) is element i, raw(i)/value()/vec(i) the 16-byte vector of
// elements i..i+W-1 for i a multiple of W).
template <class T, class Src> struct Geo {
  Src src;
  int64_t lo, mid, hi;
  __device__ __forceinline__ int64_t at(int64_t o, int64_t m) const {
    if (lo == 1) return o * mid + m;
    const int64_t h = o / lo, i = o - h * lo;
    return i + lo * (m + mid * h);
  }
};

// Segment summaries live in scratch as 4 doubles each.
struct Summ {
  double w[4];
};

__device__ __forceinline__ float fabs_t(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double fabs_t(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float sqrt_t(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return __builtin_sqrt(x); }

// Error-free sum in either precision.
template <class T> __device__ __forceinline__ void two_sum_t(T a, T b, T& s, T& e) {
  s = a + b;
  T bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}
template <class T> struct CompT {
  T s, c;
  __device__ __forceinline__ void add(T x) {
    T t, e;
    two_sum_t(s, x, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void merge(T s2, T c2) {
    T t, e;
    two_sum_t(s, s2, t, e);
    s = t;
    c = c + c2 + e;
  }
};

// ------------------------------------------------------------ sum family
template <int OP, class T> struct FSum {
  struct L {
    CompT<T> c;
  };
  __device__ static L ident() { return L{{T(0), T(0)}}; }
  __device__ static void add(L& a, T x, int64_t) {
    if (OP == F_ASUM1) {
      a.c.add(fabs_t(x));
    } else if (OP == F_ASUM2 || OP == F_NORM2) {
      T ax = fabs_t(x);
      T p = ax * ax;
      T e = fma_t(ax, ax, -p);  // exact error of the square
      a.c.add(p);
      a.c.c += e;
    } else {
      a.c.add(x);
    }
  }
  __device__ static L comb(L a, const L& b) {
    a.c.merge(b.c.s, b.c.c);
    return a;
  }
  // summaries: a compensated FP64 pair (s, c)
  __device__ static Summ summ(const L& a) {
    if (!__builtin_isfinite(a.c.s)) return Summ{{(double)a.c.s, 0.0, 0.0, 0.0}};  // see comp_value
    double s, e;
    two_sum_t((double)a.c.s, (double)a.c.c, s, e);
    return Summ{{s, e, 0.0, 0.0}};
  }
  __device__ static Summ scomb(Summ a, const Summ& b) {
    double t, e;
    two_sum_t(a.w[0], b.w[0], t, e);
    a.w[1] = a.w[1] + b.w[1] + e;
    a.w[0] = t;
    return a;
  }
  __device__ static Summ sident() { return Summ{{0.0, 0.0, 0.0, 0.0}}; }
  __device__ static T finish(const Summ& a, int64_t mid) {
    T t = (T)comp_value(a.w[0], a.w[1]);
    if (OP == F_NORM2) return sqrt_t(t);
    if (OP == F_MEAN) {
      T f = T(1) / (T)mid;  // NT2: rec(value_type(size)) * sum
      return f * t;
    }
    return t;
  }
  static constexpr bool kRescan = false;
};

// ----------------------------------------------------------- prod family
template <class T> struct FProd {
  struct L {
    double m;
    int e;
    int k;  // factors since the last renormalisation
  };
  __device__ static void norm(L& a) {
    if (a.m != 0.0 && __builtin_isfinite(a.m)) {
      int k;
      a.m = __builtin_frexp(a.m, &k);
      a.e += k;
      // keep the exponent bounded: beyond these the result is 0 or inf in either type
      a.e = a.e > (1 << 28) ? (1 << 28) : (a.e < -(1 << 28) ? -(1 << 28) : a.e);
    }
    a.k = 0;
  }
  __device__ static L ident() { return L{1.0, 0, 0}; }
  __device__ static void add(L& a, T x, int64_t) {
    if (sizeof(T) == 8) {  // a double factor enters as its mantissa in [0.5, 1)
      double xm = (double)x;
      if (xm != 0.0 && __builtin_isfinite(xm)) {
        int k;
        xm = __builtin_frexp(xm, &k);
        a.e += k;
      }
      a.m *= xm;
    } else {
      a.m *= (double)x;  // |x| < 2^128: four factors stay inside double's range
    }
    if (++a.k == 4) norm(a);
  }
  __device__ static L comb(L a, L b) {
    norm(a);
    norm(b);
    a.m *= b.m;
    a.e += b.e;
    norm(a);
    return a;
  }
  __device__ static Summ summ(L a) {
    norm(a);
    return Summ{{a.m, (double)a.e, 0.0, 0.0}};
  }
  __device__ static Summ scomb(Summ a, const Summ& b) {
    L x{a.w[0], (int)a.w[1], 0}, y{b.w[0], (int)b.w[1], 0};
    L r = comb(x, y);
    return Summ{{r.m, (double)r.e, 0.0, 0.0}};
  }
  __device__ static Summ sident() { return Summ{{1.0, 0.0, 0.0, 0.0}}; }
  __device__ static T finish(const Summ& a, int64_t) {
    double m = a.w[0];
    int e = (int)a.w[1];
    if (m == 0.0 || !__builtin_isfinite(m)) return (T)m;
    // |m| in [0.5, 1): clamp e so that ldexp is exact or saturates the same way
    if (e > 2000) e = 2000;
    if (e < -2000) e = -2000;
    return (T)__builtin_ldexp(m, e);
  }
  static constexpr bool kRescan = false;
};

// -------------------------------------------------- maximum / minimum
template <bool MAX, class T> struct FExt {
  struct L {
    int64_t nan;  // index of the last NaN, -1 if none
    int64_t j;    // index of the latest extreme, -1 if none
    T v;
  };
  __device__ static bool better(T a, T b) { return MAX ? (a > b) : (a < b); }
  __device__ static L ident() { return L{-1, -1, T(0)}; }
  // elements arrive in increasing index order within a lane
  __device__ static void add(L& a, T x, int64_t idx) {
    if (x != x) {
      a.nan = idx;
    } else if (a.j < 0 || !better(a.v, x)) {
      a.v = x;
      a.j = idx;
    }
  }
  __device__ static L comb(L a, const L& b) {
    L r;
    r.nan = a.nan > b.nan ? a.nan : b.nan;
    bool take_b;
    if (a.j < 0) take_b = true;
    else if (b.j < 0) take_b = false;
    else if (better(a.v, b.v)) take_b = false;
    else if (better(b.v, a.v)) take_b = true;
    else take_b = b.j > a.j;  // equal values: the later element wins
    r.j = take_b ? b.j : a.j;
    r.v = take_b ? b.v : a.v;
    return r;
  }
  __device__ static Summ summ(const L& a) { return Summ{{(double)a.nan, (double)a.j, (double)a.v, 0.0}}; }
  __device__ static L unsumm(const Summ& s) { return L{(int64_t)s.w[0], (int64_t)s.w[1], (T)s.w[2]}; }
  __device__ static Summ scomb(const Summ& a, const Summ& b) { return summ(comb(unsumm(a), unsumm(b))); }
  __device__ static Summ sident() { return summ(ident()); }
  // Maxinit / Valmax: the largest finite magnitude of T
  __device__ static T neutral() {
    if (sizeof(T) == 4) return (T)(MAX ? -0x1.fffffep+127f : 0x1.fffffep+127f);
    return (T)(MAX ? -0x1.fffffffffffffp+1023 : 0x1.fffffffffffffp+1023);
  }
  // no NaN: the latest extreme unless the neutral is more extreme
  __device__ static T pick(const L& a) {
    if (a.j < 0) return neutral();
    return better(neutral(), a.v) ? neutral() : a.v;
  }
  static constexpr bool kRescan = true;
};

template <int OP, class T> struct Family;
template <class T> struct Family<F_SUM, T> { using type = FSum<F_SUM, T>; };
template <class T> struct Family<F_ASUM1, T> { using type = FSum<F_ASUM1, T>; };
template <class T> struct Family<F_ASUM2, T> { using type = FSum<F_ASUM2, T>; };
template <class T> struct Family<F_NORM2, T> { using type = FSum<F_NORM2, T>; };
template <class T> struct Family<F_MEAN, T> { using type = FSum<F_MEAN, T>; };
template <class T> struct Family<F_PROD, T> { using type = FProd<T>; };
template <class T> struct Family<F_MAX, T> { using type = FExt<true, T>; };
template <class T> struct Family<F_MIN, T> { using type = FExt<false, T>; };

// Fixed-order LDS tree over a workgroup of NB lanes; every thread receives
// the result.
template <int NB = kBlock, class S, class Comb>
__device__ __forceinline__ S block_tree(S v, S* sh, Comb comb) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = comb(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  S r = sh[0];
  __syncthreads();
  return r;
}

// Finish an extreme-value fold of output o from its summary a; all threads
// of the workgroup call this together (uniform control flow) and thread 0
// gets the result.
template <class F, class T, class Src, int NB = kBlock>
__device__ T ext_finish_block(const Geo<T, Src>& g, int64_t o, typename F::L a, typename F::L* sh) {
  const int64_t mid = g.mid;
  if (a.nan < 0) return F::pick(a);
  if (a.nan == mid - 1) return g.src.at(g.at(o, a.nan));
  if (a.j > a.nan) return a.v;
  // the extreme lies before the last NaN: fold the tail after it
  typename F::L t = F::ident();
  for (int64_t m = a.nan + 1 + threadIdx.x; m < mid; m += NB) F::add(t, g.src.at(g.at(o, m)), m);
  t = block_tree<NB>(t, sh, [](typename F::L p, typename F::L q) { return F::comb(p, q); });
  return t.v;  // the tail holds no NaN and is not empty
}
template <class F, class T, class Src>
__device__ T ext_finish_thread(const Geo<T, Src>& g, int64_t o, const typename F::L& a) {
  const int64_t mid = g.mid;
  if (a.nan < 0) return F::pick(a);
  if (a.nan == mid - 1) return g.src.at(g.at(o, a.nan));
  if (a.j > a.nan) return a.v;
  typename F::L t = F::ident();
  for (int64_t m = a.nan + 1; m < mid; ++m) F::add(t, g.src.at(g.at(o, m)), m);
  return t.v;
}

// ------------------------------------------------------------ kernels
// Run kernel: lo == 1; block b -> (run o = b / segs, segment s = b % segs).
template <int OP, class T, bool VEC, class Src>
__global__ __launch_bounds__(kBlock) void fold_run_kernel(Geo<T, Src> g, int64_t segs, int64_t seglen, T* out,
                                                          Summ* summ) {
  using F = typename Family<OP, T>::type;
  using L = typename F::L;
  using V = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::n;
  __shared__ L sh[kBlock];
  const int64_t b = blockIdx.x;
  const int64_t o = b / segs, s = b - o * segs;
  const int64_t m0 = s * seglen;
  const int64_t m1 = (m0 + seglen < g.mid) ? m0 + seglen : g.mid;
  const int64_t run = o * g.mid;
  L a = F::ident();
  if (VEC) {  // run and segment start are 16-byte aligned, seglen % W == 0
    const int64_t nv = (m1 - m0) / W;
    for (int64_t k = threadIdx.x; k < nv; k += kBlock) {
      V q = g.src.vec(run + m0 + W * k);
      const int64_t m = m0 + W * k;
#pragma unroll
      for (int e = 0; e < W; ++e) F::add(a, q[e], m + e);
    }
    for (int64_t m = m0 + W * nv + threadIdx.x; m < m1; m += kBlock) F::add(a, g.src.at(run + m), m);
  } else {
    for (int64_t m = m0 + threadIdx.x; m < m1; m += kBlock) F::add(a, g.src.at(run + m), m);
  }
  a = block_tree(a, sh, [](L p, L q) { return F::comb(p, q); });
  if (segs > 1) {
    if (threadIdx.x == 0) summ[o * segs + s] = F::summ(a);
    return;
  }
  T r;
  if constexpr (F::kRescan) r = ext_finish_block<F, T, Src>(g, o, a, sh);
  else r = F::finish(F::summ(a), g.mid);
  if (threadIdx.x == 0) out[o] = r;
}

// Column kernel: one lane per (output, segment); consecutive lanes take
// consecutive outputs of the same segment.
template <int OP, class T, class Src>
__global__ __launch_bounds__(kBlock) void fold_col_kernel(Geo<T, Src> g, int64_t segs, int64_t seglen, T* out,
                                                          Summ* summ) {
  using F = typename Family<OP, T>::type;
  using L = typename F::L;
  const int64_t nout = g.lo * g.hi;
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nout * segs) return;
  const int64_t s = t / nout, o = t - s * nout;
  const int64_t m0 = s * seglen;
  const int64_t m1 = (m0 + seglen < g.mid) ? m0 + seglen : g.mid;
  L a = F::ident();
  for (int64_t m = m0; m < m1; ++m) F::add(a, g.src.at(g.at(o, m)), m);
  if (segs > 1) {
    summ[s * nout + o] = F::summ(a);  // segment-major: coalesced for the combine
    return;
  }
  if constexpr (F::kRescan) out[o] = ext_finish_thread<F, T, Src>(g, o, a);
  else out[o] = F::finish(F::summ(a), g.mid);
}

// Column kernel, 16-byte form (lo % W == 0, x 16-byte aligned; W = 4 floats
// or 2 doubles).  A workgroup is R = kBlock / P row phases of P lanes (P a
// power of two <= 64): lane (r, p) folds the W adjacent outputs of group
// q = qb * P + p — one 16-byte vector per row — over rows m0 + r, m0 + r + R,
// ... of its segment, four loads in flight.  So one wave reads 1 KiB of
// consecutive addresses per load (64 groups of one row, or 64 / P adjacent
// rows when the rows are short), and the R phases of the workgroup walk down
// the segment side by side.  Their R partial summaries per output are then
// combined in LDS by a fixed tree, so that a segment needs one summary per
// output however many lanes folded it.
template <int OP, class T, class Src>
__global__ __launch_bounds__(kBlock) void fold_tile_kernel(Geo<T, Src> g, int lgP, int64_t qblocks, int64_t segs,
                                                           int64_t seglen, T* out, Summ* summ) {
  using F = typename Family<OP, T>::type;
  using L = typename F::L;
  using V = typename Vec16<T>::type;
  constexpr int W = Vec16<T>::n;
  const int P = 1 << lgP, R = kBlock >> lgP;
  const int p = threadIdx.x & (P - 1), r = threadIdx.x >> lgP;
  const int64_t nq = g.lo * g.hi / W;
  const int64_t s = blockIdx.x / qblocks, q = (blockIdx.x - s * qblocks) * P + p;
  const bool live = q < nq;
  const int64_t o = W * q;  // first of the W outputs; the group never straddles h since lo % W == 0
  const int64_t h = o / g.lo, i = o - h * g.lo;
  const int64_t base = i + g.lo * g.mid * h;  // element (i, 0, h); rows m and m + 1 are lo apart
  const int64_t m0 = s * seglen;
  const int64_t m1 = !live ? m0 : (m0 + seglen < g.mid) ? m0 + seglen : g.mid;
  L a[W];
#pragma unroll
  for (int k = 0; k < W; ++k) a[k] = F::ident();
  int64_t m = m0 + r;
  for (; m + 3 * R < m1; m += 4 * R) {
    typename Src::Raw w[4];  // every load in flight before the first value
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = g.src.raw(base + (m + u * R) * g.lo);
    V v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = g.src.value(w[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int k = 0; k < W; ++k) F::add(a[k], v[u][k], m + u * R);
    }
  }
  for (; m < m1; m += R) {
    V v = g.src.vec(base + m * g.lo);
#pragma unroll
    for (int k = 0; k < W; ++k) F::add(a[k], v[k], m);
  }
  Summ acc[W];
#pragma unroll
  for (int k = 0; k < W; ++k) acc[k] = F::summ(a[k]);
  __shared__ Summ sh[kBlock / 2 * W];
  for (int half = R >> 1; half >= 1; half >>= 1) {
    if (r >= half && r < 2 * half) {
#pragma unroll
      for (int k = 0; k < W; ++k) sh[((r - half) * P + p) * W + k] = acc[k];
    }
    __syncthreads();
    if (r < half) {
#pragma unroll
      for (int k = 0; k < W; ++k) acc[k] = F::scomb(acc[k], sh[(r * P + p) * W + k]);
    }
    __syncthreads();
  }
  if (r != 0 || !live) return;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    if (segs > 1) {
      summ[s * (g.lo * g.hi) + o + k] = acc[k];
    } else {
      if constexpr (F::kRescan) out[o + k] = ext_finish_thread<F, T, Src>(g, o + k, F::unsumm(acc[k]));
      else out[o + k] = F::finish(acc[k], g.mid);
    }
  }
}

// Fold the summaries base[s * st], s = s0..s1-1, into a in order, loading
// eight at a time so that their latencies overlap.
template <class F>
__device__ __forceinline__ Summ scomb_range(Summ a, const Summ* base, int64_t s0, int64_t s1, int64_t st) {
  int64_t s = s0;
  for (; s + 8 <= s1; s += 8) {
    Summ b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) b[u] = base[(s + u) * st];
#pragma unroll
    for (int u = 0; u < 8; ++u) a = F::scomb(a, b[u]);
  }
  for (; s < s1; ++s) a = F::scomb(a, base[s * st]);
  return a;
}

// Combine kernel: one workgroup of NB lanes per output (NB = 1024 for the
// many segments of a whole-table fold); lane l takes a contiguous range of
// segments in order, the lanes then combine in a fixed tree (the sum and
// product families are deterministic but not sequential in rounding; the
// extreme-value summaries are order-free).
template <int OP, class T, int NB, class Src>
__global__ __launch_bounds__(NB) void fold_combine_kernel(Geo<T, Src> g, int64_t segs, T* out, const Summ* summ) {
  using F = typename Family<OP, T>::type;
  __shared__ Summ sh[NB];
  const int64_t o = blockIdx.x;
  const int64_t per = (segs + NB - 1) / NB;
  const int64_t s0 = threadIdx.x * per;
  const int64_t s1 = (s0 + per < segs) ? s0 + per : segs;
  Summ a = F::sident();
  a = scomb_range<F>(a, summ + o * segs, s0, s1, 1);
  a = block_tree<NB>(a, sh, [](Summ p, Summ q) { return F::scomb(p, q); });
  T r;
  if constexpr (F::kRescan) {
    __shared__ typename F::L shl[NB];
    r = ext_finish_block<F, T, Src, NB>(g, o, F::unsumm(a), shl);
  } else {
    r = F::finish(a, g.mid);
  }
  if (threadIdx.x == 0) out[o] = r;
}

// Combine kernel of the column paths: one lane per output folds its segment
// summaries (segment-major, so neighbouring lanes read neighbouring
// summaries) in segment order.
template <int OP, class T, class Src>
__global__ __launch_bounds__(kBlock) void fold_col_combine_kernel(Geo<T, Src> g, int64_t segs, T* out,
                                                                  const Summ* summ) {
  using F = typename Family<OP, T>::type;
  const int64_t nout = g.lo * g.hi;
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= nout) return;
  Summ a = summ[o];
  a = scomb_range<F>(a, summ + o, 1, segs, nout);
  if constexpr (F::kRescan) out[o] = ext_finish_thread<F, T, Src>(g, o, F::unsumm(a));
  else out[o] = F::finish(a, g.mid);
}

// The same with one workgroup per output, for few outputs and many segments:
// lane l folds a contiguous range of segments in order, then a fixed tree.
template <int OP, class T, class Src>
__global__ __launch_bounds__(kBlock) void fold_col_combine_block_kernel(Geo<T, Src> g, int64_t segs, T* out,
                                                                        const Summ* summ) {
  using F = typename Family<OP, T>::type;
  __shared__ Summ sh[kBlock];
  const int64_t nout = g.lo * g.hi;
  const int64_t o = blockIdx.x;
  const int64_t per = (segs + kBlock - 1) / kBlock;
  const int64_t s0 = threadIdx.x * per;
  const int64_t s1 = (s0 + per < segs) ? s0 + per : segs;
  Summ a = F::sident();
  a = scomb_range<F>(a, summ + o, s0, s1, nout);
  a = block_tree(a, sh, [](Summ p, Summ q) { return F::scomb(p, q); });
  T r;
  if constexpr (F::kRescan) {
    __shared__ typename F::L shl[kBlock];
    r = ext_finish_block<F, T, Src>(g, o, F::unsumm(a), shl);
  } else {
    r = F::finish(a, g.mid);
  }
  if (threadIdx.x == 0) out[o] = r;
}

// Output for an empty middle extent: NT2 folds nothing from the neutral
// (mean: rec(0) * 0 = NaN).
template <class T> __global__ void fold_fill_kernel(T* out, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = v;
}

// Launch plan of a fold of (lo, mid, hi) elements of src into out (lo*hi
// values) on stream s.  Env supplies the stream's scratch and reports HIP
// errors: int env.scratch(size_t bytes, void** p), int env.hip(hipError_t).
template <int OP, class T, class Src, class Env>
int fold_launch(const Src& src, int64_t lo, int64_t mid, int64_t hi, T* out, hipStream_t s, Env& env) {
  constexpr int W = Vec16<T>::n;
  const int64_t nout = lo * hi;
  Geo<T, Src> g{src, lo, mid, hi};
  auto launched = [&]() { return env.hip(hipGetLastError()); };
  if (lo == 1 && mid >= 512) {
    // enough workgroups to fill the chip; segments of at least 2048 elements
    int64_t segs = (8192 + hi - 1) / hi;
    int64_t maxsegs = (mid + 2047) / 2048;
    if (segs > maxsegs) segs = maxsegs;
    if (segs < 1) segs = 1;
    int64_t seglen = (mid + segs - 1) / segs;
    seglen = (seglen + W - 1) / W * W;
    segs = (mid + seglen - 1) / seglen;
    const bool vec = (mid % W == 0) && src.aligned();
    Summ* summ = nullptr;
    const int64_t blocks = nout * segs;
    if (blocks > 0x7fffffffLL) return NT2H_ERR_TOO_LARGE;
    if (segs > 1) {
      int st = env.scratch((size_t)(nout * segs) * sizeof(Summ), (void**)&summ);
      if (st != NT2H_OK) return st;
    }
    if (vec) hipLaunchKernelGGL((fold_run_kernel<OP, T, true, Src>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, segs, seglen, out, summ);
    else hipLaunchKernelGGL((fold_run_kernel<OP, T, false, Src>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, segs, seglen, out, summ);
    if (int st = launched()) return st;
    if (segs > 1) {
      if (segs >= 4096)
        hipLaunchKernelGGL((fold_combine_kernel<OP, T, 1024, Src>), dim3((unsigned)nout), dim3(1024), 0, s, g, segs,
                           out, (const Summ*)summ);
      else
        hipLaunchKernelGGL((fold_combine_kernel<OP, T, kBlock, Src>), dim3((unsigned)nout), dim3(kBlock), 0, s, g,
                           segs, out, (const Summ*)summ);
      if (int st = launched()) return st;
    }
    return NT2H_OK;
  }
  const bool colv = (lo % W == 0) && src.aligned();
  int64_t segs, seglen, blocks;
  int lgP = 0;
  int64_t qblocks = 0;
  if (colv) {
    // tile kernel: about 2^17 lanes with four 16-byte loads each in flight
    // (8 MiB) — measured to stream strided rows fastest, ahead of twice or
    // half that — and each lane at least four rows of its segment
    const int64_t nq = nout / W;
    while (lgP < 6 && ((int64_t)1 << lgP) < nq) ++lgP;
    const int64_t R = kBlock >> lgP;
    qblocks = (nq + ((int64_t)1 << lgP) - 1) >> lgP;
    segs = ((int64_t)1 << 17) / (qblocks * kBlock);
    const int64_t maxsegs = (mid + 4 * R - 1) / (4 * R);
    if (segs > maxsegs) segs = maxsegs;
    if (segs < 1) segs = 1;
    seglen = (mid + segs - 1) / segs;
    segs = (mid + seglen - 1) / seglen;
    blocks = qblocks * segs;
  } else {
    // scalar column kernel: aim for >= 2^18 lanes; segments of at least 16
    // elements
    segs = ((int64_t)1 << 18) / nout;
    int64_t maxsegs = (mid + 15) / 16;
    if (segs > maxsegs) segs = maxsegs;
    // (up to 2^17 segments: tables with a few outputs per row need them for
    // the lanes — 3 x 2^24 floats: sum 1085 -> 187 us against a 4096 cap)
    if (segs > ((int64_t)1 << 17)) segs = (int64_t)1 << 17;
    if (segs < 1) segs = 1;
    seglen = (mid + segs - 1) / segs;
    segs = (mid + seglen - 1) / seglen;
    blocks = (nout * segs + kBlock - 1) / kBlock;
  }
  Summ* summ = nullptr;
  if (segs > 1) {
    if (nout > 0x7fffffffLL) return NT2H_ERR_TOO_LARGE;
  }
  if (blocks > 0x7fffffffLL) return NT2H_ERR_TOO_LARGE;
  if (segs > 1) {
    int st = env.scratch((size_t)(nout * segs) * sizeof(Summ), (void**)&summ);
    if (st != NT2H_OK) return st;
  }
  if (colv) hipLaunchKernelGGL((fold_tile_kernel<OP, T, Src>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, lgP, qblocks, segs, seglen, out, summ);
  else hipLaunchKernelGGL((fold_col_kernel<OP, T, Src>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, segs, seglen, out, summ);
  if (int st = launched()) return st;
  if (segs > 1) {
    if (nout >= 8192 || segs <= 16) {  // enough outputs for one lane each
      const int64_t cblocks = (nout + kBlock - 1) / kBlock;
      hipLaunchKernelGGL((fold_col_combine_kernel<OP, T, Src>), dim3((unsigned)cblocks), dim3(kBlock), 0, s, g, segs,
                         out, (const Summ*)summ);
    } else {
      hipLaunchKernelGGL((fold_col_combine_block_kernel<OP, T, Src>), dim3((unsigned)nout), dim3(kBlock), 0, s, g, segs,
                         out, (const Summ*)summ);
    }
    if (int st = launched()) return st;
  }
  return NT2H_OK;
}

template <class T> inline T fold_neutral_result(int op) {
  const bool dbl = sizeof(T) == 8;
  switch (op) {
    case NT2H_FOLD_PROD: return T(1);
    case NT2H_FOLD_MAXIMUM: return dbl ? (T)-0x1.fffffffffffffp+1023 : (T)-0x1.fffffep+127f;
    case NT2H_FOLD_MINIMUM: return dbl ? (T)0x1.fffffffffffffp+1023 : (T)0x1.fffffep+127f;
    case NT2H_FOLD_MEAN: return (T)__builtin_nan("");
    default: return T(0);
  }
}

// A fold over (lo, mid, hi) with the argument checks of the C ABI:
// op is an nt2h_fold_op; have_src says whether the source is present.
template <class T, class Src, class Env>
int fold_entry(int op, const Src& src, bool have_src, int64_t lo, int64_t mid, int64_t hi, T* out, hipStream_t s,
               Env& env) {
  if (op < NT2H_FOLD_SUM || op > NT2H_FOLD_MEAN) return NT2H_ERR_BAD_OP;
  if (lo < 0 || mid < 0 || hi < 0) return NT2H_ERR_INVALID_ARG;
  if (lo > 0 && hi > 0 && (mid > INT64_MAX / lo || lo * mid > INT64_MAX / hi)) return NT2H_ERR_TOO_LARGE;
  const int64_t nout = lo * hi;
  if (nout == 0) return NT2H_OK;
  if (!out || (mid > 0 && !have_src)) return NT2H_ERR_NULL_POINTER;
  if (mid == 0) {
    unsigned blocks = (unsigned)((nout + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(fold_fill_kernel<T>, dim3(blocks), dim3(kBlock), 0, s, out, nout, fold_neutral_result<T>(op));
    return env.hip(hipGetLastError());
  }
  switch (op) {
    case NT2H_FOLD_SUM: return fold_launch<F_SUM, T>(src, lo, mid, hi, out, s, env);
    case NT2H_FOLD_PROD: return fold_launch<F_PROD, T>(src, lo, mid, hi, out, s, env);
    case NT2H_FOLD_MAXIMUM: return fold_launch<F_MAX, T>(src, lo, mid, hi, out, s, env);
    case NT2H_FOLD_MINIMUM: return fold_launch<F_MIN, T>(src, lo, mid, hi, out, s, env);
    case NT2H_FOLD_ASUM1: return fold_launch<F_ASUM1, T>(src, lo, mid, hi, out, s, env);
    case NT2H_FOLD_ASUM2: return fold_launch<F_ASUM2, T>(src, lo, mid, hi, out, s, env);
    case NT2H_FOLD_NORM2: return fold_launch<F_NORM2, T>(src, lo, mid, hi, out, s, env);
    case NT2H_FOLD_MEAN: return fold_launch<F_MEAN, T>(src, lo, mid, hi, out, s, env);
  }
  return NT2H_ERR_BAD_OP;
}

}  // namespace nt2h
