// nt2hip/table.hpp — Matlab-like device tables for C++ callers (compile with hipcc).
//
// The caller writes NT2's table expressions and gets NT2's results:
//
//   nt2hip::table<float> X(n), Y(n);            // device-resident, column-major, up to 4-D
//   Y = 2.5f * X + Y;                            // one launch: nt2h_saxpy_f32
//   Y = exp(-sqr(X) / 2.0f) * cos(X);            // one fused kernel for the whole tree
//   float s = globalsum(X * Y);                  // nt2h_dot_f32
//   table<float> z = sum(exp(-sqr(A)) * cos(sx(k) * A), 1);   // a fold of the tree, no temporary
//   table<float> m = maximum(abs(A - B), 3);     // every reduction, any dimension
//   table<float> c = cumsum(exp(A), 2);          // scans of trees too
//   tie(call, put) = blackscholes(S, K, T, r, sig);           // 5 -> 2 streams, one launch
//   tie(da, d1, d2, R) = tie(sqrt(T), ..., fnms(v, da, d1), ...); // N outputs, one kernel
//
// It stands where NT2's container DSL does: table<T,S> (core/container/
// table/table.hpp:34-171; up to 4 dimensions, sdk/config/include/nt2/sdk/
// parameters.hpp:21), table::operator= -> evaluate -> run_assign
// (core/container/dsl/expression.hpp:161-191, dsl/functions/elementwise/
// run.hpp:21-35), the transform loop nest (core/functions/transform.hpp:
// 55-69), terminal load/store (dsl/functions/container/terminal.hpp:28-129),
// the per-node recursion run(node, i) (dsl/functions/generic/run.hpp:
// 128-222), the multiply-adds fma/fms/fnms/fnma (boost/simd/arithmetic/
// functions/generic/{fma,fms,fnms,fnma}.hpp), the N-ary tie (core/functions/
// expr/tie.hpp:23-160), the reductions and scans of an expression along any
// dimension (dsl/functions/reduction/run.hpp:30-104) and singleton
// expansion (core/functions/sx.hpp).
//
// Lowering: an assignment is first matched against the loop nests that
// libnt2hip implements as hand-written kernels (saxpy, one functor of a
// terminal, globalsum of a terminal or of a product of two terminals,
// Black-Scholes); any other tree becomes one kernel instantiated for that
// tree from the same device functors (include/nt2hip/math.hpp): every node
// is evaluated in registers at the element's index, nothing is
// materialised between nodes.  A reduction or scan of a tree instantiates
// the library's fold and scan kernels (include/nt2hip/{fold,scan}_
// kernels.hpp) with the tree as their element source, evaluated where they
// load.  Arithmetic follows NT2's scalar path (nt2hip::exact, the default)
// or the FMA policy (nt2hip::fused) via assign<fused>(lhs, expr).  Errors
// throw nt2hip::error.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../nt2hip.h"
#include "fold_kernels.hpp"
#include "math.hpp"
#include "scan_kernels.hpp"

namespace nt2hip {

class error : public std::runtime_error {
 public:
  int code;
  error(const std::string& what, int c) : std::runtime_error(what + ": " + nt2h_strerror(c)), code(c) {}
};
inline void check(int status, const char* what) {
  if (status != NT2H_OK) throw error(what, status);
}
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw error(std::string(what) + " (" + hipGetErrorString(e) + ")", NT2H_ERR_HIP);
}

using exact = nt2h::Exact;
using fused = nt2h::Fused;

// The stream every table operation is issued on (default: the null stream).
inline hipStream_t& current_stream() {
  static hipStream_t s = nullptr;
  return s;
}
inline nt2h_stream_t abi_stream() { return reinterpret_cast<nt2h_stream_t>(current_stream()); }

// ============================================================ extents
constexpr int kMaxDims = 4;

// The extent of a loop nest as the kernels see it: column-major, element
// (c0, c1, c2, c3) at linear index c0 + d0 (c1 + d1 (c2 + d2 c3)).
struct shape {
  int64_t d[kMaxDims];
};

// Extent of a table or an expression.  A scalar matches any extent; the
// bits of `expand` mark the dimensions of extent 1 that sx() lets grow to
// the extent of the other operand (singleton expansion).
struct extent {
  int64_t d[kMaxDims] = {0, 0, 1, 1};
  bool scalar = false;
  unsigned expand = 0;
  extent() = default;
  extent(int64_t d0, int64_t d1, int64_t d2 = 1, int64_t d3 = 1) : d{d0, d1, d2, d3} {}
  static extent of_scalar() {
    extent e(1, 1);
    e.scalar = true;
    return e;
  }
  int64_t rows() const { return d[0]; }
  int64_t cols() const { return d[1]; }
  int64_t numel() const { return d[0] * d[1] * d[2] * d[3]; }
  shape sh() const { return shape{{d[0], d[1], d[2], d[3]}}; }
  bool same(const extent& o) const {
    for (int k = 0; k < kMaxDims; ++k)
      if (d[k] != o.d[k]) return false;
    return true;
  }
  // the extent of a result table: no scalar or expansion marks
  extent plain() const { return extent(d[0], d[1], d[2], d[3]); }
};

template <class T> class table;

// ============================================================ nodes
// K consecutive elements of a loop nest (the vector bodies of the kernels).
template <class T, int K = 4> struct pack {
  T v[K];
};
template <class T> using quad = pack<T, 4>;
// What fetch() of a node returns: for a table its K elements as loaded,
// for an inner node the fetches of its children, for a scalar nothing.
struct none {};
template <class A, class B> struct pair_of {
  A a;
  B b;
};
template <class A, class B, class C> struct triple_of {
  A a;
  B b;
  C c;
};

namespace detail {
// a / b for 0 <= a and 0 < b: 32-bit when both fit (the usual case), else
// an out-of-line 64-bit division (the GPU has no divide instruction, and
// the inline 64-bit sequence inside the kernels' unrolled loops would
// multiply their code size)
__device__ __noinline__ inline int64_t udiv64(int64_t a, int64_t b) { return a / b; }
NT2H_FD int64_t udiv(int64_t a, int64_t b) {
  if (((uint64_t)a | (uint64_t)b) >> 32 == 0) return (int64_t)((uint32_t)a / (uint32_t)b);
  return udiv64(a, b);
}
}  // namespace detail

namespace node {
// Every node evaluates at(i, sh): element i of a loop nest of extent sh,
// and elements i..i+K-1 for i a multiple of K (K * sizeof(T) a multiple of
// 16) in two steps: fetch<P, K>(i, sh) issues the 16-byte loads of every
// table of the tree, eval<P, K>(fetched) does the arithmetic — so that a
// kernel can have the loads of several vectors, or of several trees, in
// flight before the first arithmetic waits on memory.  aligned() says
// whether every table of the tree starts on a 16-byte boundary (the
// condition for fetch); reads(p) whether the tree reads table p,
// expanded_reads(p) whether it reads p at other elements than the one it
// evaluates (through sx()).
template <class T> struct term {  // a table read at the element index
  const T* p;
  template <class P> NT2H_FD T at(int64_t i, const shape&) const { return p[i]; }
  template <class P, int K> NT2H_FD pack<T, K> fetch(int64_t i, const shape&) const {
    using V = typename nt2h::Vec16<T>::type;
    constexpr int W = nt2h::Vec16<T>::n;
    static_assert(K % W == 0, "whole 16-byte vectors");
    pack<T, K> q;
#pragma unroll
    for (int h = 0; h < K / W; ++h) {
      V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p + i) + h);
#pragma unroll
      for (int k = 0; k < W; ++k) q.v[h * W + k] = v[k];
    }
    return q;
  }
  template <class P, int K> NT2H_FD pack<T, K> eval(const pack<T, K>& f) const { return f; }
  bool aligned() const { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
  bool reads(const void* q) const { return q == static_cast<const void*>(p); }
  bool expanded_reads(const void*) const { return false; }
};
// Singleton expansion: the child, of extent own, evaluated at the element
// that element i of the loop nest maps to — index 0 along every dimension
// where own has extent 1.  When only the first dimension is expanded (a row
// repeated down the columns) that element is i / d0.  Its elements are not
// contiguous, so fetch() evaluates them (their loads mostly hit the cache).
template <class N> struct expand {
  N n;
  shape own;
  NT2H_FD bool row_only(const shape& sh) const {
    return own.d[0] == 1 && own.d[1] == sh.d[1] && own.d[2] == sh.d[2] && own.d[3] == sh.d[3];
  }
  // the coordinates of linear index i in a loop nest of extent sh
  NT2H_FD static void coords(int64_t i, const shape& sh, int64_t* c) {
    const int64_t r = detail::udiv(i, sh.d[0]);
    c[0] = i - r * sh.d[0];
    const int64_t r2 = detail::udiv(r, sh.d[1]);
    c[1] = r - r2 * sh.d[1];
    c[3] = detail::udiv(r2, sh.d[2]);
    c[2] = r2 - c[3] * sh.d[2];
  }
  NT2H_FD int64_t own_index(const int64_t* c) const {
    return (own.d[0] == 1 ? 0 : c[0]) +
           own.d[0] * ((own.d[1] == 1 ? 0 : c[1]) +
                       own.d[1] * ((own.d[2] == 1 ? 0 : c[2]) + own.d[2] * (own.d[3] == 1 ? 0 : c[3])));
  }
  template <class P> NT2H_FD auto at(int64_t i, const shape& sh) const {
    if (row_only(sh)) return n.template at<P>(detail::udiv(i, sh.d[0]), own);
    int64_t c[4];
    coords(i, sh, c);
    return n.template at<P>(own_index(c), own);
  }
  template <class P, int K> NT2H_FD auto fetch(int64_t i, const shape& sh) const {
    using R = decltype(n.template at<P>(0, own));
    pack<R, K> q;
    if (row_only(sh)) {
      const int64_t c = detail::udiv(i, sh.d[0]);
      if (i - c * sh.d[0] + (K - 1) < sh.d[0]) {  // the K elements share a column
        const R v = n.template at<P>(c, own);
#pragma unroll
        for (int k = 0; k < K; ++k) q.v[k] = v;
        return q;
      }
    }
    // one decomposition, then step the coordinates with carries
    int64_t c[4];
    coords(i, sh, c);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      q.v[k] = n.template at<P>(own_index(c), own);
      if (k < K - 1 && ++c[0] == sh.d[0]) {
        c[0] = 0;
        if (++c[1] == sh.d[1]) {
          c[1] = 0;
          if (++c[2] == sh.d[2]) {
            c[2] = 0;
            ++c[3];
          }
        }
      }
    }
    return q;
  }
  template <class P, int K, class R> NT2H_FD pack<R, K> eval(const pack<R, K>& f) const { return f; }
  bool aligned() const { return true; }
  bool reads(const void* q) const { return n.reads(q); }
  bool expanded_reads(const void* q) const { return n.reads(q); }
};
template <class T> struct scalar {
  T v;
  template <class P> NT2H_FD T at(int64_t, const shape&) const { return v; }
  template <class P, int K> NT2H_FD none fetch(int64_t, const shape&) const { return {}; }
  template <class P, int K> NT2H_FD pack<T, K> eval(const none&) const {
    pack<T, K> q;
#pragma unroll
    for (int k = 0; k < K; ++k) q.v[k] = v;
    return q;
  }
  bool aligned() const { return true; }
  bool reads(const void*) const { return false; }
  bool expanded_reads(const void*) const { return false; }
};
template <class F, class A> struct unary {
  A a;
  template <class P> NT2H_FD auto at(int64_t i, const shape& r) const {
    return F::template apply<P>(a.template at<P>(i, r));
  }
  template <class P, int K> NT2H_FD auto fetch(int64_t i, const shape& r) const {
    return a.template fetch<P, K>(i, r);
  }
  template <class P, int K, class FA> NT2H_FD auto eval(const FA& f) const {
    const auto x = a.template eval<P, K>(f);
    pack<decltype(F::template apply<P>(x.v[0])), K> q;
#pragma unroll
    for (int k = 0; k < K; ++k) q.v[k] = F::template apply<P>(x.v[k]);
    return q;
  }
  bool aligned() const { return a.aligned(); }
  bool reads(const void* p) const { return a.reads(p); }
  bool expanded_reads(const void* p) const { return a.expanded_reads(p); }
};
template <class F, class A, class B> struct binary {
  A a;
  B b;
  template <class P> NT2H_FD auto at(int64_t i, const shape& r) const {
    return F::template apply<P>(a.template at<P>(i, r), b.template at<P>(i, r));
  }
  template <class P, int K> NT2H_FD auto fetch(int64_t i, const shape& r) const {
    using FA = decltype(a.template fetch<P, K>(i, r));
    using FB = decltype(b.template fetch<P, K>(i, r));
    return pair_of<FA, FB>{a.template fetch<P, K>(i, r), b.template fetch<P, K>(i, r)};
  }
  template <class P, int K, class FF> NT2H_FD auto eval(const FF& f) const {
    const auto x = a.template eval<P, K>(f.a);
    const auto y = b.template eval<P, K>(f.b);
    pack<decltype(F::template apply<P>(x.v[0], y.v[0])), K> q;
#pragma unroll
    for (int k = 0; k < K; ++k) q.v[k] = F::template apply<P>(x.v[k], y.v[k]);
    return q;
  }
  bool aligned() const { return a.aligned() && b.aligned(); }
  bool reads(const void* p) const { return a.reads(p) || b.reads(p); }
  bool expanded_reads(const void* p) const { return a.expanded_reads(p) || b.expanded_reads(p); }
};
template <class F, class A, class B, class C> struct ternary {
  A a;
  B b;
  C c;
  template <class P> NT2H_FD auto at(int64_t i, const shape& r) const {
    return F::template apply<P>(a.template at<P>(i, r), b.template at<P>(i, r), c.template at<P>(i, r));
  }
  template <class P, int K> NT2H_FD auto fetch(int64_t i, const shape& r) const {
    using FA = decltype(a.template fetch<P, K>(i, r));
    using FB = decltype(b.template fetch<P, K>(i, r));
    using FC = decltype(c.template fetch<P, K>(i, r));
    return triple_of<FA, FB, FC>{a.template fetch<P, K>(i, r), b.template fetch<P, K>(i, r),
                                 c.template fetch<P, K>(i, r)};
  }
  template <class P, int K, class FF> NT2H_FD auto eval(const FF& f) const {
    const auto x = a.template eval<P, K>(f.a);
    const auto y = b.template eval<P, K>(f.b);
    const auto z = c.template eval<P, K>(f.c);
    pack<decltype(F::template apply<P>(x.v[0], y.v[0], z.v[0])), K> q;
#pragma unroll
    for (int k = 0; k < K; ++k) q.v[k] = F::template apply<P>(x.v[k], y.v[k], z.v[k]);
    return q;
  }
  bool aligned() const { return a.aligned() && b.aligned() && c.aligned(); }
  bool reads(const void* p) const { return a.reads(p) || b.reads(p) || c.reads(p); }
  bool expanded_reads(const void* p) const {
    return a.expanded_reads(p) || b.expanded_reads(p) || c.expanded_reads(p);
  }
};
}  // namespace node

// ========================================================== functors
namespace fn {
#define NT2HIP_UNARY(NAME, F32, F64, OPID)                                                   \
  struct NAME {                                                                              \
    static constexpr int op = OPID;                                                          \
    template <class P> NT2H_FD static float apply(float x) { return nt2h::F32<P>(x); }       \
    template <class P> NT2H_FD static double apply(double x) { return nt2h::F64<P>(x); }     \
  };
NT2HIP_UNARY(exp_, exp_f, exp_d, NT2H_OP_EXP)
NT2HIP_UNARY(log_, log_f, log_d, NT2H_OP_LOG)
NT2HIP_UNARY(erf_, erf_f, erf_d, NT2H_OP_ERF)
NT2HIP_UNARY(erfc_, erfc_f, erfc_d, NT2H_OP_ERFC)
NT2HIP_UNARY(normcdf_, normcdf_f, normcdf_d, NT2H_OP_NORMCDF)
NT2HIP_UNARY(normpdf_, normpdf_f, normpdf_d, NT2H_OP_NORMPDF)
NT2HIP_UNARY(tanh_, tanh_f, tanh_d, NT2H_OP_TANH)
NT2HIP_UNARY(sin_, sin_f, sin_d, NT2H_OP_SIN)
NT2HIP_UNARY(cos_, cos_f, cos_d, NT2H_OP_COS)
#undef NT2HIP_UNARY
struct sqrt_ {
  static constexpr int op = NT2H_OP_SQRT;
  template <class P, class T> NT2H_FD static T apply(T x) { return P::sqrt(x); }
};
struct fastnormcdf_ {  // float only (statistics/.../fastnormcdf.hpp)
  static constexpr int op = NT2H_OP_FASTNORMCDF;
  template <class P> NT2H_FD static float apply(float x) { return nt2h::fastnormcdf_f<P>(x); }
};
struct exp2_ {
  static constexpr int op = NT2H_OP_EXP2;
  template <class P> NT2H_FD static float apply(float x) { return nt2h::exp2_f<P>(x); }
};
struct exp10_ {
  static constexpr int op = NT2H_OP_EXP10;
  template <class P> NT2H_FD static float apply(float x) { return nt2h::exp10_f<P>(x); }
};
struct sqr_ {
  static constexpr int op = -1;
  template <class P, class T> NT2H_FD static T apply(T x) { return x * x; }
};
struct abs_ {
  static constexpr int op = -1;
  template <class P, class T> NT2H_FD static T apply(T x) { return x < T(0) ? -x : (x == T(0) ? T(0) : x); }
};
struct neg_ {
  static constexpr int op = -1;
  template <class P, class T> NT2H_FD static T apply(T x) { return -x; }
};
struct plus_ {
  template <class P, class T> NT2H_FD static T apply(T a, T b) { return a + b; }
};
struct minus_ {
  template <class P, class T> NT2H_FD static T apply(T a, T b) { return a - b; }
};
struct multiplies_ {
  template <class P, class T> NT2H_FD static T apply(T a, T b) { return a * b; }
};
struct divides_ {
  template <class P, class T> NT2H_FD static T apply(T a, T b) { return P::div(a, b); }
};
// The multiply-adds of Boost.SIMD (arithmetic/functions/generic/{fma,fms,
// fnma,fnms}.hpp:32): NT2's scalar path forms the product and rounds twice;
// the fused policy issues one v_fma.
struct fma_ {  // a*b + c
  template <class P, class T> NT2H_FD static T apply(T a, T b, T c) { return P::madd(a, b, c); }
};
struct fms_ {  // a*b - c
  template <class P, class T> NT2H_FD static T apply(T a, T b, T c) { return P::madd(a, b, -c); }
};
struct fnms_ {  // c - a*b
  template <class P, class T> NT2H_FD static T apply(T a, T b, T c) { return P::nms(a, b, c); }
};
struct fnma_ {  // -c - a*b
  template <class P, class T> NT2H_FD static T apply(T a, T b, T c) { return P::nms(a, b, -c); }
};
}  // namespace fn

// Functor tags of sx(tag, a, b), as NT2's tag::plus_, tag::multiplies_, ...
namespace tag {
using plus_ = fn::plus_;
using minus_ = fn::minus_;
using multiplies_ = fn::multiplies_;
using divides_ = fn::divides_;
}  // namespace tag

// ==================================================== expression wrapper
template <class T, class N> struct expr {
  N n;
  extent ext;
};

template <class T> struct is_table : std::false_type {};
template <class T> struct is_table<table<T>> : std::true_type {};
template <class X> struct is_expr : std::false_type {};
template <class T, class N> struct is_expr<expr<T, N>> : std::true_type {};
template <class X> struct is_operand : std::integral_constant<bool, is_table<X>::value || is_expr<X>::value> {};

template <class T> inline expr<T, node::term<T>> as_expr(const table<T>& t) {
  return {node::term<T>{t.data()}, t.extent_()};
}
template <class T, class N> inline const expr<T, N>& as_expr(const expr<T, N>& e) { return e; }

// Extent of an elementwise node: equal extents, a scalar against anything,
// or along each dimension an sx()-marked extent 1 against any extent.
inline extent merge(const extent& a, const extent& b) {
  if (a.scalar) return b;
  if (b.scalar) return a;
  extent r;
  for (int k = 0; k < kMaxDims; ++k) {
    const unsigned bit = 1u << k;
    if (a.d[k] == b.d[k]) {
      r.d[k] = a.d[k];
      r.expand |= a.expand & b.expand & bit;
    } else if (a.d[k] == 1 && (a.expand & bit)) {
      r.d[k] = b.d[k];
      r.expand |= b.expand & bit;
    } else if (b.d[k] == 1 && (b.expand & bit)) {
      r.d[k] = a.d[k];
      r.expand |= a.expand & bit;
    } else {
      throw error("nt2hip: size mismatch in a table expression", NT2H_ERR_INVALID_ARG);
    }
  }
  return r;
}

// value type of a table or expression
template <class X> struct value_of;
template <class T> struct value_of<table<T>> { using type = T; };
template <class T, class N> struct value_of<expr<T, N>> { using type = T; };

// ----------------------------------------------------------- unary builders
#define NT2HIP_UFUN(NAME, F)                                                                  \
  template <class T> inline auto NAME(const table<T>& a) { return NAME(as_expr(a)); }       \
  template <class T, class N> inline expr<T, node::unary<fn::F, N>> NAME(const expr<T, N>& a) { \
    return {node::unary<fn::F, N>{a.n}, a.ext};                                               \
  }
NT2HIP_UFUN(exp, exp_)
NT2HIP_UFUN(log, log_)
NT2HIP_UFUN(sqrt, sqrt_)
NT2HIP_UFUN(erf, erf_)
NT2HIP_UFUN(erfc, erfc_)
NT2HIP_UFUN(normcdf, normcdf_)
NT2HIP_UFUN(normpdf, normpdf_)
NT2HIP_UFUN(fastnormcdf, fastnormcdf_)
NT2HIP_UFUN(exp2, exp2_)
NT2HIP_UFUN(exp10, exp10_)
NT2HIP_UFUN(tanh, tanh_)
NT2HIP_UFUN(sin, sin_)
NT2HIP_UFUN(cos, cos_)
NT2HIP_UFUN(sqr, sqr_)
NT2HIP_UFUN(abs, abs_)
#undef NT2HIP_UFUN
template <class T> inline auto operator-(const table<T>& a) { return -as_expr(a); }
template <class T, class N> inline expr<T, node::unary<fn::neg_, N>> operator-(const expr<T, N>& a) {
  return {node::unary<fn::neg_, N>{a.n}, a.ext};
}

// ---------------------------------------------------------- binary builders
template <class F, class TA, class NA, class NB>
inline expr<TA, node::binary<F, NA, NB>> make_bin(const expr<TA, NA>& a, const expr<TA, NB>& b) {
  return {node::binary<F, NA, NB>{a.n, b.n}, merge(a.ext, b.ext)};
}
template <class T, class S> inline expr<T, node::scalar<T>> scal(S v) {
  return {node::scalar<T>{T(v)}, extent::of_scalar()};
}

#define NT2HIP_BOP(OP, F)                                                                           \
  template <class A, class B, class = std::enable_if_t<is_operand<A>::value && is_operand<B>::value>> \
  inline auto operator OP(const A& a, const B& b) {                                                 \
    return make_bin<fn::F>(as_expr(a), as_expr(b));                                                 \
  }                                                                                                 \
  template <class A, class S,                                                                       \
            class = std::enable_if_t<is_operand<A>::value && std::is_arithmetic<S>::value>>         \
  inline auto operator OP(const A& a, S s) {                                                        \
    using T = typename value_of<A>::type;                                                           \
    return make_bin<fn::F>(as_expr(a), scal<T>(s));                                                 \
  }                                                                                                 \
  template <class S, class B,                                                                       \
            class = std::enable_if_t<std::is_arithmetic<S>::value && is_operand<B>::value>,         \
            class = void>                                                                           \
  inline auto operator OP(S s, const B& b) {                                                        \
    using T = typename value_of<B>::type;                                                           \
    return make_bin<fn::F>(scal<T>(s), as_expr(b));                                                 \
  }
NT2HIP_BOP(+, plus_)
NT2HIP_BOP(-, minus_)
NT2HIP_BOP(*, multiplies_)
NT2HIP_BOP(/, divides_)
#undef NT2HIP_BOP

// ------------------------------------------------------- ternary builders
// fma(a, b, c), fms, fnms, fnma with any mix of tables, expressions and
// scalars (NT2's demo writes fma(sqr(v), T(0.5), r) and fnms(v, da, d1)).
template <class... Xs> struct first_value;
template <class X, class... R> struct first_value<X, R...> {
  using type = typename std::conditional<is_operand<X>::value, value_of<X>, first_value<R...>>::type::type;
};
template <class T, class X> inline auto operand(const X& x) {
  if constexpr (is_operand<X>::value) {
    static_assert(std::is_same<typename value_of<X>::type, T>::value, "operands of one value type");
    return as_expr(x);
  } else {
    return scal<T>(x);
  }
}
template <class F, class A, class B, class C> inline auto make_ter(const A& a, const B& b, const C& c) {
  using T = typename first_value<A, B, C>::type;
  const auto ea = operand<T>(a);
  const auto eb = operand<T>(b);
  const auto ec = operand<T>(c);
  using N = node::ternary<F, decltype(ea.n), decltype(eb.n), decltype(ec.n)>;
  return expr<T, N>{N{ea.n, eb.n, ec.n}, merge(merge(ea.ext, eb.ext), ec.ext)};
}
#define NT2HIP_TFUN(NAME, F)                                                                              \
  template <class A, class B, class C,                                                                    \
            class = std::enable_if_t<is_operand<A>::value || is_operand<B>::value || is_operand<C>::value>> \
  inline auto NAME(const A& a, const B& b, const C& c) {                                                  \
    return make_ter<fn::F>(a, b, c);                                                                      \
  }
NT2HIP_TFUN(fma, fma_)
NT2HIP_TFUN(fms, fms_)
NT2HIP_TFUN(fnms, fnms_)
NT2HIP_TFUN(fnma, fnma_)
#undef NT2HIP_TFUN

// ---------------------------------------------------- singleton expansion
// sx(x): x (a table or an expression) with every dimension of extent 1
// free to grow to the extent of the operand it meets — a 1 x n row repeated
// down the rows, an m x 1 column across the columns, a 1 x 1 x p pillar
// across both, in up to four dimensions.  sx(tag, a, b) is NT2's form: the
// binary functor of a and b with both expanded (core/functions/sx.hpp).
template <class T, class N> inline expr<T, node::expand<N>> sx(const expr<T, N>& e) {
  if (e.ext.scalar) throw error("nt2hip::sx: expects a table or an expression", NT2H_ERR_INVALID_ARG);
  expr<T, node::expand<N>> r{node::expand<N>{e.n, e.ext.sh()}, e.ext.plain()};
  for (int k = 0; k < kMaxDims; ++k)
    if (e.ext.d[k] == 1) r.ext.expand |= 1u << k;
  return r;
}
template <class T> inline auto sx(const table<T>& t) { return sx(as_expr(t)); }
template <class F, class A, class B, class = std::enable_if_t<is_operand<A>::value && is_operand<B>::value>>
inline auto sx(F, const A& a, const B& b) {
  return make_bin<F>(sx(as_expr(a)), sx(as_expr(b)));
}

// ============================================================ kernels
namespace detail {
template <class P, class T, class N>
__global__ __launch_bounds__(256) void k_transform(N n, T* out, int64_t numel, shape sh) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < numel; i += stride) out[i] = n.template at<P>(i, sh);
}
// Vector body: lane l evaluates elements 4q..4q+3 of quad q (16-byte loads
// and stores, one quad per lane and step); the last numel % 4 elements go
// through at().
template <class P, class T, class N>
__global__ __launch_bounds__(256) void k_transform4(N n, T* out, int64_t numel, shape sh) {
  using V = typename nt2h::Vec16<T>::type;
  constexpr int W = nt2h::Vec16<T>::n;
  const int64_t nq = numel / 4;
  const int64_t stride = (int64_t)gridDim.x * 256;
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (int64_t q = g; q < nq; q += stride) {
    const quad<T> r = n.template eval<P, 4>(n.template fetch<P, 4>(4 * q, sh));
#pragma unroll
    for (int h = 0; h < 4 / W; ++h) {
      V v;
#pragma unroll
      for (int k = 0; k < W; ++k) v[k] = r.v[h * W + k];
      __builtin_nontemporal_store(v, reinterpret_cast<V*>(out + 4 * q) + h);
    }
  }
  const int64_t t = 4 * nq + g;
  if (t < numel) out[t] = n.template at<P>(t, sh);
}
// tie's loop nests: at every element (every quad in the vector body) all
// children are evaluated before any output is stored, so an output may be
// an input of another child (core/functions/expr/tie.hpp:80-115 reads the
// whole fusion sequence, then stores it).
template <class... Ns> struct nodes;
template <> struct nodes<> {
  bool aligned() const { return true; }
  bool expanded_reads(const void*) const { return false; }
};
template <class H, class... R> struct nodes<H, R...> {
  H h;
  nodes<R...> t;
  bool aligned() const { return h.aligned() && t.aligned(); }
  bool expanded_reads(const void* p) const { return h.expanded_reads(p) || t.expanded_reads(p); }
};
template <class T, int N> struct out_ptrs {
  T* p[N];
};
template <class P, class T> NT2H_FD void eval_all(const nodes<>&, int64_t, const shape&, T*) {}
template <class P, class T, class H, class... R>
NT2H_FD void eval_all(const nodes<H, R...>& ns, int64_t i, const shape& sh, T* v) {
  v[0] = ns.h.template at<P>(i, sh);
  eval_all<P, T>(ns.t, i, sh, v + 1);
}
// the vector body: every child's loads first, then every child's arithmetic
template <class P> NT2H_FD none fetch_all4(const nodes<>&, int64_t, const shape&) { return {}; }
template <class P, class H, class... R> NT2H_FD auto fetch_all4(const nodes<H, R...>& ns, int64_t i, const shape& sh) {
  using FH = decltype(ns.h.template fetch<P, 4>(i, sh));
  using FT = decltype(fetch_all4<P>(ns.t, i, sh));
  return pair_of<FH, FT>{ns.h.template fetch<P, 4>(i, sh), fetch_all4<P>(ns.t, i, sh)};
}
template <class P, class T> NT2H_FD void eval_all4(const nodes<>&, const none&, quad<T>*) {}
template <class P, class T, class H, class... R, class FF>
NT2H_FD void eval_all4(const nodes<H, R...>& ns, const FF& f, quad<T>* q) {
  q[0] = ns.h.template eval<P, 4>(f.a);
  eval_all4<P, T>(ns.t, f.b, q + 1);
}
template <class P, class T, int N, class Ns>
__global__ __launch_bounds__(256) void k_tie4(Ns ns, out_ptrs<T, N> o, int64_t numel, shape sh) {
  using V = typename nt2h::Vec16<T>::type;
  constexpr int W = nt2h::Vec16<T>::n;
  const int64_t nq = numel / 4;
  const int64_t stride = (int64_t)gridDim.x * 256;
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (int64_t q = g; q < nq; q += stride) {
    quad<T> v[N];
    eval_all4<P, T>(ns, fetch_all4<P>(ns, 4 * q, sh), v);
#pragma unroll
    for (int j = 0; j < N; ++j) {
#pragma unroll
      for (int h = 0; h < 4 / W; ++h) {
        V x;
#pragma unroll
        for (int k = 0; k < W; ++k) x[k] = v[j].v[h * W + k];
        __builtin_nontemporal_store(x, reinterpret_cast<V*>(o.p[j] + 4 * q) + h);
      }
    }
  }
  const int64_t t = 4 * nq + g;
  if (t < numel) {
    T v[N];
    eval_all<P, T>(ns, t, sh, v);
#pragma unroll
    for (int j = 0; j < N; ++j) o.p[j][t] = v[j];
  }
}
template <class P, class T, int N, class Ns>
__global__ __launch_bounds__(256) void k_tie(Ns ns, out_ptrs<T, N> o, int64_t numel, shape sh) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < numel; i += stride) {
    T v[N];
    eval_all<P, T>(ns, i, sh, v);
#pragma unroll
    for (int j = 0; j < N; ++j) o.p[j][i] = v[j];
  }
}

inline unsigned grid_for(int64_t numel) {
  int64_t b = (numel + 255) / 256;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (unsigned)b;
}
inline unsigned quad_grid(int64_t numel) {
  int64_t b = (numel / 4 + 255) / 256;
  if (b < 1) b = 1;
  if (b > (int64_t(1) << 30)) b = int64_t(1) << 30;
  return (unsigned)b;
}

// An expression tree as the element source of the library's fold and scan
// kernels: element i of the loop nest of extent sh, evaluated in registers
// where the kernel reads it; raw(i) is the fetch of the tree's tables for
// the 16-byte vector of elements i..i+W-1, value() its arithmetic.
template <class P, class T, class N> struct ExprSrc {
  static constexpr int W = nt2h::Vec16<T>::n;
  using Raw = decltype(std::declval<const N&>().template fetch<P, W>(0, std::declval<const shape&>()));
  N n;
  shape sh;
  NT2H_FD T at(int64_t i) const { return n.template at<P>(i, sh); }
  NT2H_FD Raw raw(int64_t i) const { return n.template fetch<P, W>(i, sh); }
  NT2H_FD typename nt2h::Vec16<T>::type value(const Raw& r) const {
    const pack<T, W> q = n.template eval<P, W>(r);
    typename nt2h::Vec16<T>::type v;
#pragma unroll
    for (int k = 0; k < W; ++k) v[k] = q.v[k];
    return v;
  }
  NT2H_FD typename nt2h::Vec16<T>::type vec(int64_t i) const { return value(raw(i)); }
  bool aligned() const { return n.aligned(); }
};
// The launch environment of those kernels: the scratch and look-back state
// the library keeps for the current stream.
struct FrontEnv {
  int scratch(size_t bytes, void** p) { return nt2h_stream_buffer(abi_stream(), 0, bytes, p); }
  int lookback(size_t bytes, void** p) { return nt2h_stream_buffer(abi_stream(), 1, bytes, p); }
  void lookback_wait(unsigned* cap, unsigned* mask) {
    *cap = 0;  // the kernels' own poll limit
    *mask = 0;
  }
  int hip(hipError_t e) { return e == hipSuccess ? NT2H_OK : NT2H_ERR_HIP; }
};
}  // namespace detail

// ============================================================= table
template <class T> class table {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "table<float> or table<double>");

 public:
  table() = default;
  explicit table(int64_t n) { resize(extent(n, 1)); }
  table(int64_t d0, int64_t d1, int64_t d2 = 1, int64_t d3 = 1) { resize(extent(d0, d1, d2, d3)); }
  explicit table(const extent& e) { resize(e); }
  // host data, column-major
  table(const std::vector<T>& host, int64_t rows, int64_t cols) : table(rows, cols) { upload(host); }
  table(const std::vector<T>& host, const extent& e) : table(e) { upload(host); }
  explicit table(const std::vector<T>& host) : table(host, (int64_t)host.size(), 1) {}
  table(const table& o) : table(o.ext_) { copy_from(o); }
  table(table&& o) noexcept : p_(o.p_), ext_(o.ext_), cap_(o.cap_) { o.p_ = nullptr; o.cap_ = 0; o.ext_ = {}; }
  template <class N> table(const expr<T, N>& e) { *this = e; }
  ~table() { nt2h_free(p_); }

  table& operator=(const table& o) {
    if (this != &o) { resize(o.ext_); copy_from(o); }
    return *this;
  }
  table& operator=(table&& o) noexcept {
    std::swap(p_, o.p_); std::swap(ext_, o.ext_); std::swap(cap_, o.cap_);
    return *this;
  }
  template <class N> table& operator=(const expr<T, N>& e);
  table& operator=(T v) { return *this = expr<T, node::scalar<T>>{node::scalar<T>{v}, ext_}; }

  T* data() { return p_; }
  const T* data() const { return p_; }
  int64_t numel() const { return ext_.numel(); }
  int64_t rows() const { return ext_.d[0]; }
  int64_t cols() const { return ext_.d[1]; }
  // size(x, k): 1-based, 1 beyond the fourth dimension
  int64_t dim(int k) const { return (k >= 1 && k <= kMaxDims) ? ext_.d[k - 1] : 1; }
  extent extent_() const { return ext_; }

  void resize(extent e) {
    e = e.plain();
    for (int k = 0; k < kMaxDims; ++k)
      if (e.d[k] < 0) throw error("nt2hip::table resize", NT2H_ERR_INVALID_ARG);
    if (e.numel() > cap_) {
      nt2h_free(p_);
      p_ = nullptr;
      check(nt2h_malloc((void**)&p_, sizeof(T) * e.numel()), "nt2h_malloc");
      cap_ = e.numel();
    }
    ext_ = e;
  }
  // host <-> device through the library's pinned staging (synchronous)
  void upload(const std::vector<T>& host) {
    if ((int64_t)host.size() != numel()) throw error("nt2hip::table upload size", NT2H_ERR_INVALID_ARG);
    check(nt2h_upload(p_, host.data(), sizeof(T) * host.size(), abi_stream()), "nt2h_upload");
  }
  std::vector<T> to_host() const {
    std::vector<T> h(numel());
    check(nt2h_sync(abi_stream()), "nt2h_sync");  // work queued on the table's stream first
    check(nt2h_download(h.data(), p_, sizeof(T) * h.size(), abi_stream()), "nt2h_download");
    return h;
  }

 private:
  void copy_from(const table& o) {
    check(nt2h_memcpy(p_, o.p_, sizeof(T) * numel(), NT2H_D2D, abi_stream()), "nt2h_memcpy");
  }
  T* p_ = nullptr;
  extent ext_{};
  int64_t cap_ = 0;
};

// ========================================================== lowering
namespace detail {
template <class P> constexpr int precision_id() { return P::fused ? NT2H_FUSED : NT2H_EXACT; }

// generic: one kernel for the tree — the vector body when the output and
// every table of the tree are 16-byte aligned (one quad per lane: the grid
// covers the table), else the scalar body
template <class P, class T, class N> inline void run_generic(T* out, const N& n, const extent& x) {
  if (x.numel() == 0) return;
  if (n.aligned() && (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
    hipLaunchKernelGGL((k_transform4<P, T, N>), dim3(quad_grid(x.numel())), dim3(256), 0, current_stream(), n, out,
                       x.numel(), x.sh());
  } else {
    hipLaunchKernelGGL((k_transform<P, T, N>), dim3(grid_for(x.numel())), dim3(256), 0, current_stream(), n, out,
                       x.numel(), x.sh());
  }
  hip_check(hipGetLastError(), "transform launch");
}
template <class P, class T, class N> struct lower {
  static void run(T* out, const N& n, const extent& x) { run_generic<P, T, N>(out, n, x); }
};
// Y = alpha*X + Y  ->  nt2h_saxpy_f32 (when the last operand is the target)
template <class P>
struct lower<P, float,
             node::binary<fn::plus_, node::binary<fn::multiplies_, node::scalar<float>, node::term<float>>,
                          node::term<float>>> {
  using N = node::binary<fn::plus_, node::binary<fn::multiplies_, node::scalar<float>, node::term<float>>,
                         node::term<float>>;
  static void run(float* out, const N& n, const extent& x) {
    if (!P::fused && n.b.p == out && n.a.b.p != out) {
      check(nt2h_saxpy_f32(n.a.a.v, n.a.b.p, out, x.numel(), abi_stream()), "nt2h_saxpy_f32");
      return;
    }
    run_generic<P, float, N>(out, n, x);
  }
};
// Y = f(X) for a functor with a map kernel  ->  nt2h_map_f32 / nt2h_map_f64
template <class P, class T, class F> struct lower<P, T, node::unary<F, node::term<T>>> {
  using N = node::unary<F, node::term<T>>;
  static void run(T* out, const N& n, const extent& x) {
    const int64_t numel = x.numel();
    constexpr bool f64_ok = F::op >= 0 && F::op <= NT2H_OP_NORMPDF;
    if (F::op >= 0 && (std::is_same<T, float>::value || f64_ok)) {
      int st = std::is_same<T, float>::value
                   ? nt2h_map_f32(F::op, precision_id<P>(), (const float*)n.a.p, (float*)out, numel, abi_stream())
                   : nt2h_map_f64(F::op, precision_id<P>(), (const double*)n.a.p, (double*)out, numel, abi_stream());
      check(st, "nt2h_map");
      return;
    }
    run_generic<P, T, N>(out, n, x);
  }
};
// The assignment's loop nest.  A tree that reads the target at other
// elements than the one it writes (through sx()) is evaluated into a fresh
// table, which then takes the target's place.
template <class P, class T, class N> inline void assign_to(table<T>& lhs, const N& n, const extent& e) {
  const extent x = e.plain();
  if (lhs.data() && n.expanded_reads(lhs.data())) {
    table<T> tmp(x);
    lower<P, T, N>::run(tmp.data(), n, x);
    lhs = std::move(tmp);
    return;
  }
  lhs.resize(x);
  lower<P, T, N>::run(lhs.data(), n, x);
}
}  // namespace detail

template <class T> template <class N> table<T>& table<T>::operator=(const expr<T, N>& e) {
  detail::assign_to<exact>(*this, e.n, e.ext);
  return *this;
}

// assign<fused>(lhs, expr): the same loop nest with the FMA policy
template <class P, class T, class N> inline table<T>& assign(table<T>& lhs, const expr<T, N>& e) {
  detail::assign_to<P>(lhs, e.n, e.ext);
  return lhs;
}

// ======================================== reductions and scans (any dim)
// NT2 folds or scans along dimension k by viewing the table as (lo, mid,
// hi): the product of the dimensions before k, dimension k, the product of
// those after (reduction/run.hpp:48-93); a k beyond the table's dimensions
// is a singleton.  Tables lower to nt2h_fold_f32/f64 and nt2h_scan_f32/f64;
// expressions to the same kernels with the tree as the element source (no
// temporary table); globalsum of a terminal or of a product of two
// terminals to nt2h_sum / nt2h_dot.
namespace detail {
inline void lmh(const extent& x, int dim, int64_t& lo, int64_t& mid, int64_t& hi) {
  if (dim < 1) throw error("nt2hip: dimensions are 1-based", NT2H_ERR_INVALID_ARG);
  lo = 1;
  mid = 1;
  hi = 1;
  if (dim > kMaxDims) {
    lo = x.numel();
    return;
  }
  for (int k = 0; k < dim - 1; ++k) lo *= x.d[k];
  mid = x.d[dim - 1];
  for (int k = dim; k < kMaxDims; ++k) hi *= x.d[k];
}
inline extent reduced(const extent& x, int dim) {
  extent r = x.plain();
  if (dim >= 1 && dim <= kMaxDims) r.d[dim - 1] = 1;
  return r;
}
// the first dimension whose extent is not 1 (NT2's default for sum(x) & co)
inline int first_nonsingleton(const extent& x) {
  for (int k = 0; k < kMaxDims; ++k)
    if (x.d[k] != 1) return k + 1;
  return 1;
}
inline int fold_abi(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* out) {
  return nt2h_fold_f32(op, x, lo, mid, hi, out, abi_stream());
}
inline int fold_abi(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* out) {
  return nt2h_fold_f64(op, x, lo, mid, hi, out, abi_stream());
}
inline int scan_abi(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* y) {
  return nt2h_scan_f32(op, x, lo, mid, hi, y, abi_stream());
}
inline int scan_abi(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* y) {
  return nt2h_scan_f64(op, x, lo, mid, hi, y, abi_stream());
}
// fold OP of an expression over (lo, mid, hi) into out (lo*hi values): the
// fold kernels of the library instantiated over the tree
template <int OP, class T, class N> inline void fold_expr(const N& n, const extent& x, int64_t lo, int64_t mid,
                                                          int64_t hi, T* out) {
  if (lo * hi == 0) return;
  FrontEnv env;
  if (mid == 0) {
    hipLaunchKernelGGL(nt2h::fold_fill_kernel<T>, dim3((unsigned)((lo * hi + 255) / 256)), dim3(256), 0,
                       current_stream(), out, lo * hi, nt2h::fold_neutral_result<T>(OP));
    hip_check(hipGetLastError(), "fold launch");
    return;
  }
  check(nt2h::fold_launch<OP, T>(ExprSrc<exact, T, N>{n, x.sh()}, lo, mid, hi, out, current_stream(), env),
        "fold of an expression");
}
template <int OP, class T> inline void fold_into(const table<T>& t, int dim, table<T>& out) {
  int64_t lo, mid, hi;
  lmh(t.extent_(), dim, lo, mid, hi);
  out.resize(reduced(t.extent_(), dim));
  check(fold_abi(OP, t.data(), lo, mid, hi, out.data()), "nt2h_fold");
}
template <int OP, class T, class N> inline void fold_into(const expr<T, N>& e, int dim, table<T>& out) {
  const extent x = e.ext.plain();
  int64_t lo, mid, hi;
  lmh(x, dim, lo, mid, hi);
  out.resize(reduced(x, dim));
  fold_expr<OP, T>(e.n, x, lo, mid, hi, out.data());
}
template <int OP, class T> inline void scan_into(const table<T>& t, int dim, table<T>& out) {
  int64_t lo, mid, hi;
  lmh(t.extent_(), dim, lo, mid, hi);
  out.resize(t.extent_());
  check(scan_abi(OP, t.data(), lo, mid, hi, out.data()), "nt2h_scan");
}
template <int OP, class T, class N> inline void scan_into(const expr<T, N>& e, int dim, table<T>& out) {
  const extent x = e.ext.plain();
  int64_t lo, mid, hi;
  lmh(x, dim, lo, mid, hi);
  out.resize(x);
  if (lo * mid * hi == 0) return;
  FrontEnv env;
  check(nt2h::scan_launch<OP, T>(ExprSrc<exact, T, N>{e.n, x.sh()}, out.data(), lo, mid, hi, current_stream(), env),
        "scan of an expression");
}
// fold OP of every element, as a (numel, 1) view folded along dimension 1
template <int OP, class T> inline T fold_all(const table<T>& t) {
  table<T> out(1, 1);
  check(fold_abi(OP, t.data(), 1, t.numel(), 1, out.data()), "nt2h_fold");
  return out.to_host()[0];
}
template <int OP, class T, class N> inline T fold_all(const expr<T, N>& e) {
  table<T> out(1, 1);
  const extent x = e.ext.plain();
  fold_expr<OP, T>(e.n, x, 1, x.numel(), 1, out.data());
  return out.to_host()[0];
}
template <class X> inline extent ext_of(const X& x) { return as_expr(x).ext.plain(); }

template <class T, class N> struct gsum {
  static T run(const expr<T, N>& e) { return fold_all<NT2H_FOLD_SUM>(e); }
};
template <class T> struct gsum<T, node::term<T>> {  // globalsum(X) -> nt2h_sum
  static T run(const expr<T, node::term<T>>& e) {
    double out = 0.0;
    const int64_t numel = e.ext.numel();
    int st = std::is_same<T, float>::value
                 ? nt2h_sum_f32((const float*)e.n.p, numel, NT2H_SUM_KAHAN, &out, abi_stream())
                 : nt2h_sum_f64((const double*)e.n.p, numel, &out, abi_stream());
    check(st, "nt2h_sum");
    return (T)out;
  }
};
template <class T> struct gsum<T, node::binary<fn::multiplies_, node::term<T>, node::term<T>>> {  // -> nt2h_dot
  using N = node::binary<fn::multiplies_, node::term<T>, node::term<T>>;
  static T run(const expr<T, N>& e) {
    double out = 0.0;
    const int64_t numel = e.ext.numel();
    int st = std::is_same<T, float>::value
                 ? nt2h_dot_f32((const float*)e.n.a.p, (const float*)e.n.b.p, numel, NT2H_SUM_KAHAN, &out,
                                abi_stream())
                 : nt2h_dot_f64((const double*)e.n.a.p, (const double*)e.n.b.p, numel, &out, abi_stream());
    check(st, "nt2h_dot");
    return (T)out;
  }
};
}  // namespace detail

// globalsum(e): sum of every element (core/functions/globalsum.hpp:45),
// compensated per lane and combined in FP64, rounded to T.
template <class T, class N> inline T globalsum(const expr<T, N>& e) { return detail::gsum<T, N>::run(e); }
template <class T> inline T globalsum(const table<T>& t) { return globalsum(as_expr(t)); }
// dot(a, b) = globalsum(a .* b) (reduction/functions/expr/dot.hpp:28)
template <class A, class B> inline auto dot(const A& a, const B& b) { return globalsum(as_expr(a) * as_expr(b)); }

// sum, prod, maximum, minimum, asum1, asum2, norm2, mean (core/reduction/
// include/nt2/core/functions/{sum,prod,...}.hpp) along dimension dim (the
// result has extent 1 there), or along the first non-singleton dimension
// when none is given, as NT2's; globalX folds every element.  cumsum,
// cumprod, cummax, cummin (core/swar) keep the extent.
#define NT2HIP_FOLD_DIM(NAME, OP)                                                                 \
  template <class X, class = std::enable_if_t<is_operand<X>::value>>                              \
  inline table<typename value_of<X>::type> NAME(const X& x, int dim) {                            \
    table<typename value_of<X>::type> out;                                                        \
    detail::fold_into<OP>(x, dim, out);                                                           \
    return out;                                                                                   \
  }                                                                                               \
  template <class X, class = std::enable_if_t<is_operand<X>::value>>                              \
  inline table<typename value_of<X>::type> NAME(const X& x) {                                     \
    return NAME(x, detail::first_nonsingleton(detail::ext_of(x)));                                \
  }
#define NT2HIP_FOLD_ALL(NAME, OP)                                                                 \
  template <class X, class = std::enable_if_t<is_operand<X>::value>>                              \
  inline typename value_of<X>::type NAME(const X& x) {                                            \
    return detail::fold_all<OP>(as_expr(x));                                                      \
  }
NT2HIP_FOLD_DIM(sum, NT2H_FOLD_SUM)
NT2HIP_FOLD_DIM(prod, NT2H_FOLD_PROD)
NT2HIP_FOLD_DIM(maximum, NT2H_FOLD_MAXIMUM)
NT2HIP_FOLD_DIM(minimum, NT2H_FOLD_MINIMUM)
NT2HIP_FOLD_DIM(asum1, NT2H_FOLD_ASUM1)
NT2HIP_FOLD_DIM(asum2, NT2H_FOLD_ASUM2)
NT2HIP_FOLD_DIM(norm2, NT2H_FOLD_NORM2)
NT2HIP_FOLD_DIM(mean, NT2H_FOLD_MEAN)
NT2HIP_FOLD_ALL(globalprod, NT2H_FOLD_PROD)
NT2HIP_FOLD_ALL(globalmax, NT2H_FOLD_MAXIMUM)
NT2HIP_FOLD_ALL(globalmin, NT2H_FOLD_MINIMUM)
NT2HIP_FOLD_ALL(globalasum1, NT2H_FOLD_ASUM1)
NT2HIP_FOLD_ALL(globalasum2, NT2H_FOLD_ASUM2)
NT2HIP_FOLD_ALL(globalnorm2, NT2H_FOLD_NORM2)
NT2HIP_FOLD_ALL(globalmean, NT2H_FOLD_MEAN)
#undef NT2HIP_FOLD_DIM
#undef NT2HIP_FOLD_ALL

#define NT2HIP_SCAN(NAME, OP)                                                                     \
  template <class X, class = std::enable_if_t<is_operand<X>::value>>                              \
  inline table<typename value_of<X>::type> NAME(const X& x, int dim) {                            \
    table<typename value_of<X>::type> out;                                                        \
    detail::scan_into<OP>(x, dim, out);                                                           \
    return out;                                                                                   \
  }                                                                                               \
  template <class X, class = std::enable_if_t<is_operand<X>::value>>                              \
  inline table<typename value_of<X>::type> NAME(const X& x) {                                     \
    return NAME(x, detail::first_nonsingleton(detail::ext_of(x)));                                \
  }
NT2HIP_SCAN(cumsum, NT2H_SCAN_CUMSUM)
NT2HIP_SCAN(cumprod, NT2H_SCAN_CUMPROD)
NT2HIP_SCAN(cummax, NT2H_SCAN_CUMMAX)
NT2HIP_SCAN(cummin, NT2H_SCAN_CUMMIN)
#undef NT2HIP_SCAN

// The same reductions and scans into an existing table, whose storage is
// reused when it is large enough (no allocation in a loop):
//   fold<NT2H_FOLD_SUM>(z, exp(X) * sx(w), 2);   scan<NT2H_SCAN_CUMMAX>(c, X, 1);
template <int OP, class X, class = std::enable_if_t<is_operand<X>::value>>
inline void fold(table<typename value_of<X>::type>& out, const X& x, int dim) {
  static_assert(OP >= NT2H_FOLD_SUM && OP <= NT2H_FOLD_MEAN, "an nt2h_fold_op");
  if (as_expr(x).n.reads(out.data()) && out.data()) {  // the result would overwrite an operand
    table<typename value_of<X>::type> tmp;
    detail::fold_into<OP>(x, dim, tmp);
    out = std::move(tmp);
    return;
  }
  detail::fold_into<OP>(x, dim, out);
}
template <int OP, class X, class = std::enable_if_t<is_operand<X>::value>>
inline void scan(table<typename value_of<X>::type>& out, const X& x, int dim) {
  static_assert(OP >= NT2H_SCAN_CUMSUM && OP <= NT2H_SCAN_CUMMIN, "an nt2h_scan_op");
  if (as_expr(x).n.reads(out.data()) && out.data()) {  // a scan reads ahead of the element it writes
    table<typename value_of<X>::type> tmp;
    detail::scan_into<OP>(x, dim, tmp);
    out = std::move(tmp);
    return;
  }
  detail::scan_into<OP>(x, dim, out);
}

// =============================================================== tie
// tie(o1, ..., oN) = tie(e1, ..., eN): N outputs from one loop nest
// (core/container/table/include/nt2/core/functions/expr/tie.hpp:23-160),
// every child read at an element before any output is written there.
template <class T, class... Ns> struct tie_exprs {
  detail::nodes<Ns...> ns;
  extent ext;
};
template <class T> struct bs_call_put {
  const table<T> *S, *K, *T_, *r, *sig;
  int precision;
};
namespace detail {
inline detail::nodes<> make_nodes() { return {}; }
template <class H, class... R> inline detail::nodes<H, R...> make_nodes(const H& h, const R&... r) {
  return {h, make_nodes(r...)};
}
inline extent merge_all(const extent& a) { return a; }
template <class... R> inline extent merge_all(const extent& a, const extent& b, const R&... r) {
  return merge_all(merge(a, b), r...);
}
}  // namespace detail

template <class T, int N> struct tie_t {
  table<T>* o[N];

  template <class... Ns> void operator=(const tie_exprs<T, Ns...>& es) {
    static_assert(sizeof...(Ns) == N, "tie: as many expressions as outputs");
    assign_all<exact>(es);
  }
  // tie(a, b) = tie(b, a): the right-hand tables as terminals
  tie_t& operator=(const tie_t& rhs) {
    assign_all<exact>(rhs.as_exprs(std::make_integer_sequence<int, N>()));
    return *this;
  }
  template <int... I> auto as_exprs(std::integer_sequence<int, I...>) const {
    return tie_exprs<T, decltype(as_expr(*o[I]).n)...>{detail::make_nodes(as_expr(*o[I]).n...),
                                                       detail::merge_all(o[I]->extent_()...)};
  }
  template <class P, class... Ns> void assign_all(const tie_exprs<T, Ns...>& es) {
    const extent x = es.ext.plain();
    // an output that some child reads through sx() (at other elements) is
    // written only after the whole loop nest: evaluate into fresh tables
    bool stage = false;
    for (int j = 0; j < N; ++j) stage = stage || es.ns.expanded_reads(o[j]->data());
    table<T> tmp[N];
    detail::out_ptrs<T, N> dst;
    for (int j = 0; j < N; ++j) {
      table<T>& t = stage ? tmp[j] : *o[j];
      t.resize(x);
      dst.p[j] = t.data();
    }
    const int64_t numel = x.numel();
    if (numel > 0) {
      bool al = es.ns.aligned();
      for (int j = 0; j < N; ++j) al = al && (reinterpret_cast<uintptr_t>(dst.p[j]) & 15u) == 0;
      using NS = detail::nodes<Ns...>;
      if (al) {
        hipLaunchKernelGGL((detail::k_tie4<P, T, N, NS>), dim3(detail::quad_grid(numel)), dim3(256), 0,
                           current_stream(), es.ns, dst, numel, x.sh());
      } else {
        hipLaunchKernelGGL((detail::k_tie<P, T, N, NS>), dim3(detail::grid_for(numel)), dim3(256), 0,
                           current_stream(), es.ns, dst, numel, x.sh());
      }
      hip_check(hipGetLastError(), "tie launch");
    }
    if (stage)
      for (int j = 0; j < N; ++j) *o[j] = std::move(tmp[j]);
  }

  // tie(call, put) = blackscholes(S, K, T, r, sig): nt2h_bs_fwd_f32/f64
  void operator=(const bs_call_put<T>& q) {
    static_assert(N == 2, "blackscholes yields two tables: tie(call, put)");
    const extent x = q.S->extent_();
    o[0]->resize(x);
    o[1]->resize(x);
    int st = std::is_same<T, float>::value
                 ? nt2h_bs_fwd_f32(q.precision, (const float*)q.S->data(), (const float*)q.K->data(),
                                   (const float*)q.T_->data(), (const float*)q.r->data(), (const float*)q.sig->data(),
                                   (float*)o[0]->data(), (float*)o[1]->data(), x.numel(), abi_stream())
                 : nt2h_bs_fwd_f64(q.precision, (const double*)q.S->data(), (const double*)q.K->data(),
                                   (const double*)q.T_->data(), (const double*)q.r->data(),
                                   (const double*)q.sig->data(), (double*)o[0]->data(), (double*)o[1]->data(),
                                   x.numel(), abi_stream());
    check(st, "nt2h_bs_fwd");
  }
};
// assign<fused>(tie(...), tie(...)): the same loop nest with the FMA policy
template <class P, class T, int N, class... Ns> inline void assign(tie_t<T, N>&& lhs, const tie_exprs<T, Ns...>& es) {
  static_assert(sizeof...(Ns) == N, "tie: as many expressions as outputs");
  lhs.template assign_all<P>(es);
}

// tie of writable tables: the left-hand side (or, on the right, their values)
template <class T, class... R,
          class = std::enable_if_t<(std::is_same<R, table<T>>::value && ...)>>
inline tie_t<T, 1 + (int)sizeof...(R)> tie(table<T>& a, R&... rest) {
  return {{&a, &rest...}};
}
// tie of expressions (tables, expressions and scalars mixed): the right-hand side
template <class A, class... R,
          class = std::enable_if_t<is_operand<A>::value && (is_operand<R>::value && ...)>, class = void>
inline auto tie(const A& a, const R&... rest) {
  using T = typename value_of<A>::type;
  return tie_exprs<T, decltype(as_expr(a).n), decltype(as_expr(rest).n)...>{
      detail::make_nodes(as_expr(a).n, as_expr(rest).n...),
      detail::merge_all(as_expr(a).ext, as_expr(rest).ext...)};
}

// Black-Scholes call and put of the five input tables (one launch):
//   tie(call, put) = blackscholes(S, K, T, r, sig);
template <class T>
inline bs_call_put<T> blackscholes(const table<T>& S, const table<T>& K, const table<T>& Tm, const table<T>& r,
                                   const table<T>& sig, int precision = NT2H_EXACT) {
  const extent x = S.extent_();
  if (!(K.extent_().same(x) && Tm.extent_().same(x) && r.extent_().same(x) && sig.extent_().same(x)))
    throw error("nt2hip::blackscholes: size mismatch", NT2H_ERR_INVALID_ARG);
  return {&S, &K, &Tm, &r, &sig, precision};
}

// Forward + adjoint of the portfolio sum(w .* call): price and the
// w-weighted cotangents of every input; returns {sum dr, sum dsig}.
struct bs_totals {
  double dP_dr, dP_dsig;
};
template <class T>
inline bs_totals blackscholes_adjoint(const table<T>& S, const table<T>& K, const table<T>& Tm, const table<T>& r,
                                      const table<T>& sig, const table<T>& w, table<T>& price, table<T>& dS,
                                      table<T>& dK, table<T>& dT, table<T>& dr, table<T>& dsig) {
  const extent x = S.extent_();
  for (table<T>* o : {&price, &dS, &dK, &dT, &dr, &dsig}) o->resize(x);
  double tot[2] = {0.0, 0.0};
  int st = std::is_same<T, float>::value
               ? nt2h_bs_fwd_adj_f32((const float*)S.data(), (const float*)K.data(), (const float*)Tm.data(),
                                     (const float*)r.data(), (const float*)sig.data(), (const float*)w.data(),
                                     (float*)price.data(), (float*)dS.data(), (float*)dK.data(), (float*)dT.data(),
                                     (float*)dr.data(), (float*)dsig.data(), tot, x.numel(), abi_stream())
               : nt2h_bs_fwd_adj_f64((const double*)S.data(), (const double*)K.data(), (const double*)Tm.data(),
                                     (const double*)r.data(), (const double*)sig.data(), (const double*)w.data(),
                                     (double*)price.data(), (double*)dS.data(), (double*)dK.data(),
                                     (double*)dT.data(), (double*)dr.data(), (double*)dsig.data(), tot, x.numel(),
                                     abi_stream());
  check(st, "nt2h_bs_fwd_adj");
  return {tot[0], tot[1]};
}

// Column fold z = sum(exp(-sqr(X)/(2*sqr(sx(s)))) .* cos(sx(k).*X + sx(p)), 1) and,
// when the gradient tables are given, its adjoint for cotangent gz (empty = ones).
inline void colfold_adjoint(const table<float>& X, const table<float>& s, const table<float>& k,
                            const table<float>& p, const table<float>* gz, table<float>& z, table<float>* dz_ds,
                            table<float>* dz_dk, table<float>* dz_dp) {
  const extent out(1, X.cols());
  z.resize(out);
  if (dz_ds) { dz_ds->resize(out); dz_dk->resize(out); dz_dp->resize(out); }
  check(nt2h_colfold_gc_f32(X.data(), X.rows(), X.cols(), X.rows(), s.data(), k.data(), p.data(),
                            gz ? gz->data() : nullptr, z.data(), dz_ds ? dz_ds->data() : nullptr,
                            dz_dk ? dz_dk->data() : nullptr, dz_dp ? dz_dp->data() : nullptr, abi_stream()),
        "nt2h_colfold_gc_f32");
}

}  // namespace nt2hip
