// nt2hip/table.hpp — Matlab-like device tables for C++ callers (compile with hipcc).
//
// The caller writes NT2's table expressions and gets NT2's results:
//
//   nt2hip::table<float> X(n), Y(n);            // device-resident, column-major
//   Y = 2.5f * X + Y;                            // one launch: nt2h_saxpy_f32
//   Y = exp(-sqr(X) / 2.0f) * cos(X);            // one fused kernel for the whole tree
//   float s = globalsum(X * Y);                  // nt2h_dot_f32
//   table<float> z = sum(exp(-sqr(A)) * cos(sx(k) * A), 1);   // inner_fold over rows
//   tie(call, put) = blackscholes(S, K, T, r, sig);           // 5 -> 2 streams, one launch
//   tie(da, d1, d2, R) = tie(sqrt(T), ..., fnms(v, da, d1), ...); // N outputs, one kernel
//
// It stands where NT2's container DSL does: table<T,S> (core/container/
// table/table.hpp:34-171), table::operator= -> evaluate -> run_assign
// (core/container/dsl/expression.hpp:161-191, dsl/functions/elementwise/
// run.hpp:21-35), the transform loop nest (core/functions/transform.hpp:
// 55-69), terminal load/store (dsl/functions/container/terminal.hpp:28-129),
// the per-node recursion run(node, i) (dsl/functions/generic/run.hpp:
// 128-222), the multiply-adds fma/fms/fnms/fnma (boost/simd/arithmetic/
// functions/generic/{fma,fms,fnms,fnma}.hpp), the N-ary tie (core/functions/
// expr/tie.hpp:23-160), the reductions
// (dsl/functions/reduction/run.hpp:30-104: inner_fold along dimension 1,
// outer_fold along dimension 2) and the sx broadcast (core/functions/sx.hpp).
//
// Lowering: an assignment is first matched against the loop nests that
// libnt2hip implements as hand-written kernels (saxpy, one functor of a
// terminal, globalsum of a terminal or of a product of two terminals,
// Black-Scholes); any other tree becomes one kernel instantiated for that
// tree from the same device functors (include/nt2hip/math.hpp): every node
// is evaluated in registers at the element's index, nothing is
// materialised between nodes.  Arithmetic follows NT2's scalar path
// (nt2hip::exact, the default) or the FMA policy (nt2hip::fused) via
// assign<fused>(lhs, expr).  Errors throw nt2hip::error.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../nt2hip.h"
#include "fold.hpp"
#include "math.hpp"

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

struct extent {
  int64_t rows = 0, cols = 0;
  bool row_bcast = false;  // a 1 x cols row that sx() repeats down every column
  int64_t numel() const { return rows * cols; }
  bool operator==(const extent& o) const { return rows == o.rows && cols == o.cols; }
};

template <class T> class table;

// ============================================================ nodes
// Four consecutive elements of a loop nest (the vector body of the kernels).
template <class T> struct quad {
  T v[4];
};

namespace node {
// Every node evaluates at(i, rows): linear index i of a column-major
// (rows x cols) loop nest (rows lets an sx-broadcast row find its column),
// and at4(i, rows): elements i..i+3 for i a multiple of 4, with 16-byte loads
// of the tables.  aligned() says whether every table of the tree starts on
// a 16-byte boundary (the condition for at4).
template <class T> struct term {  // a table read at the element index
  const T* p;
  template <class P> NT2H_FD T at(int64_t i, int64_t) const { return p[i]; }
  template <class P> NT2H_FD quad<T> at4(int64_t i, int64_t) const {
    using V = typename nt2h::Vec16<T>::type;
    constexpr int W = nt2h::Vec16<T>::n;
    quad<T> q;
#pragma unroll
    for (int h = 0; h < 4 / W; ++h) {
      V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p + i) + h);
#pragma unroll
      for (int k = 0; k < W; ++k) q.v[h * W + k] = v[k];
    }
    return q;
  }
  bool aligned() const { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
  bool bcast_reads(const void*) const { return false; }
};
template <class T> struct row_bcast {  // sx: a 1 x cols row repeated down every column
  const T* p;
  template <class P> NT2H_FD T at(int64_t i, int64_t rows) const { return p[i / rows]; }
  template <class P> NT2H_FD quad<T> at4(int64_t i, int64_t rows) const {
    quad<T> q;
    const int64_t c0 = i / rows;
    if ((i + 3) / rows == c0) {  // the four elements share a column
      const T v = p[c0];
#pragma unroll
      for (int k = 0; k < 4; ++k) q.v[k] = v;
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) q.v[k] = p[(i + k) / rows];
    }
    return q;
  }
  bool aligned() const { return true; }
  bool bcast_reads(const void* q) const { return q == static_cast<const void*>(p); }
};
template <class T> struct scalar {
  T v;
  template <class P> NT2H_FD T at(int64_t, int64_t) const { return v; }
  template <class P> NT2H_FD quad<T> at4(int64_t, int64_t) const { return quad<T>{{v, v, v, v}}; }
  bool aligned() const { return true; }
  bool bcast_reads(const void*) const { return false; }
};
template <class F, class A> struct unary {
  A a;
  template <class P> NT2H_FD auto at(int64_t i, int64_t r) const { return F::template apply<P>(a.template at<P>(i, r)); }
  template <class P> NT2H_FD auto at4(int64_t i, int64_t r) const {
    const auto x = a.template at4<P>(i, r);
    quad<decltype(F::template apply<P>(x.v[0]))> q;
#pragma unroll
    for (int k = 0; k < 4; ++k) q.v[k] = F::template apply<P>(x.v[k]);
    return q;
  }
  bool aligned() const { return a.aligned(); }
  bool bcast_reads(const void* p) const { return a.bcast_reads(p); }
};
template <class F, class A, class B> struct binary {
  A a;
  B b;
  template <class P> NT2H_FD auto at(int64_t i, int64_t r) const {
    return F::template apply<P>(a.template at<P>(i, r), b.template at<P>(i, r));
  }
  template <class P> NT2H_FD auto at4(int64_t i, int64_t r) const {
    const auto x = a.template at4<P>(i, r);
    const auto y = b.template at4<P>(i, r);
    quad<decltype(F::template apply<P>(x.v[0], y.v[0]))> q;
#pragma unroll
    for (int k = 0; k < 4; ++k) q.v[k] = F::template apply<P>(x.v[k], y.v[k]);
    return q;
  }
  bool aligned() const { return a.aligned() && b.aligned(); }
  bool bcast_reads(const void* p) const { return a.bcast_reads(p) || b.bcast_reads(p); }
};
template <class F, class A, class B, class C> struct ternary {
  A a;
  B b;
  C c;
  template <class P> NT2H_FD auto at(int64_t i, int64_t r) const {
    return F::template apply<P>(a.template at<P>(i, r), b.template at<P>(i, r), c.template at<P>(i, r));
  }
  template <class P> NT2H_FD auto at4(int64_t i, int64_t r) const {
    const auto x = a.template at4<P>(i, r);
    const auto y = b.template at4<P>(i, r);
    const auto z = c.template at4<P>(i, r);
    quad<decltype(F::template apply<P>(x.v[0], y.v[0], z.v[0]))> q;
#pragma unroll
    for (int k = 0; k < 4; ++k) q.v[k] = F::template apply<P>(x.v[k], y.v[k], z.v[k]);
    return q;
  }
  bool aligned() const { return a.aligned() && b.aligned() && c.aligned(); }
  bool bcast_reads(const void* p) const { return a.bcast_reads(p) || b.bcast_reads(p) || c.bcast_reads(p); }
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

// ==================================================== expression wrapper
template <class T, class N> struct expr {
  N n;
  extent ext;  // {0, 0} for a scalar
};

template <class T> struct is_table : std::false_type {};
template <class T> struct is_table<table<T>> : std::true_type {};
template <class X> struct is_expr : std::false_type {};
template <class T, class N> struct is_expr<expr<T, N>> : std::true_type {};

template <class T> inline expr<T, node::term<T>> as_expr(const table<T>& t) {
  return {node::term<T>{t.data()}, t.extent_()};
}
template <class T, class N> inline const expr<T, N>& as_expr(const expr<T, N>& e) { return e; }

template <class T> inline extent merge(const extent& a, const extent& b) {
  if (a.numel() == 0 && a.rows == 0 && !a.row_bcast) return b;  // scalar operand
  if (b.numel() == 0 && b.rows == 0 && !b.row_bcast) return a;
  if (a.row_bcast && !b.row_bcast && a.cols == b.cols) return b;  // sx(row) against a full table
  if (b.row_bcast && !a.row_bcast && a.cols == b.cols) return a;
  if (!(a == b)) throw error("nt2hip: size mismatch in a table expression", NT2H_ERR_INVALID_ARG);
  return a;
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
  return {node::binary<F, NA, NB>{a.n, b.n}, merge<TA>(a.ext, b.ext)};
}
template <class T, class S> inline expr<T, node::scalar<T>> scal(S v) { return {node::scalar<T>{T(v)}, extent{}}; }

#define NT2HIP_BOP(OP, F)                                                                           \
  template <class A, class B,                                                                       \
            class = std::enable_if_t<(is_table<A>::value || is_expr<A>::value) &&                   \
                                     (is_table<B>::value || is_expr<B>::value)>>                    \
  inline auto operator OP(const A& a, const B& b) {                                                 \
    return make_bin<fn::F>(as_expr(a), as_expr(b));                                                 \
  }                                                                                                 \
  template <class A, class S,                                                                       \
            class = std::enable_if_t<(is_table<A>::value || is_expr<A>::value) &&                   \
                                     std::is_arithmetic<S>::value>>                                 \
  inline auto operator OP(const A& a, S s) {                                                        \
    using T = typename value_of<A>::type;                                                           \
    return make_bin<fn::F>(as_expr(a), scal<T>(s));                                                 \
  }                                                                                                 \
  template <class S, class B,                                                                       \
            class = std::enable_if_t<std::is_arithmetic<S>::value &&                                \
                                     (is_table<B>::value || is_expr<B>::value)>,                    \
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
template <class X> struct is_operand : std::integral_constant<bool, is_table<X>::value || is_expr<X>::value> {};
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
  return expr<T, N>{N{ea.n, eb.n, ec.n}, merge<T>(merge<T>(ea.ext, eb.ext), ec.ext)};
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

// sx(row): a 1 x cols table repeated down the rows of whatever (rows x cols)
// operand it meets (core/functions/sx.hpp); on its own it is 1 x cols.
template <class T> inline expr<T, node::row_bcast<T>> sx(const table<T>& row) {
  if (row.rows() != 1) throw error("nt2hip::sx: expects a 1 x cols row", NT2H_ERR_INVALID_ARG);
  return {node::row_bcast<T>{row.data()}, extent{1, row.cols(), true}};
}

// ============================================================ kernels
namespace detail {
template <class P, class T, class N>
__global__ __launch_bounds__(256) void k_transform(N n, T* out, int64_t numel, int64_t rows) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < numel; i += stride) out[i] = n.template at<P>(i, rows);
}
// Vector body: lane l evaluates elements 4q..4q+3 of quad q (16-byte loads
// and stores, one quad per lane and step); the last numel % 4 elements go
// through at().
template <class P, class T, class N>
__global__ __launch_bounds__(256) void k_transform4(N n, T* out, int64_t numel, int64_t rows) {
  using V = typename nt2h::Vec16<T>::type;
  constexpr int W = nt2h::Vec16<T>::n;
  const int64_t nq = numel / 4;
  const int64_t stride = (int64_t)gridDim.x * 256;
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (int64_t q = g; q < nq; q += stride) {
    const quad<T> r = n.template at4<P>(4 * q, rows);
#pragma unroll
    for (int h = 0; h < 4 / W; ++h) {
      V v;
#pragma unroll
      for (int k = 0; k < W; ++k) v[k] = r.v[h * W + k];
      __builtin_nontemporal_store(v, reinterpret_cast<V*>(out + 4 * q) + h);
    }
  }
  const int64_t t = 4 * nq + g;
  if (t < numel) out[t] = n.template at<P>(t, rows);
}
// tie's loop nests: at every element (every quad in the vector body) all
// children are evaluated before any output is stored, so an output may be
// an input of another child (core/functions/expr/tie.hpp:80-115 reads the
// whole fusion sequence, then stores it).
template <class... Ns> struct nodes;
template <> struct nodes<> {
  bool aligned() const { return true; }
  bool bcast_reads(const void*) const { return false; }
};
template <class H, class... R> struct nodes<H, R...> {
  H h;
  nodes<R...> t;
  bool aligned() const { return h.aligned() && t.aligned(); }
  bool bcast_reads(const void* p) const { return h.bcast_reads(p) || t.bcast_reads(p); }
};
template <class T, int N> struct out_ptrs {
  T* p[N];
};
template <class P, class T> NT2H_FD void eval_all(const nodes<>&, int64_t, int64_t, T*) {}
template <class P, class T, class H, class... R>
NT2H_FD void eval_all(const nodes<H, R...>& ns, int64_t i, int64_t rows, T* v) {
  v[0] = ns.h.template at<P>(i, rows);
  eval_all<P, T>(ns.t, i, rows, v + 1);
}
template <class P, class T> NT2H_FD void eval_all4(const nodes<>&, int64_t, int64_t, quad<T>*) {}
template <class P, class T, class H, class... R>
NT2H_FD void eval_all4(const nodes<H, R...>& ns, int64_t i, int64_t rows, quad<T>* q) {
  q[0] = ns.h.template at4<P>(i, rows);
  eval_all4<P, T>(ns.t, i, rows, q + 1);
}
template <class P, class T, int N, class Ns>
__global__ __launch_bounds__(256) void k_tie4(Ns ns, out_ptrs<T, N> o, int64_t numel, int64_t rows) {
  using V = typename nt2h::Vec16<T>::type;
  constexpr int W = nt2h::Vec16<T>::n;
  const int64_t nq = numel / 4;
  const int64_t stride = (int64_t)gridDim.x * 256;
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (int64_t q = g; q < nq; q += stride) {
    quad<T> v[N];
    eval_all4<P, T>(ns, 4 * q, rows, v);
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
    eval_all<P, T>(ns, t, rows, v);
#pragma unroll
    for (int j = 0; j < N; ++j) o.p[j][t] = v[j];
  }
}
template <class P, class T, int N, class Ns>
__global__ __launch_bounds__(256) void k_tie(Ns ns, out_ptrs<T, N> o, int64_t numel, int64_t rows) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < numel; i += stride) {
    T v[N];
    eval_all<P, T>(ns, i, rows, v);
#pragma unroll
    for (int j = 0; j < N; ++j) o.p[j][i] = v[j];
  }
}

// Per-lane accumulation in the value type's working precision (float:
// TwoSum-compensated pair; double: plain double), and the workgroup
// reduction to one double in thread 0.
template <class T> struct Acc;
template <> struct Acc<float> {
  nt2h::Comp c{0.0f, 0.0f};
  NT2H_FD void add(float x) { c.add(x); }
  NT2H_FD double finish() const { return (double)c.s + (double)c.c; }
  NT2H_FD double block_total() {
    __shared__ float lds[2 * nt2h::kWaves];
    nt2h::Comp a[1] = {c};
    double t[1] = {0.0};  // written by thread 0 only; defined everywhere
    nt2h::block_reduce_comp<1>(a, lds, t);
    return t[0];
  }
};
template <> struct Acc<double> {
  double s = 0.0;
  NT2H_FD void add(double x) { s += x; }
  NT2H_FD double finish() const { return s; }
  NT2H_FD double block_total() {
    __shared__ double lds[nt2h::kWaves];
    double v = s;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += nt2h::shfl_xor_d(v, m);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
      for (int w = 0; w < nt2h::kWaves; ++w) t += lds[w];
    return t;
  }
};

template <class P, class T, class N>
__global__ __launch_bounds__(256) void k_globalsum(N n, int64_t numel, int64_t rows, nt2h::Workspace ws,
                                                   double* out) {
  Acc<T> acc;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < numel; i += stride) acc.add(n.template at<P>(i, rows));
  double bsum[1] = {acc.block_total()};
  double res[1];
  if (nt2h::grid_combine<1>(bsum, ws, res)) out[0] = res[0];
}
// inner_fold along dimension 1: one workgroup per column
template <class P, class T, class N>
__global__ __launch_bounds__(256) void k_colsum(N n, int64_t rows, T* out) {
  Acc<T> acc;
  const int64_t base = (int64_t)blockIdx.x * rows;
  for (int64_t i = threadIdx.x; i < rows; i += 256) acc.add(n.template at<P>(base + i, rows));
  double t = acc.block_total();
  if (threadIdx.x == 0) out[blockIdx.x] = (T)t;
}
// outer_fold along dimension 2: one lane per row, columns walked in order
// (consecutive lanes read consecutive rows of a column: coalesced)
template <class P, class T, class N>
__global__ __launch_bounds__(256) void k_rowsum(N n, int64_t rows, int64_t cols, T* out) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  Acc<T> acc;
  for (int64_t j = 0; j < cols; ++j) acc.add(n.template at<P>(j * rows + i, rows));
  out[i] = (T)acc.finish();
}

inline unsigned grid_for(int64_t numel) {
  int64_t b = (numel + 255) / 256;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (unsigned)b;
}

// Reduction workspace of the front end, one per device (calls on one
// device are serialised on the current stream).
inline nt2h::Workspace workspace() {
  static std::mutex m;
  static std::map<int, nt2h::Workspace> cache;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(m);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  nt2h::Workspace w{};
  hip_check(hipMalloc((void**)&w.partials, sizeof(double) * nt2h::kMaxPartials * nt2h::kMaxFoldBlocks), "hipMalloc");
  hip_check(hipMalloc((void**)&w.gpartials, sizeof(double) * nt2h::kMaxPartials * nt2h::kMaxGroups), "hipMalloc");
  size_t cb = sizeof(unsigned) * (nt2h::kMaxGroups + 1) * nt2h::kCounterStride;
  hip_check(hipMalloc((void**)&w.counters, cb), "hipMalloc");
  hip_check(hipMalloc((void**)&w.result, sizeof(double) * nt2h::kMaxPartials), "hipMalloc");
  hip_check(hipMemset(w.counters, 0, cb), "hipMemset");
  cache[dev] = w;
  return w;
}
}  // namespace detail

// ============================================================= table
template <class T> class table {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "table<float> or table<double>");

 public:
  table() = default;
  explicit table(int64_t n) { resize(extent{n, 1}); }
  table(int64_t rows, int64_t cols) { resize(extent{rows, cols}); }
  // host data, column-major
  table(const std::vector<T>& host, int64_t rows, int64_t cols) : table(rows, cols) { upload(host); }
  explicit table(const std::vector<T>& host) : table(host, (int64_t)host.size(), 1) {}
  table(const table& o) : table(o.ext_.rows, o.ext_.cols) { copy_from(o); }
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
  int64_t rows() const { return ext_.rows; }
  int64_t cols() const { return ext_.cols; }
  extent extent_() const { return ext_; }

  void resize(extent e) {
    if (e.rows < 0 || e.cols < 0) throw error("nt2hip::table resize", NT2H_ERR_INVALID_ARG);
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
template <class P, class T, class N> inline void run_generic(T* out, const N& n, extent x) {
  if (x.numel() == 0) return;
  if (n.aligned() && (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
    int64_t b = (x.numel() / 4 + 255) / 256;
    if (b < 1) b = 1;
    if (b > (int64_t(1) << 30)) b = int64_t(1) << 30;
    hipLaunchKernelGGL((k_transform4<P, T, N>), dim3((unsigned)b), dim3(256), 0, current_stream(), n, out,
                       x.numel(), x.rows);
  } else {
    hipLaunchKernelGGL((k_transform<P, T, N>), dim3(grid_for(x.numel())), dim3(256), 0, current_stream(), n, out,
                       x.numel(), x.rows);
  }
  hip_check(hipGetLastError(), "transform launch");
}
template <class P, class T, class N> struct lower {
  static void run(T* out, const N& n, extent x) { run_generic<P, T, N>(out, n, x); }
};
// Y = alpha*X + Y  ->  nt2h_saxpy_f32 (when the last operand is the target)
template <class P>
struct lower<P, float,
             node::binary<fn::plus_, node::binary<fn::multiplies_, node::scalar<float>, node::term<float>>,
                          node::term<float>>> {
  using N = node::binary<fn::plus_, node::binary<fn::multiplies_, node::scalar<float>, node::term<float>>,
                         node::term<float>>;
  static void run(float* out, const N& n, extent x) {
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
  static void run(T* out, const N& n, extent x) {
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
}  // namespace detail

template <class T> template <class N> table<T>& table<T>::operator=(const expr<T, N>& e) {
  extent x = e.ext;
  x.row_bcast = false;
  resize(x);
  detail::lower<exact, T, N>::run(p_, e.n, x);
  return *this;
}

// assign<fused>(lhs, expr): the same loop nest with the FMA policy
template <class P, class T, class N> inline table<T>& assign(table<T>& lhs, const expr<T, N>& e) {
  extent x = e.ext;
  x.row_bcast = false;
  lhs.resize(x);
  detail::lower<P, T, N>::run(lhs.data(), e.n, x);
  return lhs;
}

// ========================================================= reductions
namespace detail {
template <class P, class T, class N> inline double globalsum_generic(const N& n, int64_t numel, int64_t rows) {
  nt2h::Workspace ws = workspace();
  double out = 0.0;
  if (numel > 0) {
    hipLaunchKernelGGL((k_globalsum<P, T, N>), dim3(nt2h::fold_blocks(numel)), dim3(256), 0, current_stream(), n,
                       numel, rows, ws, ws.result);
    hip_check(hipGetLastError(), "globalsum launch");
    hip_check(hipMemcpyAsync(&out, ws.result, sizeof(double), hipMemcpyDeviceToHost, current_stream()), "copy");
    hip_check(hipStreamSynchronize(current_stream()), "sync");
  }
  return out;
}
template <class T, class N> struct gsum {
  static double run(const N& n, int64_t numel, int64_t rows) { return globalsum_generic<exact, T, N>(n, numel, rows); }
};
template <class T> struct gsum<T, node::term<T>> {  // globalsum(X) -> nt2h_sum
  static double run(const node::term<T>& n, int64_t numel, int64_t) {
    double out = 0.0;
    int st = std::is_same<T, float>::value
                 ? nt2h_sum_f32((const float*)n.p, numel, NT2H_SUM_KAHAN, &out, abi_stream())
                 : nt2h_sum_f64((const double*)n.p, numel, &out, abi_stream());
    check(st, "nt2h_sum");
    return out;
  }
};
template <class T> struct gsum<T, node::binary<fn::multiplies_, node::term<T>, node::term<T>>> {  // -> nt2h_dot
  static double run(const node::binary<fn::multiplies_, node::term<T>, node::term<T>>& n, int64_t numel, int64_t) {
    double out = 0.0;
    int st = std::is_same<T, float>::value
                 ? nt2h_dot_f32((const float*)n.a.p, (const float*)n.b.p, numel, NT2H_SUM_KAHAN, &out, abi_stream())
                 : nt2h_dot_f64((const double*)n.a.p, (const double*)n.b.p, numel, &out, abi_stream());
    check(st, "nt2h_dot");
    return out;
  }
};
}  // namespace detail

// globalsum(e): sum of every element (core/functions/globalsum.hpp:45); the
// value is the compensated FP64-combined sum rounded to T.
template <class T, class N> inline T globalsum(const expr<T, N>& e) {
  return (T)detail::gsum<T, N>::run(e.n, e.ext.numel(), e.ext.rows);
}
template <class T> inline T globalsum(const table<T>& t) { return globalsum(as_expr(t)); }
// dot(a, b) = globalsum(a .* b) (reduction/functions/expr/dot.hpp:28)
template <class A, class B> inline auto dot(const A& a, const B& b) { return globalsum(as_expr(a) * as_expr(b)); }

// sum(e, dim): dim 1 -> 1 x cols (inner_fold), dim 2 -> rows x 1 (outer_fold)
template <class T, class N> inline table<T> sum(const expr<T, N>& e, int dim) {
  const extent x = e.ext;
  if (dim == 1) {
    table<T> out(1, x.cols);
    if (x.cols > 0 && x.rows > 0) {
      hipLaunchKernelGGL((detail::k_colsum<exact, T, N>), dim3((unsigned)x.cols), dim3(256), 0, current_stream(), e.n,
                         x.rows, out.data());
      hip_check(hipGetLastError(), "sum(.,1) launch");
    } else if (x.cols > 0) {
      out = T(0);
    }
    return out;
  }
  if (dim == 2) {
    table<T> out(x.rows, 1);
    if (x.rows > 0) {
      hipLaunchKernelGGL((detail::k_rowsum<exact, T, N>), dim3((unsigned)((x.rows + 255) / 256)), dim3(256), 0,
                         current_stream(), e.n, x.rows, x.cols, out.data());
      hip_check(hipGetLastError(), "sum(.,2) launch");
    }
    return out;
  }
  throw error("nt2hip::sum: dimension 1 or 2", NT2H_ERR_INVALID_ARG);
}
template <class T> inline table<T> sum(const table<T>& t, int dim) { return sum(as_expr(t), dim); }
// sum(e): along the first non-singleton dimension, as NT2's sum
template <class T, class N> inline table<T> sum(const expr<T, N>& e) {
  if (e.ext.rows == 1 || e.ext.cols == 1) {
    table<T> out(1, 1);
    out = globalsum(e);
    return out;
  }
  return sum(e, 1);
}
template <class T> inline table<T> sum(const table<T>& t) { return sum(as_expr(t)); }

// ================================== other reduction_ functors and scans
// prod, maximum, minimum, asum1, asum2, norm2, mean (core/reduction/include/
// nt2/core/functions/{prod,...}.hpp) and cumsum, cumprod, cummax, cummin
// (core/swar) along dimension 1 (1 x cols result), 2 (rows x 1) or a
// singleton dimension >= 3 (the element itself through the functor), or
// along the first non-singleton dimension when none is given, as NT2's; the
// globalX forms fold every element.  Expressions are evaluated into a
// temporary first; these lower to nt2h_fold_f32/f64 and nt2h_scan_f32/f64.
namespace detail {
inline void lmh2(extent x, int dim, int64_t& lo, int64_t& mid, int64_t& hi) {
  if (dim < 1) throw error("nt2hip: dimensions are 1-based", NT2H_ERR_INVALID_ARG);
  if (dim == 1) { lo = 1; mid = x.rows; hi = x.cols; }
  else if (dim == 2) { lo = x.rows; mid = x.cols; hi = 1; }
  else { lo = x.numel(); mid = 1; hi = 1; }
}
inline int first_nonsingleton(extent x) { return x.rows != 1 ? 1 : (x.cols != 1 ? 2 : 1); }
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
template <class T> inline table<T> fold_dim(int op, const table<T>& t, int dim) {
  const extent x = t.extent_();
  int64_t lo, mid, hi;
  lmh2(x, dim, lo, mid, hi);
  table<T> out(dim == 1 ? 1 : x.rows, dim == 2 ? 1 : x.cols);
  check(fold_abi(op, t.data(), lo, mid, hi, out.data()), "nt2h_fold");
  return out;
}
template <class T> inline T fold_all(int op, const table<T>& t) {
  table<T> out(1, 1);
  check(fold_abi(op, t.data(), 1, t.numel(), 1, out.data()), "nt2h_fold");
  return out.to_host()[0];
}
template <class T> inline table<T> scan_dim(int op, const table<T>& t, int dim) {
  const extent x = t.extent_();
  int64_t lo, mid, hi;
  lmh2(x, dim, lo, mid, hi);
  table<T> out(x.rows, x.cols);
  check(scan_abi(op, t.data(), lo, mid, hi, out.data()), "nt2h_scan");
  return out;
}
}  // namespace detail

#define NT2HIP_FOLD_FUNCTOR(NAME, GLOBAL, OP)                                                                  \
  template <class T> inline table<T> NAME(const table<T>& t, int dim) { return detail::fold_dim(OP, t, dim); } \
  template <class T> inline table<T> NAME(const table<T>& t) {                                                 \
    return detail::fold_dim(OP, t, detail::first_nonsingleton(t.extent_()));                                   \
  }                                                                                                            \
  template <class T, class N> inline table<T> NAME(const expr<T, N>& e, int dim) { return NAME(table<T>(e), dim); } \
  template <class T, class N> inline table<T> NAME(const expr<T, N>& e) { return NAME(table<T>(e)); }         \
  template <class T> inline T GLOBAL(const table<T>& t) { return detail::fold_all(OP, t); }                    \
  template <class T, class N> inline T GLOBAL(const expr<T, N>& e) { return GLOBAL(table<T>(e)); }
NT2HIP_FOLD_FUNCTOR(prod, globalprod, NT2H_FOLD_PROD)
NT2HIP_FOLD_FUNCTOR(maximum, globalmax, NT2H_FOLD_MAXIMUM)
NT2HIP_FOLD_FUNCTOR(minimum, globalmin, NT2H_FOLD_MINIMUM)
NT2HIP_FOLD_FUNCTOR(asum1, globalasum1, NT2H_FOLD_ASUM1)
NT2HIP_FOLD_FUNCTOR(asum2, globalasum2, NT2H_FOLD_ASUM2)
NT2HIP_FOLD_FUNCTOR(norm2, globalnorm2, NT2H_FOLD_NORM2)
NT2HIP_FOLD_FUNCTOR(mean, globalmean, NT2H_FOLD_MEAN)
#undef NT2HIP_FOLD_FUNCTOR

#define NT2HIP_SCAN_FUNCTOR(NAME, OP)                                                                          \
  template <class T> inline table<T> NAME(const table<T>& t, int dim) { return detail::scan_dim(OP, t, dim); } \
  template <class T> inline table<T> NAME(const table<T>& t) {                                                 \
    return detail::scan_dim(OP, t, detail::first_nonsingleton(t.extent_()));                                   \
  }                                                                                                            \
  template <class T, class N> inline table<T> NAME(const expr<T, N>& e, int dim) { return NAME(table<T>(e), dim); } \
  template <class T, class N> inline table<T> NAME(const expr<T, N>& e) { return NAME(table<T>(e)); }
NT2HIP_SCAN_FUNCTOR(cumsum, NT2H_SCAN_CUMSUM)
NT2HIP_SCAN_FUNCTOR(cumprod, NT2H_SCAN_CUMPROD)
NT2HIP_SCAN_FUNCTOR(cummax, NT2H_SCAN_CUMMAX)
NT2HIP_SCAN_FUNCTOR(cummin, NT2H_SCAN_CUMMIN)
#undef NT2HIP_SCAN_FUNCTOR

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
template <class T> inline extent merge_all(const extent& a) { return a; }
template <class T, class... R> inline extent merge_all(const extent& a, const extent& b, const R&... r) {
  return merge_all<T>(merge<T>(a, b), r...);
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
                                                       detail::merge_all<T>(o[I]->extent_()...)};
  }
  template <class P, class... Ns> void assign_all(const tie_exprs<T, Ns...>& es) {
    extent x = es.ext;
    x.row_bcast = false;
    // an output that some child reads through sx() (at other elements) is
    // written only after the whole loop nest: evaluate into fresh tables
    bool stage = false;
    for (int j = 0; j < N; ++j) stage = stage || es.ns.bcast_reads(o[j]->data());
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
        int64_t nb = (numel / 4 + 255) / 256;
        if (nb < 1) nb = 1;
        if (nb > (int64_t(1) << 30)) nb = int64_t(1) << 30;
        hipLaunchKernelGGL((detail::k_tie4<P, T, N, NS>), dim3((unsigned)nb), dim3(256), 0, current_stream(), es.ns,
                           dst, numel, x.rows);
      } else {
        hipLaunchKernelGGL((detail::k_tie<P, T, N, NS>), dim3(detail::grid_for(numel)), dim3(256), 0,
                           current_stream(), es.ns, dst, numel, x.rows);
      }
      hip_check(hipGetLastError(), "tie launch");
    }
    if (stage)
      for (int j = 0; j < N; ++j) *o[j] = std::move(tmp[j]);
  }

  // tie(call, put) = blackscholes(S, K, T, r, sig): nt2h_bs_fwd_f32/f64
  void operator=(const bs_call_put<T>& q) {
    static_assert(N == 2, "blackscholes yields two tables: tie(call, put)");
    extent x = q.S->extent_();
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
          class = std::enable_if_t<(is_table<A>::value || is_expr<A>::value) &&
                                   ((is_table<R>::value || is_expr<R>::value) && ...)>,
          class = void>
inline auto tie(const A& a, const R&... rest) {
  using T = typename value_of<A>::type;
  return tie_exprs<T, decltype(as_expr(a).n), decltype(as_expr(rest).n)...>{
      detail::make_nodes(as_expr(a).n, as_expr(rest).n...),
      detail::merge_all<T>(as_expr(a).ext, as_expr(rest).ext...)};
}

// Black-Scholes call and put of the five input tables (one launch):
//   tie(call, put) = blackscholes(S, K, T, r, sig);
template <class T>
inline bs_call_put<T> blackscholes(const table<T>& S, const table<T>& K, const table<T>& Tm, const table<T>& r,
                                   const table<T>& sig, int precision = NT2H_EXACT) {
  extent x = S.extent_();
  if (!(K.extent_() == x && Tm.extent_() == x && r.extent_() == x && sig.extent_() == x))
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
  extent x = S.extent_();
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
  extent out{1, X.cols()};
  z.resize(out);
  if (dz_ds) { dz_ds->resize(out); dz_dk->resize(out); dz_dp->resize(out); }
  check(nt2h_colfold_gc_f32(X.data(), X.rows(), X.cols(), X.rows(), s.data(), k.data(), p.data(),
                            gz ? gz->data() : nullptr, z.data(), dz_ds ? dz_ds->data() : nullptr,
                            dz_dk ? dz_dk->data() : nullptr, dz_dp ? dz_dp->data() : nullptr, abi_stream()),
        "nt2h_colfold_gc_f32");
}

}  // namespace nt2hip
