// Shared by the C++ front-end check program (test_table.cpp and the nd_*.cpp
// units, compiled separately so that they build in parallel): the oracle's
// entry points, the report, input generators and bit comparisons, and the
// fold/scan checks of N-D expression trees.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/nt2hip/table.hpp"

extern "C" {
void nt2ref_map_f32(int op, const float* x, float* y, int64_t n);
void nt2ref_map_f64(int op, const double* x, double* y, int64_t n);
void nt2ref_saxpy_f32(float a, const float* x, float* y, int64_t n);
void nt2ref_bs_fwd_f32(const float*, const float*, const float*, const float*, const float*, float*, float*, int64_t);
void nt2ref_exact_dot_f32(const float* x, const float* y, int64_t n, double* out);
void nt2ref_exact_sum_f32(const float* x, int64_t n, double* out);
void nt2ref_fold_lmh_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* out);
void nt2ref_scan_lmh_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* y);
void nt2ref_fold_lmh_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* out);
void nt2ref_scan_lmh_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* y);
}

using nt2hip::check;
using nt2hip::table;
inline void hip_ok(hipError_t e) { nt2hip::hip_check(e, "hip"); }
// one "PASS name" / "FAIL name detail" line per check (test_table.cpp)
void report(const char* name, bool ok, const char* detail = "");
inline uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline bool same_bits(const std::vector<float>& a, const std::vector<float>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (bits(a[i]) != bits(b[i]) && !(std::isnan(a[i]) && std::isnan(b[i]))) return false;
  return true;
}
inline std::vector<float> uniform(size_t n, float lo, float hi, uint64_t seed) {
  std::vector<float> v(n);
  uint64_t z = seed;
  for (auto& x : v) {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t t = z;
    t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
    t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
    t ^= t >> 31;
    x = (float)(lo + (hi - lo) * ((t >> 11) * 0x1.0p-53));
  }
  return v;
}

// ---------------------------------------------------------------------------
// N-D tables: singleton expansion along any dimension, and every reduction
// and scan of an expression tree along every dimension.  A fold or scan of
// a tree runs the library's kernels with the tree as their element source:
// its bits must equal those of nt2h_fold/nt2h_scan on the materialised
// table (same launch plan, same order), which is itself checked against the
// oracle (maximum/minimum/cummax/cummin bit for bit, sums to a bound).
struct Ext4 {
  int64_t d[4];
  int64_t n() const { return d[0] * d[1] * d[2] * d[3]; }
};
inline void lmh_of(const Ext4& e, int dim, int64_t& lo, int64_t& mid, int64_t& hi) {
  lo = mid = hi = 1;
  if (dim > 4) { lo = e.n(); return; }
  for (int k = 0; k < dim - 1; ++k) lo *= e.d[k];
  mid = e.d[dim - 1];
  for (int k = dim; k < 4; ++k) hi *= e.d[k];
}
// index, in a table of extent own, of the element that linear index i of
// extent sh reads through sx(): coordinate 0 where own has extent 1
inline int64_t expanded_index(int64_t i, const Ext4& sh, const Ext4& own) {
  int64_t c[4];
  for (int k = 0; k < 4; ++k) { c[k] = i % sh.d[k]; i /= sh.d[k]; }
  int64_t j = 0;
  for (int k = 3; k >= 0; --k) j = j * own.d[k] + (own.d[k] == 1 ? 0 : c[k]);
  return j;
}
template <class T> bool same_bits_t(const std::vector<T>& a, const std::vector<T>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (std::memcmp(&a[i], &b[i], sizeof(T)) != 0 && !(std::isnan(a[i]) && std::isnan(b[i]))) return false;
  return true;
}
template <class X> table<typename nt2hip::value_of<X>::type> fold_named(int op, const X& x, int dim) {
  switch (op) {
    case NT2H_FOLD_SUM: return nt2hip::sum(x, dim);
    case NT2H_FOLD_PROD: return nt2hip::prod(x, dim);
    case NT2H_FOLD_MAXIMUM: return nt2hip::maximum(x, dim);
    case NT2H_FOLD_MINIMUM: return nt2hip::minimum(x, dim);
    case NT2H_FOLD_ASUM1: return nt2hip::asum1(x, dim);
    case NT2H_FOLD_ASUM2: return nt2hip::asum2(x, dim);
    case NT2H_FOLD_NORM2: return nt2hip::norm2(x, dim);
    default: return nt2hip::mean(x, dim);
  }
}
template <class X> table<typename nt2hip::value_of<X>::type> scan_named(int op, const X& x, int dim) {
  switch (op) {
    case NT2H_SCAN_CUMSUM: return nt2hip::cumsum(x, dim);
    case NT2H_SCAN_CUMPROD: return nt2hip::cumprod(x, dim);
    case NT2H_SCAN_CUMMAX: return nt2hip::cummax(x, dim);
    default: return nt2hip::cummin(x, dim);
  }
}
inline int lib_fold(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* o) {
  return nt2h_fold_f32(op, x, lo, mid, hi, o, nullptr);
}
inline int lib_fold(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* o) {
  return nt2h_fold_f64(op, x, lo, mid, hi, o, nullptr);
}
inline int lib_scan(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* o) {
  return nt2h_scan_f32(op, x, lo, mid, hi, o, nullptr);
}
inline int lib_scan(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* o) {
  return nt2h_scan_f64(op, x, lo, mid, hi, o, nullptr);
}
inline void ref_fold(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* o) {
  nt2ref_fold_lmh_f32(op, x, lo, mid, hi, o);
}
inline void ref_fold(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* o) {
  nt2ref_fold_lmh_f64(op, x, lo, mid, hi, o);
}
inline void ref_scan(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* o) {
  nt2ref_scan_lmh_f32(op, x, lo, mid, hi, o);
}
inline void ref_scan(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* o) {
  nt2ref_scan_lmh_f64(op, x, lo, mid, hi, o);
}
// Every fold (dims 1..5) and scan (dims 1..4) of the tree e, whose values on
// the host are want (already checked bit for bit against Z = e).
template <class T, class EX>
void reduce_checks(const char* name, const EX& e, const std::vector<T>& want, const Ext4& sh) {
  table<T> Z;
  Z = e;
  bool ident = true, oracle = true;
  char buf[256] = "";
  for (int dim = 1; dim <= 5; ++dim) {
    int64_t lo, mid, hi;
    lmh_of(sh, dim, lo, mid, hi);
    for (int op = NT2H_FOLD_SUM; op <= NT2H_FOLD_MEAN; ++op) {
      table<T> got = fold_named(op, e, dim);
      table<T> mat(lo * hi);
      check(lib_fold(op, Z.data(), lo, mid, hi, mat.data()), "nt2h_fold");
      const bool shape_ok = got.numel() == lo * hi && (dim > 4 || got.dim(dim) == 1);
      const auto g = got.to_host(), m = mat.to_host();
      if (!shape_ok || !same_bits_t(g, m)) {
        if (ident) std::snprintf(buf, sizeof buf, "fold op %d dim %d differs from the materialised fold", op, dim);
        ident = false;
      }
      bool ok = true;
      if (op == NT2H_FOLD_MAXIMUM || op == NT2H_FOLD_MINIMUM) {  // NT2's fold, bit for bit
        std::vector<T> r(lo * hi);
        ref_fold(op, want.data(), lo, mid, hi, r.data());
        ok = same_bits_t(g, r);
      } else if (op == NT2H_FOLD_SUM || op == NT2H_FOLD_MEAN || op == NT2H_FOLD_ASUM1) {
        // the compensated sums against an extended-precision sum: within two
        // units in the last place of the result (plus 2^-10 of one of the
        // magnitude's, for cancellation)
        for (int64_t o = 0; o < lo * hi && ok; ++o) {
          const int64_t h = o / lo, i = o - h * lo;
          long double s = 0, mag = 0;
          for (int64_t k = 0; k < mid; ++k) {
            const long double x = want[i + lo * (k + mid * h)];
            s += op == NT2H_FOLD_ASUM1 ? std::fabs(x) : x;
            mag += std::fabs(x);
          }
          if (op == NT2H_FOLD_MEAN) { s /= mid; mag /= mid; }
          if (std::isnan((double)s)) { ok = std::isnan(g[o]); continue; }
          const double ulp = std::is_same<T, float>::value ? 0x1p-23 : 0x1p-52;
          const double r = (double)s;
          ok = std::fabs((double)g[o] - r) <= 2 * ulp * std::fabs(r) + ulp * 0x1p-10 * (double)mag;
        }
      }
      if (!ok) {
        if (oracle) std::snprintf(buf, sizeof buf, "fold op %d dim %d differs from the reference", op, dim);
        oracle = false;
      }
    }
  }
  for (int dim = 1; dim <= 4; ++dim) {
    int64_t lo, mid, hi;
    lmh_of(sh, dim, lo, mid, hi);
    for (int op = NT2H_SCAN_CUMSUM; op <= NT2H_SCAN_CUMMIN; ++op) {
      table<T> got = scan_named(op, e, dim);
      table<T> mat(sh.n());
      check(lib_scan(op, Z.data(), lo, mid, hi, mat.data()), "nt2h_scan");
      const auto g = got.to_host(), m = mat.to_host();
      if (got.numel() != sh.n() || got.dim(dim) != sh.d[dim - 1] || !same_bits_t(g, m)) {
        if (ident) std::snprintf(buf, sizeof buf, "scan op %d dim %d differs from the materialised scan", op, dim);
        ident = false;
      }
      if (op == NT2H_SCAN_CUMMAX || op == NT2H_SCAN_CUMMIN) {
        std::vector<T> r(sh.n());
        ref_scan(op, want.data(), lo, mid, hi, r.data());
        if (!same_bits_t(g, r)) {
          if (oracle) std::snprintf(buf, sizeof buf, "scan op %d dim %d differs from the oracle (cummax/cummin)", op, dim);
          oracle = false;
        }
      }
    }
  }
  report(name, ident && oracle, buf);
}
