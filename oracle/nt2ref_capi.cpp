// C entry points of the oracle for ctypes — TEST INFRASTRUCTURE ONLY.
//
// Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
// never by the product library.  Every function here evaluates NT2's scalar
// path (nt2ref.hpp) element by element, or an FP64 closed form, on host
// arrays.  Loops are spread over OpenMP threads only where each element is
// independent, so results do not depend on the thread count; folds are
// sequential, left to right, as NT2's scalar inner_fold is.
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nt2ref.hpp"

using namespace nt2ref;

// op ids shared with include/nt2hip.h (NT2H_OP_*)
enum { OP_EXP = 0, OP_LOG = 1, OP_SQRT = 2, OP_ERF = 3, OP_ERFC = 4, OP_NORMCDF = 5, OP_TANH = 6,
       OP_SIN = 7, OP_COS = 8, OP_NORMPDF = 9, OP_FASTNORMCDF = 10, OP_EXP2 = 11, OP_EXP10 = 12 };

template <class T> static inline T apply(int op, T x);
template <> inline float apply<float>(int op, float x) {
  switch (op) {
    case OP_EXP: return nt2ref::exp(x);
    case OP_LOG: return nt2ref::log(x);
    case OP_SQRT: return nt2ref::sqrt(x);
    case OP_ERF: return nt2ref::erf(x);
    case OP_ERFC: return nt2ref::erfc(x);
    case OP_NORMCDF: return nt2ref::normcdf(x);
    case OP_TANH: return nt2ref::tanh(x);
    case OP_SIN: return nt2ref::sin(x);
    case OP_COS: return nt2ref::cos(x);
    case OP_NORMPDF: return nt2ref::normpdf(x);
    case OP_FASTNORMCDF: return nt2ref::fastnormcdf(x);
    case OP_EXP2: return nt2ref::exp2(x);
    case OP_EXP10: return nt2ref::exp10(x);
  }
  return qnan<float>();
}
template <> inline double apply<double>(int op, double x) {
  switch (op) {
    case OP_EXP: return nt2ref::exp(x);
    case OP_LOG: return nt2ref::log(x);
    case OP_SQRT: return nt2ref::sqrt(x);
    case OP_ERF: return nt2ref::erf(x);
    case OP_ERFC: return nt2ref::erfc(x);
    case OP_NORMCDF: return nt2ref::normcdf(x);
    case OP_TANH: return nt2ref::tanh(x);
    case OP_SIN: return nt2ref::sin(x);
    case OP_COS: return nt2ref::cos(x);
    case OP_NORMPDF: return nt2ref::normpdf(x);
  }
  return qnan<double>();
}

// Black-Scholes, config formula (not the demo's drift-only d1), composed
// from NT2 functors as a user writes it on tables:
//   d1 = (log(S/K) + (r + sqr(sig)/2)*T) / (sig*sqrt(T)); d2 = d1 - sig*sqrt(T)
//   call = S*normcdf(d1) - K*exp(-r*T)*normcdf(d2)
//   put  = K*exp(-r*T)*normcdf(-d2) - S*normcdf(-d1)
template <class T> struct BSFwd {
  T sqT, sst, d1, d2, ert, kert, n1, n2;
  void eval(T S, T K, T Tm, T r, T sig) {
    sqT = nt2ref::sqrt(Tm);
    sst = sig * sqT;
    d1 = (nt2ref::log(S / K) + (r + (sig * sig) * T(0.5)) * Tm) / sst;
    d2 = d1 - sst;
    ert = nt2ref::exp(-r * Tm);
    kert = K * ert;
    n1 = nt2ref::normcdf(d1);
    n2 = nt2ref::normcdf(d2);
  }
};

extern "C" {

int nt2ref_version(void) { return 1; }
int nt2ref_max_threads(void) { return omp_get_max_threads(); }

void nt2ref_map_f32(int op, const float* x, float* y, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) y[i] = apply<float>(op, x[i]);
}
void nt2ref_map_f64(int op, const double* x, double* y, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) y[i] = apply<double>(op, x[i]);
}

// demo/saxpy/nt2/saxpy_nt2.cpp:41 — Y = alpha*X + Y, two roundings
void nt2ref_saxpy_f32(float a, const float* x, float* y, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) { float p = a * x[i]; y[i] = p + y[i]; }
}

// NT2 scalar inner_fold order (fold_step.hpp:39-40): sequential float.
float nt2ref_fold_sum_f32(const float* x, int64_t n) { return fold_sum(x, n); }
void nt2ref_fold_lmh_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* out) {
  fold_lmh(op, x, lo, mid, hi, out);
}
void nt2ref_fold_lmh_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* out) {
  fold_lmh(op, x, lo, mid, hi, out);
}
void nt2ref_scan_lmh_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* y) {
  scan_lmh(op, x, lo, mid, hi, y);
}
void nt2ref_scan_lmh_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* y) {
  scan_lmh(op, x, lo, mid, hi, y);
}
float nt2ref_fold_dot_f32(const float* x, const float* y, int64_t n) { return fold_dot(x, y, n); }
double nt2ref_fold_sum_f64(const double* x, int64_t n) { return fold_sum(x, n); }
double nt2ref_fold_dot_f64(const double* x, const double* y, int64_t n) { return fold_dot(x, y, n); }

// Near-exact sums for the "correctly rounded" reference: double-double
// accumulation (TwoSum) of exact double terms (a float product is exact in
// double).  hi+lo is returned in out[0], out[1].
static inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b; double bb = s - a; e = (a - (s - bb)) + (b - bb);
}
void nt2ref_exact_sum_f32(const float* x, int64_t n, double* out) {
  double hi = 0, lo = 0;
  for (int64_t i = 0; i < n; ++i) { double s, e; two_sum(hi, (double)x[i], s, e); hi = s; lo += e; }
  double s, e; two_sum(hi, lo, s, e); out[0] = s; out[1] = e;
}
void nt2ref_exact_dot_f32(const float* x, const float* y, int64_t n, double* out) {
  double hi = 0, lo = 0;
  for (int64_t i = 0; i < n; ++i) {
    double p = (double)x[i] * (double)y[i];
    double s, e; two_sum(hi, p, s, e); hi = s; lo += e;
  }
  double s, e; two_sum(hi, lo, s, e); out[0] = s; out[1] = e;
}

// ---------------------------------------------------------- Black-Scholes
void nt2ref_bs_fwd_f32(const float* S, const float* K, const float* Tm, const float* r, const float* sig,
                       float* call, float* put, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    BSFwd<float> b; b.eval(S[i], K[i], Tm[i], r[i], sig[i]);
    call[i] = S[i] * b.n1 - b.kert * b.n2;
    put[i] = b.kert * nt2ref::normcdf(-b.d2) - S[i] * nt2ref::normcdf(-b.d1);
  }
}
void nt2ref_bs_fwd_f64(const double* S, const double* K, const double* Tm, const double* r, const double* sig,
                       double* call, double* put, int64_t n) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    BSFwd<double> b; b.eval(S[i], K[i], Tm[i], r[i], sig[i]);
    call[i] = S[i] * b.n1 - b.kert * b.n2;
    put[i] = b.kert * nt2ref::normcdf(-b.d2) - S[i] * nt2ref::normcdf(-b.d1);
  }
}

}  // extern "C"

// Forward + w-weighted analytic adjoint of price = call, in type T.
// dS = w N(d1); dK = -w e^{-rT} N(d2); dT = w (S phi(d1) sig/(2 sqrt T) + r K e^{-rT} N(d2));
// dr = w K T e^{-rT} N(d2); dsig = w S phi(d1) sqrt T.  Totals of dr and dsig
// are summed sequentially in double (the portfolio gradient for scalar r, sigma).
template <class T>
static void bs_fwd_adj(const T* S, const T* K, const T* Tm, const T* r, const T* sig, const T* w, T* price,
                       T* dS, T* dK, T* dT, T* dr, T* dsig, double* totals, int64_t n, bool parallel) {
#pragma omp parallel for schedule(static) if (parallel)
  for (int64_t i = 0; i < n; ++i) {
    BSFwd<T> b; b.eval(S[i], K[i], Tm[i], r[i], sig[i]);
    T phi = nt2ref::normpdf(b.d1);
    T kn2 = b.kert * b.n2;
    price[i] = S[i] * b.n1 - kn2;
    dS[i] = w[i] * b.n1;
    dK[i] = -(w[i] * (b.ert * b.n2));
    T sphi = S[i] * phi;
    dT[i] = w[i] * (sphi * sig[i] / (T(2) * b.sqT) + r[i] * kn2);
    dr[i] = w[i] * (Tm[i] * kn2);
    dsig[i] = w[i] * (sphi * b.sqT);
  }
  if (totals) {
    double tr = 0, ts = 0;
    for (int64_t i = 0; i < n; ++i) { tr += (double)dr[i]; ts += (double)dsig[i]; }
    totals[0] = tr; totals[1] = ts;
  }
}
extern "C" {
void nt2ref_bs_fwd_adj_f32(const float* S, const float* K, const float* Tm, const float* r, const float* sig,
                           const float* w, float* price, float* dS, float* dK, float* dT, float* dr, float* dsig,
                           double* totals, int64_t n) {
  bs_fwd_adj<float>(S, K, Tm, r, sig, w, price, dS, dK, dT, dr, dsig, totals, n, true);
}
void nt2ref_bs_fwd_adj_f64(const double* S, const double* K, const double* Tm, const double* r,
                           const double* sig, const double* w, double* price, double* dS, double* dK, double* dT,
                           double* dr, double* dsig, double* totals, int64_t n) {
  bs_fwd_adj<double>(S, K, Tm, r, sig, w, price, dS, dK, dT, dr, dsig, totals, n, true);
}

// CPU baseline: the float fwd+adj above under NT2's shared-memory policy
// (arch/shared_memory/include/nt2/core/functions/shared_memory/transform.hpp:39-49):
// grain = L2 bytes / sizeof(T); parallel only above 8 grains; static blocks.
void nt2ref_cpu_bs_fwd_adj_f32(const float* S, const float* K, const float* Tm, const float* r,
                               const float* sig, const float* w, float* price, float* dS, float* dK, float* dT,
                               float* dr, float* dsig, int64_t n, int64_t l2_bytes) {
  int64_t grain = l2_bytes / (int64_t)sizeof(float);
  if (grain < 1) grain = 1;
  if (n <= 8 * grain) {
    bs_fwd_adj<float>(S, K, Tm, r, sig, w, price, dS, dK, dT, dr, dsig, nullptr, n, false);
    return;
  }
  int64_t nblk = (n + grain - 1) / grain;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nblk; ++b) {
    int64_t o = b * grain, c = std::min(grain, n - o);
    bs_fwd_adj<float>(S + o, K + o, Tm + o, r + o, sig + o, w + o, price + o, dS + o, dK + o, dT + o, dr + o,
                      dsig + o, nullptr, c, false);
  }
}

// ------------------------------------------------ 2-D column fold (C5)
// z_j = sum_i exp(-sqr(X_ij)/(2*sqr(s_j))) * cos(k_j*X_ij + p_j), NT2 float
// functors, sequential float fold down each column (column-major, ld).
void nt2ref_colfold_f32(const float* X, int64_t rows, int64_t cols, int64_t ld, const float* s, const float* k,
                        const float* p, float* z) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < cols; ++j) {
    const float* c = X + j * ld;
    float ss2 = 2.0f * (s[j] * s[j]);
    float acc = 0.0f;
    for (int64_t i = 0; i < rows; ++i) {
      float x = c[i];
      float e = nt2ref::exp(-(x * x) / ss2);
      float t = e * nt2ref::cos(k[j] * x + p[j]);
      acc = acc + t;
    }
    z[j] = acc;
  }
}
// FP64 oracle of the same fold and its gradients (closed form):
//   dz/ds = sum e c X^2 / s^3,  dz/dk = -sum e sin(kX+p) X,  dz/dp = -sum e sin(kX+p)
// each scaled by the cotangent g_j (1 when g is null).  Sums are double-double.
void nt2ref_colfold_gc_f64(const float* X, int64_t rows, int64_t cols, int64_t ld, const float* s,
                           const float* k, const float* p, const float* g, double* z, double* ds, double* dk,
                           double* dp) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < cols; ++j) {
    const float* c = X + j * ld;
    double sj = s[j], kj = k[j], pj = p[j];
    double inv2s2 = 1.0 / (2.0 * sj * sj);
    double acc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int64_t i = 0; i < rows; ++i) {
      double x = c[i];
      double e = nt2ref::exp(-(x * x) * inv2s2);
      double a = kj * x + pj;
      double co = nt2ref::cos(a), si = nt2ref::sin(a);
      double t[4] = {e * co, e * co * x * x, -e * si * x, -e * si};
      for (int q = 0; q < 4; ++q) { double sm, er; two_sum(acc[q][0], t[q], sm, er); acc[q][0] = sm; acc[q][1] += er; }
    }
    double gj = g ? (double)g[j] : 1.0;
    double s3 = sj * sj * sj;
    z[j] = acc[0][0] + acc[0][1];
    ds[j] = gj * (acc[1][0] + acc[1][1]) / s3;
    dk[j] = gj * (acc[2][0] + acc[2][1]);
    dp[j] = gj * (acc[3][0] + acc[3][1]);
  }
}

// ------------------------------------------------------------ ULP stats
// got (float) against an FP64 reference: out = {max ulp_vs, argmax, max rel err, n_nonfinite_mismatch}
void nt2ref_ulp_stats_f32(const float* got, const double* ref, int64_t n, double* out) {
  double mu = 0, mr = 0; int64_t am = -1, bad = 0;
#pragma omp parallel
  {
    double lmu = 0, lmr = 0; int64_t lam = -1, lbad = 0;
#pragma omp for schedule(static) nowait
    for (int64_t i = 0; i < n; ++i) {
      double u = ulp_vs(got[i], ref[i]);
      if (!(u <= 1e300)) { lbad++; continue; }
      if (u > lmu || lam < 0) { lmu = u; lam = i; }
      double a = std::fabs(ref[i]);
      double rel = a > 0 ? std::fabs((double)got[i] - ref[i]) / a : std::fabs((double)got[i]);
      if (rel > lmr) lmr = rel;
    }
#pragma omp critical
    {
      if (lmu > mu || (lmu == mu && lam >= 0 && (am < 0 || lam < am))) { mu = lmu; am = lam; }
      if (lmr > mr) mr = lmr;
      bad += lbad;
    }
  }
  out[0] = mu; out[1] = (double)am; out[2] = mr; out[3] = (double)bad;
}
void nt2ref_ulp_stats_f64(const double* got, const double* ref, int64_t n, double* out) {
  double mu = 0; int64_t am = -1, bad = 0; double mr = 0;
  for (int64_t i = 0; i < n; ++i) {
    double u = std_ulp(got[i], ref[i]);
    if (!(u <= 1e300)) { bad++; continue; }
    if (u > mu || am < 0) { mu = u; am = i; }
    double a = std::fabs(ref[i]);
    double rel = a > 0 ? std::fabs(got[i] - ref[i]) / a : std::fabs(got[i]);
    if (rel > mr) mr = rel;
  }
  out[0] = mu; out[1] = (double)am; out[2] = mr; out[3] = (double)bad;
}
// Two float arrays: out = {max std-ULP, argmax, max NT2-ULP, n_nan_mismatch}
void nt2ref_cmp_f32(const float* a, const float* b, int64_t n, double* out) {
  double ms = 0, mn = 0; int64_t am = -1, bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    double s = std_ulp(a[i], b[i]);
    if (!(s <= 1e300)) { bad++; continue; }
    if (s > ms || am < 0) { ms = s; am = i; }
    double u = nt2_ulpdist(a[i], b[i]);
    if (u > mn) mn = u;
  }
  out[0] = ms; out[1] = (double)am; out[2] = mn; out[3] = (double)bad;
}
double nt2ref_nt2_ulpdist_f32(float a, float b) { return nt2_ulpdist(a, b); }
double nt2ref_nt2_ulpdist_f64(double a, double b) { return nt2_ulpdist(a, b); }

// ---------------------------------------- exhaustive ULP histogram (float)
// modules/test/exhaustive/include/nt2/sdk/exhaustive/exhaustive.hpp:103-269:
// every float in [lo, hi] once, bucket = min(128, ilog2(round(2*ulpdist+1))),
// bucket 129 = invalid distance; reference float(std::f(double(x))).
void nt2ref_exhaustive_hist_f32(int op, float lo, float hi, uint64_t* hist /*130*/) {
  for (int i = 0; i < 130; ++i) hist[i] = 0;
  int64_t olo = ord(lo), ohi = ord(hi);
#pragma omp parallel
  {
    uint64_t h[130] = {0};
#pragma omp for schedule(static)
    for (int64_t o = olo; o <= ohi; ++o) {
      int32_t bits = o < 0 ? (int32_t)((int64_t)(int32_t)0x80000000 - o) : (int32_t)o;
      float x = f_bits((uint32_t)bits);
      float got = apply<float>(op, x);
      double dx = x, rd;
      switch (op) {
        case OP_EXP: rd = std::exp(dx); break;
        case OP_EXP2: rd = std::exp2(dx); break;
        case OP_EXP10: rd = std::pow(10.0, dx); break;
        case OP_LOG: rd = std::log(dx); break;
        case OP_SQRT: rd = std::sqrt(dx); break;
        case OP_ERFC: rd = std::erfc(dx); break;
        case OP_ERF: rd = std::erf(dx); break;
        case OP_TANH: rd = std::tanh(dx); break;
        case OP_SIN: rd = std::sin(dx); break;
        case OP_COS: rd = std::cos(dx); break;
        default: rd = std::nan("");
      }
      float ref = (float)rd;
      double u = nt2_ulpdist(ref, got) * 2.0 + 1.0;
      if (!(std::fabs(u) <= 1e308)) { h[129]++; continue; }
      long iu = std::lround(u);
      int p = 0;
      while ((1L << (p + 1)) <= iu) ++p;
      if (p > 128) p = 128;
      h[p]++;
    }
#pragma omp critical
    for (int i = 0; i < 130; ++i) hist[i] += h[i];
  }
}

}  // extern "C"
