// nt2ref — CPU restatement of NT2's *scalar* evaluation path (the oracle).
//
// TEST INFRASTRUCTURE ONLY.  This file is the reference the HIP evaluator is
// compared against: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may build, load or call it.  The product library
// (nt2_amd/, include/nt2hip.h) never links or falls back to it.
//
// What it restates (paths relative to the NT2 checkout):
//   * two-rounding primitives fma/fnms and static Horner:
//       modules/boost/simd/base/include/boost/simd/arithmetic/functions/generic/fma.hpp:32
//       modules/boost/simd/base/include/boost/simd/arithmetic/functions/generic/fnms.hpp:32
//       modules/core/polynomials/include/nt2/polynomials/functions/scalar/impl/horner.hpp:37-62
//   * exp / exp2 / exp10:
//       modules/core/exponential/include/nt2/exponential/functions/scalar/impl/expo/
//         f_expo_reduction.hpp:49-157, d_expo_reduction.hpp:44-84,
//         expo_base.hpp:34-69, expo_scale.hpp:27-35
//   * log: .../scalar/impl/logs/f_log.hpp:59-87, f_kernel.hpp:43-77, d_log.hpp:52-112
//   * erf / erfc: modules/core/euler/include/nt2/euler/functions/scalar/erfc.hpp:39-105,
//       scalar/erf.hpp:39-100, details/erf_kernel.hpp:29-183
//   * normcdf / fastnormcdf / normpdf:
//       modules/core/statistics/include/nt2/statistics/functions/generic/normcdf.hpp:34,
//       fastnormcdf.hpp:39-54, normpdf.hpp:31
//   * tanh: modules/core/hyperbolic/include/nt2/hyperbolic/functions/scalar/tanh.hpp:45-54,
//       details/tanh_kernel.hpp:25-80
//   * sin / cos with the full reduction ladder:
//       modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/impl/trigo/
//         trig_base.hpp:131-161, trig_reduction.hpp:94-247, f_trig_evaluation.hpp:40-48,
//         d_trig_evaluation.hpp:25-60;
//       generic/rem_pio2_{straight,cephes,medium}.hpp; scalar/rem_pio2.hpp:83-913
//       (the fdlibm argument reduction, Sun Microsystems 1993, restated from its
//        published algorithm: Payne-Hanek with 8-bit float / 24-bit double chunks)
//   * sqrt: std::sqrt (modules/boost/simd/base/include/boost/simd/arithmetic/functions/scalar/sqrt.hpp:46)
//   * sequential fold: modules/core/reduction/include/nt2/core/functions/details/fold_step.hpp:31-43
//   * ULP metric: modules/test/unit/src/ulp.cpp:18-32
//
// Build with -ffp-contract=off and without -mfma: NT2's fma/fnms/horner are
// a*b+c with two roundings on x86 (no FMA3 override), so contraction would
// change results.  Every function below evaluates its operations in exactly
// the order and type the cited file does.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace nt2ref {

// ---------------------------------------------------------------- bit casts
inline float    f_bits(uint32_t u) { float f;  std::memcpy(&f, &u, 4); return f; }
inline double   d_bits(uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }
inline uint32_t bits_of(float f)   { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline uint64_t bits_of(double d)  { uint64_t u; std::memcpy(&u, &d, 8); return u; }

template <class T> struct ft;
template <> struct ft<float> {
  typedef uint32_t u; typedef int32_t i;
  static float from(uint32_t b) { return f_bits(b); }
  static const int mant = 23, bias = 127, biasm1 = 126;
  static uint32_t expmask() { return 0x7f800000u; }
  static uint32_t halfbits() { return 0x3f000000u; }
  static uint32_t signmask() { return 0x80000000u; }
};
template <> struct ft<double> {
  typedef uint64_t u; typedef int64_t i;
  static double from(uint64_t b) { return d_bits(b); }
  static const int mant = 52, bias = 1023, biasm1 = 1022;
  static uint64_t expmask() { return 0x7ff0000000000000ull; }
  static uint64_t halfbits() { return 0x3fe0000000000000ull; }
  static uint64_t signmask() { return 0x8000000000000000ull; }
};

template <class T> inline T inf() { return std::numeric_limits<T>::infinity(); }
template <class T> inline T qnan() { return std::numeric_limits<T>::quiet_NaN(); }

// ------------------------------------------------ two-rounding primitives
// fma(a,b,c) = a*b + c and fnms(a,b,c) = c - a*b, each operation rounded.
template <class T> inline T fma2(T a, T b, T c) { T p = a * b; return p + c; }
template <class T> inline T fnms2(T a, T b, T c) { T p = a * b; return c - p; }

// Static Horner: coefficients in decreasing powers (horner.hpp:37-62);
// horner(c0..cn)(x) = fma(x, fma(x, ... fma(x, c0, c1) ..., c_{n-1}), c_n).
template <class T, size_t N>
inline T horner(const typename ft<T>::u (&c)[N], T x) {
  T r = ft<T>::from(c[0]);
  for (size_t i = 1; i < N; ++i) r = fma2(x, r, ft<T>::from(c[i]));
  return r;
}

// round2even as Boost.SIMD writes it (arithmetic/functions/scalar/round2even.hpp):
// |a| + 2^nmb - 2^nmb below 2^nmb, sign restored.
template <class T> inline T round2even(T a0) {
  const T t2n = (sizeof(T) == 4) ? T(8388608.0) : T(4503599627370496.0);
  T v = std::fabs(a0);
  T d0 = v + t2n;
  T d = d0 - t2n;
  d = (v < t2n) ? d : v;
  return a0 < T(0) ? -d : d;
}

// x * 2^k with 2^k built from exponent bits (expo_scale.hpp:27-35).
inline float scale(float y, float k) {
  int32_t ik = (int32_t)k + 127;
  uint32_t b = (uint32_t)ik << 23;
  return y * f_bits(b);
}
inline double scale(double y, double k) {
  int64_t ik = (int64_t)k + 1023;
  uint64_t b = (uint64_t)ik << 52;
  return y * d_bits(b);
}

// fast_frexp (boost/simd/ieee/functions/scalar/fast_frexp.hpp): mantissa in
// [0.5,1) by forcing the exponent field, exponent = field - (bias-1).
template <class T> inline T fast_frexp(T a0, typename ft<T>::i& e) {
  typedef typename ft<T>::u U;
  U b = bits_of(a0);
  e = (typename ft<T>::i)((b & ft<T>::expmask()) >> ft<T>::mant) - ft<T>::biasm1;
  U m = (b & ~ft<T>::expmask()) | ft<T>::halfbits();
  T r; std::memcpy(&r, &m, sizeof(T));
  return r;
}

// ================================================================== exp
namespace cst {
// exponential/constants/*.hpp
static const uint32_t f_maxlog = 0x42b0c0a5u, f_minlog = 0xc2b0c0a5u;
static const uint32_t f_log2hi = 0x3f318000u, f_log2lo = 0xb95e8083u;
static const uint32_t f_invlog2 = 0x3fb8aa3bu, f_log2 = 0x3f317218u;
static const uint32_t f_maxlog2 = 0x42fe0000u, f_minlog2 = 0xc2fe0000u;
static const uint32_t f_maxlog10 = 0x4218ec59u, f_minlog10 = 0xc2179999u;
static const uint32_t f_invlog10_2 = 0x40549a78u, f_log10_2hi = 0x3e9a0000u, f_log10_2lo = 0x39826a14u;
static const uint64_t d_maxlog = 0x40862e42fefa39efull, d_minlog = 0xc086232bdd7abcd2ull;
static const uint64_t d_log2hi = 0x3fe62e42fee00000ull, d_log2lo = 0x3dea39ef35793c76ull;
static const uint64_t d_invlog2 = 0x3ff71547652b82feull;
}

// f_expo_reduction.hpp:59-67 — remez deg 4 for (exp(x)-1-x)/x^2 on [-0.5,0.5]
static const uint32_t kExpF[5] = {0x3ab778cf, 0x3c098d8b, 0x3d2aa957, 0x3e2aa9a5, 0x3f000000};
// f_expo_reduction.hpp:97-108 — exp2
static const uint32_t kExp2F[5] = {0x3922c8c4, 0x3ab01218, 0x3c1d9422, 0x3d6356eb, 0x3e75fdf1};
// f_expo_reduction.hpp:145-157 — exp10
static const uint32_t kExp10F[6] = {0x3e54dff1, 0x3f0aacef, 0x3f95eb4c, 0x400237da, 0x4029a926, 0x40135d8e};
// d_expo_reduction.hpp:69-75
static const uint64_t kExpD[5] = {0x3e66376972bea4d0ull, 0xbebbbd41c5d26bf1ull, 0x3f11566aaf25de2cull,
                                  0xbf66c16c16bebd93ull, 0x3fc555555555553eull};

// expo_base.hpp:54-69 (float driver) + f_expo_reduction natural_tag
inline float exp(float a0) {
  if (a0 >= f_bits(cst::f_maxlog)) return inf<float>();
  if (a0 <= f_bits(cst::f_minlog)) return 0.0f;
  if (a0 != a0) return a0;
  float k = round2even(f_bits(cst::f_invlog2) * a0);
  float x = fnms2(k, f_bits(cst::f_log2hi), a0);
  x = fnms2(k, f_bits(cst::f_log2lo), x);
  float y = horner(kExpF, x);
  float c = 1.0f + fma2(y, x * x, x);
  return scale(c, k);
}
inline float exp2(float a0) {
  if (a0 >= f_bits(cst::f_maxlog2)) return inf<float>();
  if (a0 <= f_bits(cst::f_minlog2)) return 0.0f;
  if (a0 != a0) return a0;
  float k = round2even(a0);
  float x = a0 - k;
  float y = horner(kExp2F, x);
  float c = 1.0f + fma2(y, x * x, x * f_bits(cst::f_log2));
  return scale(c, k);
}
inline float exp10(float a0) {
  if (a0 >= f_bits(cst::f_maxlog10)) return inf<float>();
  if (a0 <= f_bits(cst::f_minlog10)) return 0.0f;
  if (a0 != a0) return a0;
  float k = round2even(f_bits(cst::f_invlog10_2) * a0);
  float x = fnms2(k, f_bits(cst::f_log10_2hi), a0);
  x -= k * f_bits(cst::f_log10_2lo);
  float c = 1.0f + horner(kExp10F, x) * x;
  return scale(c, k);
}
// expo_base.hpp:34-50 (double driver) + d_expo_reduction.hpp:44-84
inline double exp(double a0) {
  if (a0 >= d_bits(cst::d_maxlog)) return inf<double>();
  if (a0 <= d_bits(cst::d_minlog)) return 0.0;
  if (a0 != a0) return a0;
  double k = round2even(d_bits(cst::d_invlog2) * a0);
  double hi = fnms2(k, d_bits(cst::d_log2hi), a0);
  double lo = k * d_bits(cst::d_log2lo);
  double x = hi - lo;
  double t = x * x;
  double c = fnms2(t, horner(kExpD, t), x);
  c = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
  return scale(c, k);
}

// ================================================================== log
// f_kernel.hpp:66-76
static const uint32_t kLogF[8] = {0xbda5dff0, 0x3e0229f9, 0xbe04d6b7, 0x3e116e80,
                                  0xbe2a6aa0, 0x3e4cd0a3, 0xbe800064, 0x3eaaaaa9};
// d_log.hpp:75-86
static const uint64_t kLogD1[3] = {0x3fc39a09d078c69full, 0x3fcc71c51d8e78afull, 0x3fd999999997fa04ull};
static const uint64_t kLogD2[4] = {0x3fc2f112df3e5244ull, 0x3fc7466496cb03deull, 0x3fd2492494229359ull,
                                   0x3fe5555555555593ull};

// f_log.hpp:59-87 + f_kernel.hpp:43-77
inline float log(float a0) {
  if (a0 == inf<float>()) return a0;
  if (a0 == 0.0f) return -inf<float>();
  if (a0 != a0 || a0 < 0.0f) return qnan<float>();
  float z = a0, t = 0.0f;
  if (std::fabs(z) < f_bits(0x00800000u)) {  // Smallestposval
    z *= f_bits(0x4b000000u);                 // Twotonmb = 2^23
    t = f_bits(0xc17f1402u);                  // Mlogtwo2nmb = -23 ln 2
  }
  int32_t e;
  float x = fast_frexp(z, e);
  bool lt = x < f_bits(0x3f3504f3u);  // Sqrt_2o_2
  float fe = lt ? (float)e + (-1.0f) : (float)e;
  x = (lt ? x + x : x) - 1.0f;
  float x2 = x * x;
  float y = horner(kLogF, x) * x * x2;
  y = fma2(fe, f_bits(cst::f_log2lo), y);
  y = fma2(-0.5f, x2, y);
  return fma2(f_bits(cst::f_log2hi), fe, x + y + t);
}
// d_log.hpp:52-112 (no subnormal prescale in the double path)
inline double log(double a0) {
  if (a0 == inf<double>()) return a0;
  if (a0 == 0.0) return -inf<double>();
  if (a0 != a0 || a0 < 0.0) return qnan<double>();
  int64_t k;
  double x = fast_frexp(a0, k);
  bool lt = d_bits(0x3FE6A09E667F3BCDull) > x;
  if (lt) k -= 1;
  double f = (lt ? x + x : x + 0.0) - 1.0;
  double dk = (double)k;
  double s = f / (2.0 + f);
  double z = s * s;
  double w = z * z;
  double t1 = w * horner(kLogD1, w);
  double t2 = z * horner(kLogD2, w);
  double R = t2 + t1;
  double hfsq = 0.5 * (f * f);
  return dk * d_bits(0x3fe62e42fee00000ull) -
         ((hfsq - (s * (hfsq + R) + dk * d_bits(0x3dea39ef35793c76ull))) - f);
}

inline float sqrt(float x) { return std::sqrt(x); }
inline double sqrt(double x) { return std::sqrt(x); }

// ============================================================= erf / erfc
// erf_kernel.hpp (float)
static const uint32_t kErf1F[6] = {0xba1fc83b, 0x3ba4468d, 0xbcdb61f4, 0x3de70f22, 0xbec0937e, 0x3f906eba};
static const uint32_t kErfc2F[11] = {0xc0283611, 0x40d67e3b, 0xc0cd1a85, 0x4050b063, 0xbfaea865, 0x3e2006f0,
                                     0x3e1175c7, 0x3ec4ca6e, 0x3e243828, 0xbf918a62, 0x3f0a0e8b};
static const uint32_t kErfc3F[9] = {0x41aa8e55, 0x401b5680, 0xc010d956, 0x3f2cff3b, 0xc016c985,
                                    0xbffc9284, 0xbfa11698, 0xbe036d7e, 0x3f7ffffe};
// erf_kernel.hpp (double)
static const uint64_t kErf1DP[5] = {0x3f110512d5b20332ull, 0x3f53b7664358865aull, 0x3fa4a59a4f02579cull,
                                    0x3fc16500f106c0a5ull, 0x3ff20dd750429b61ull};
static const uint64_t kErf1DQ[5] = {0x3f37ea4332348252ull, 0x3f8166f75999dbd1ull, 0x3fb64536ca92ea2full,
                                    0x3fdd0a84eb1ca867ull, 0x3ff0000000000000ull};
static const uint64_t kErfc2DP[7] = {0x0ull, 0x3f7cf4cfe0aacbb4ull, 0x3fb2488a6b5cb5e5ull, 0x3fd53dd7a67c7e9full,
                                     0x3fec1986509e687bull, 0x3ff54dfe9b258a60ull, 0x3feffffffbbb552bull};
static const uint64_t kErfc2DQ[7] = {0x3f89a996639b0d00ull, 0x3fc033c113a7deeeull, 0x3fe307622fcff772ull,
                                     0x3ff9e677c2777c3cull, 0x40053b1052dca8bdull, 0x4003adeae79b9708ull,
                                     0x3ff0000000000000ull};
static const uint64_t kErfc3DP[7] = {0x0ull, 0x3f971d0907ea7a92ull, 0x3fc42210f88b9d43ull, 0x3fe29be1cff90d94ull,
                                     0x3ff44744306832aeull, 0x3ff9fa202deb88e5ull, 0x3fefff5a9e697ae2ull};
static const uint64_t kErfc3DQ[7] = {0x3fa47bd61bbb3843ull, 0x3fd1d7ab774bb837ull, 0x3ff0cfd4cb6cde9full,
                                     0x400315ffdfd5ce91ull, 0x400afd487397568full, 0x400602f24bf3fdb6ull,
                                     0x3ff0000000000000ull};
static const uint64_t kErfc4D[14] = {0xc0505fce04ec4ec5ull, 0x4055803d26c4ec4full, 0xc0457e03041e9d8bull,
                                     0x4015ac7631f7ac4full, 0x40074fbabc514212ull, 0x3fb505a857e9ccc8ull,
                                     0xbff105e6b8ef1a63ull, 0x3f34a451701654f1ull, 0x3fdb1494cac06d39ull,
                                     0x3e907e71e046a820ull, 0xbfd20dd753ae5dfdull, 0x3db60000e984b501ull,
                                     0x3fe20dd750429a16ull, 0xbc7e4ad1ec7d0000ull};

inline float erfc3_f(float z) { return (1.0f - z) * horner(kErfc3F, z); }
inline float erfc2_f(float z) { return horner(kErfc2F, z); }
inline float erf1_f(float xx) { return horner(kErf1F, xx); }

static const uint32_t kTwothirdF = 0x3f2aaaabu;
static const uint64_t kPiD = 0x400921fb54442d18ull;

// scalar/erfc.hpp:77-105 (float)
inline float erfc(float a0) {
  float x = std::fabs(a0);
  float r1 = 0.0f;
  float z = x / (1.0f + x);
  if (x < f_bits(kTwothirdF)) {
    r1 = erfc3_f(z);
  } else if (x == inf<float>()) {
    r1 = 0.0f;
  } else {
    z -= 0.4f;
    r1 = exp(-(x * x)) * erfc2_f(z);
  }
  return (a0 < 0.0f) ? 2.0f - r1 : r1;
}
// scalar/erf.hpp:76-100 (float)
inline float erf(float a0) {
  if (a0 != a0) return a0;
  if (std::isinf(a0)) return a0 > 0 ? 1.0f : -1.0f;
  float x = std::fabs(a0);
  if (x < f_bits(kTwothirdF)) return a0 * erf1_f(x * x);
  float z = x / (1.0f + x) - 0.4f;
  float r2 = 1.0f - exp(-(x * x)) * erfc2_f(z);
  if (a0 < 0.0f) r2 = -r2;
  return r2;
}
// scalar/erfc.hpp:39-74 (double)
inline double erfc(double a0) {
  if (a0 != a0) return a0;
  double x = std::fabs(a0);
  double xx = x * x;
  double z = 0.0;
  if (x <= 0.0000000001) {
    z = 1.0 - (x * 2.0) / std::sqrt(d_bits(kPiD));
  } else if (x < 0.65) {
    z = 1.0 - x * (horner(kErf1DP, xx) / horner(kErf1DQ, xx));
  } else if (x < 2.2) {
    z = exp(-xx) * (horner(kErfc2DP, x) / horner(kErfc2DQ, x));
  } else if (x < 6.0) {
    z = exp(-xx) * (horner(kErfc3DP, x) / horner(kErfc3DQ, x));
  } else {
    z = exp(-xx) * horner(kErfc4D, 1.0 / x);
  }
  return (a0 < 0.0) ? 2.0 - z : z;
}
// scalar/erf.hpp:39-71 (double)
inline double erf(double a0) {
  if (a0 != a0) return a0;
  double x = std::fabs(a0);
  double xx = x * x;
  if (x <= 0.0000000001) return (a0 * 2.0) / std::sqrt(d_bits(kPiD));
  if (x <= 0.65) return a0 * (horner(kErf1DP, xx) / horner(kErf1DQ, xx));
  if (x <= 2.2) {
    double z = 1.0 - exp(-xx) * (horner(kErfc2DP, x) / horner(kErfc2DQ, x));
    return a0 < 0.0 ? -z : z;
  }
  if (x <= 6.0) {
    double z = 1.0 - exp(-xx) * (horner(kErfc3DP, x) / horner(kErfc3DQ, x));
    return a0 < 0.0 ? -z : z;
  }
  return 1.0;  // One<A0>() as the reference returns, whatever the sign
}

// ================================================= normcdf & friends
// generic/normcdf.hpp:34 — Half * erfc(-Sqrt_2o_2 * a0)
inline float normcdf(float a0) { return 0.5f * erfc(-f_bits(0x3F3504F3u) * a0); }
inline double normcdf(double a0) { return 0.5 * erfc(-d_bits(0x3FE6A09E667F3BCDull) * a0); }
// generic/normpdf.hpp:31 — exp(Mhalf*sqr(a0)) * Invsqrt_2pi
inline float normpdf(float a0) { return exp(-0.5f * (a0 * a0)) * f_bits(0x3ecc422au); }
inline double normpdf(double a0) { return exp(-0.5 * (a0 * a0)) * d_bits(0x3fd9884533d43651ull); }
// generic/fastnormcdf.hpp:39-54 (float only) — Abramowitz & Stegun 26.2.17
static const uint32_t kFastNcdf[5] = {0x3faa466f, 0xbfe91eea, 0x3fe40778, 0xbeb68f87, 0x3ea385fa};
inline float fastnormcdf(float a0) {
  float l = std::fabs(a0);
  float k = 1.0f / fma2(0.2316419f, l, 1.0f);
  float w = horner(kFastNcdf, k);
  w *= k * f_bits(0x3ecc422au) * exp(-(l * l) * 0.5f);
  return a0 > 0.0f ? 1.0f - w : w;
}

// ================================================================ tanh
static const uint32_t kTanhF[5] = {0xbbbaf0ea, 0x3ca9134e, 0xbd5c1e2d, 0x3e088393, 0xbeaaaa99};
static const uint64_t kTanhDP[3] = {0xbfeedc5baafd6f4bull, 0xc058d26a0e26682dull, 0xc0993ac030580563ull};
static const uint64_t kTanhDQ[4] = {0x3FF0000000000000ull, 0x405c33f28a581b86ull, 0x40a176fa0e5535faull,
                                    0x40b2ec102442040cull};
// scalar/tanh.hpp:45-54 + details/tanh_kernel.hpp
inline float tanh(float a0) {
  float x = std::fabs(a0);
  if (x < 0.625f) {
    float x2 = x * x;
    return fma2(horner(kTanhF, x2) * x2, a0, a0);
  }
  float r = fma2(-2.0f, 1.0f / (1.0f + exp(x + x)), 1.0f);
  return f_bits(bits_of(r) ^ (bits_of(a0) & 0x80000000u));
}
inline double tanh(double a0) {
  double x = std::fabs(a0);
  if (x < 0.625) {
    double x2 = x * x;
    return fma2(x2 * horner(kTanhDP, x2) / horner(kTanhDQ, x2), a0, a0);
  }
  double r = fma2(-2.0, 1.0 / (1.0 + exp(x + x)), 1.0);
  return d_bits(bits_of(r) ^ (bits_of(a0) & 0x8000000000000000ull));
}

// =========================================================== trigonometry
namespace trig {
// trigonometric/constants (float, double)
static const uint32_t f_pio4 = 0x3f490fdb, f_pio2 = 0x3fc90fdb, f_20pi = 0x427b53d1, f_mpi = 0x43490fdb,
                      f_dmpi = 0x49490fdb, f_twoopi = 0x3f22f983, f_invpio2 = 0x3f22f984;
static const uint32_t f_p1 = 0x3fc90f80, f_p2 = 0x37354400, f_p3 = 0x2e85a300;
static const uint32_t f_p1t = 0x37354443, f_p2t = 0x2e85a308, f_p3t = 0x248d3132;
static const uint64_t d_pio4 = 0x3fe921fb54442d18ull, d_pio2 = 0x3ff921fb54442d18ull,
                      d_20pi = 0x404f6a7a2955385eull, d_mpi = 0x412921fb54442d18ull,
                      d_twoopi = 0x3fe45f306dc9c883ull, d_invpio2 = 0x3FE45F306DC9C883ull;
static const uint64_t d_p1 = 0x3FF921FB54400000ull, d_p2 = 0x3DD0B4611A600000ull, d_p3 = 0x3BA3198A2E000000ull;
static const uint64_t d_p1t = 0x3DD0B4611A626331ull, d_p2t = 0x3BA3198A2E037073ull, d_p3t = 0x397B839A252049C1ull;

// 2/pi digits: 8-bit chunks for the float kernel, 24-bit chunks for double.
// These are the binary expansion of 2/pi (a mathematical constant).
extern const int32_t two_over_pi_24[66];
int32_t kernel_rem_pio2f(const float* x, float* y, int e0, int nx);
int32_t kernel_rem_pio2d(const double* x, double* y, int e0, int nx);
int32_t ieee754_rem_pio2f(float x, float* y);
int32_t ieee754_rem_pio2d(double x, double* y);

template <class T> struct C;
template <> struct C<float> {
  static float pio4() { return f_bits(f_pio4); } static float pio2() { return f_bits(f_pio2); }
  static float p20() { return f_bits(f_20pi); } static float mpi() { return f_bits(f_mpi); }
  static float twoopi() { return f_bits(f_twoopi); } static float invpio2() { return f_bits(f_invpio2); }
  static float p1() { return f_bits(f_p1); } static float p2() { return f_bits(f_p2); } static float p3() { return f_bits(f_p3); }
  static float p1t() { return f_bits(f_p1t); } static float p2t() { return f_bits(f_p2t); } static float p3t() { return f_bits(f_p3t); }
};
template <> struct C<double> {
  static double pio4() { return d_bits(d_pio4); } static double pio2() { return d_bits(d_pio2); }
  static double p20() { return d_bits(d_20pi); } static double mpi() { return d_bits(d_mpi); }
  static double twoopi() { return d_bits(d_twoopi); } static double invpio2() { return d_bits(d_invpio2); }
  static double p1() { return d_bits(d_p1); } static double p2() { return d_bits(d_p2); } static double p3() { return d_bits(d_p3); }
  static double p1t() { return d_bits(d_p1t); } static double p2t() { return d_bits(d_p2t); } static double p3t() { return d_bits(d_p3t); }
};

// generic/rem_pio2_cephes.hpp: 3-part Cody-Waite with xi = round2even(x*2/pi)
template <class T> inline int32_t rem_cephes(T x, T& xr) {
  T xi = round2even(x * C<T>::twoopi());
  xr = x - xi * C<T>::p1();
  xr -= xi * C<T>::p2();
  xr -= xi * C<T>::p3();
  return ((int32_t)(typename ft<T>::i)xi) & 3;
}
// generic/rem_pio2_medium.hpp: fdlibm-style medium reduction with tails
template <class T> inline int32_t rem_medium(T t, T& xr) {
  const T fn = round2even(t * C<T>::invpio2());
  T r = t - fn * C<T>::p1();
  T w = fn * C<T>::p1t();
  T t2 = r;
  w = fn * C<T>::p2();
  r = t2 - w;
  w = fn * C<T>::p2t() - ((t2 - r) - w);
  t2 = r;
  w = fn * C<T>::p3();
  r = t2 - w;
  w = fn * C<T>::p3t() - ((t2 - r) - w);
  xr = r - w;
  return ((int32_t)(typename ft<T>::i)fn) & 3;
}

// trig_reduction.hpp:94-247 — the scalar ladder for mode big_, x = |a0| finite.
inline int32_t reduce(double x, double& xr);
inline int32_t reduce(float x, float& xr) {
  if (!(x > C<float>::pio4())) { xr = x; return 0; }
  if (!(x > C<float>::pio2())) {  // rem_pio2_straight
    xr = x - C<float>::p1(); xr -= C<float>::p2(); xr -= C<float>::p3();
    return 1;
  }
  if (!(x > C<float>::p20())) return rem_cephes(x, xr);
  if (!(x > C<float>::mpi())) return rem_medium(x, xr);
  if (!(x > f_bits(f_dmpi))) {  // convert to double and reduce there
    double uxr; int32_t n = reduce((double)x, uxr);
    xr = (float)uxr;
    return n;
  }
  float y[2];
  int32_t n = ieee754_rem_pio2f(x, y);
  xr = y[0];
  return n & 3;
}
inline int32_t reduce(double x, double& xr) {
  if (!(x > C<double>::pio4())) { xr = x; return 0; }
  if (!(x > C<double>::pio2())) {
    xr = x - C<double>::p1(); xr -= C<double>::p2(); xr -= C<double>::p3();
    return 1;
  }
  if (!(x > C<double>::p20())) return rem_cephes(x, xr);
  if (!(x > C<double>::mpi())) return rem_medium(x, xr);
  // no upgrade exists for double: fdlibm rem_pio2 (scalar/rem_pio2.hpp:406-430)
  if (x == inf<double>()) { xr = qnan<double>(); return 0; }
  double y[2];
  int32_t n = ieee754_rem_pio2d(x, y);
  xr = y[0];
  return n & 3;
}

// f_trig_evaluation.hpp:40-48, d_trig_evaluation.hpp:25-50
static const uint32_t kCosF[3] = {0x37ccf5ce, 0xbab60619, 0x3d2aaaa5};
static const uint32_t kSinF[3] = {0xb94ca1f9, 0x3c08839d, 0xbe2aaaa2};
static const uint64_t kCosD[7] = {0x3da8ff831ad9b219ull, 0xbe21eea7c1e514d4ull, 0x3e927e4f8e06d9a5ull,
                                  0xbefa01a019ddbcd9ull, 0x3f56c16c16c15d47ull, 0xbfa5555555555551ull,
                                  0x3fe0000000000000ull};
static const uint64_t kSinD[6] = {0x3de5d8fd1fcf0ec1ull, 0xbe5ae5e5a9291691ull, 0x3ec71de3567d4896ull,
                                  0xbf2a01a019bfdf03ull, 0x3f8111111110f7d0ull, 0xbfc5555555555548ull};
inline float cos_eval(float z) {
  float y = horner(kCosF, z);
  return 1.0f + fma2(z, -0.5f, y * (z * z));
}
inline float sin_eval(float z, float x) { return fma2(horner(kSinF, z) * z, x, x); }
inline double cos_eval(double z) { return 1.0 - horner(kCosD, z) * z; }
inline double sin_eval(double z, double x) { return fma2(horner(kSinD, z) * z, x, x); }
}  // namespace trig

template <class T> inline bool is_invalid(T a) { return !(std::fabs(a) <= std::numeric_limits<T>::max()); }
template <class T> inline T xor_sign(T z, typename ft<T>::u s) {
  typename ft<T>::u b = bits_of(z) ^ s; T r; std::memcpy(&r, &b, sizeof(T)); return r;
}

// trig_base.hpp:131-145 (cosa, regular style)
template <class T> inline T cos_t(T a0) {
  typedef typename ft<T>::u U;
  if (is_invalid(a0)) return qnan<T>();
  T x = std::fabs(a0), xr;
  int32_t n = trig::reduce(x, xr);
  U swap = (U)(n & 1);
  U sign = (swap ^ (U)((n & 2) >> 1)) << (sizeof(T) * 8 - 1);
  T z = xr * xr;
  z = swap ? trig::sin_eval(z, xr) : trig::cos_eval(z);
  return xor_sign(z, sign);
}
// trig_base.hpp:147-161 (sina, regular style)
template <class T> inline T sin_t(T a0) {
  typedef typename ft<T>::u U;
  if (is_invalid(a0)) return qnan<T>();
  T x = std::fabs(a0), xr;
  int32_t n = trig::reduce(x, xr);
  int32_t swap = n & 1;
  U sign = (bits_of(a0) & ft<T>::signmask()) ^ ((U)(n & 2) << (sizeof(T) * 8 - 2));
  T z = xr * xr;
  z = swap ? trig::cos_eval(z) : trig::sin_eval(z, xr);
  return xor_sign(z, sign);
}
inline float sin(float a) { return sin_t(a); }
inline double sin(double a) { return sin_t(a); }
inline float cos(float a) { return cos_t(a); }
inline double cos(double a) { return cos_t(a); }

// ============================================================ ULP metrics
// modules/test/unit/src/ulp.cpp:18-32 — frexp-mantissa distance over epsilon
// (adjacent floats are 0.5 apart: "NT2-ULP").
template <class T> inline double nt2_ulpdist(T a0, T a1) {
  if ((a0 == a1) || ((a0 != a0) && (a1 != a1))) return 0.;
  int e1, e2;
  T m1 = std::frexp(a0, &e1);
  T m2 = std::frexp(a1, &e2);
  int expo = -(e1 > e2 ? e1 : e2);
  T e = (e1 == e2) ? std::fabs(m1 - m2) : std::fabs(std::ldexp(a0, expo) - std::ldexp(a1, expo));
  return double(e / std::numeric_limits<T>::epsilon());
}
// IEEE spacing distance (nextafter steps), NaN==NaN counts as 0.
inline int64_t ord(float f) { int32_t i = (int32_t)bits_of(f); return i < 0 ? (int64_t)(int32_t)0x80000000 - i : i; }
inline int64_t ord(double d) {
  int64_t i = (int64_t)bits_of(d);
  return i < 0 ? (int64_t)0x8000000000000000ull - i : i;
}
template <class T> inline double std_ulp(T a, T b) {
  if (a != a && b != b) return 0.0;
  if (a != a || b != b) return std::numeric_limits<double>::infinity();
  int64_t d = ord(a) - ord(b);
  return (double)(d < 0 ? -d : d);
}
// |got - ref| in units of the spacing of T at ref, with |ref| floored at the
// smallest normal (underflow regions are then measured absolutely).
template <class T> inline double ulp_vs(T got, double ref) {
  if (got != got && ref != ref) return 0.0;
  if (got != got || ref != ref) return std::numeric_limits<double>::infinity();
  if (std::isinf(ref) || std::isinf(got)) return ((double)got == ref) ? 0.0 : std::numeric_limits<double>::infinity();
  double a = std::fabs(ref);
  double mn = (double)std::numeric_limits<T>::min();
  if (a < mn) a = mn;
  int e; std::frexp(a, &e);
  double sp = std::ldexp(1.0, e - std::numeric_limits<T>::digits);
  return std::fabs((double)got - ref) / sp;
}

// ======================================================= sequential fold
// details/fold_step.hpp:31-43 + scalar/inner_fold.hpp:37-52: acc = acc + in(i), left to right.
template <class T> inline T fold_sum(const T* x, int64_t n) {
  T acc = T(0);
  for (int64_t i = 0; i < n; ++i) acc = acc + x[i];
  return acc;
}
template <class T> inline T fold_dot(const T* x, const T* y, int64_t n) {
  T acc = T(0);
  for (int64_t i = 0; i < n; ++i) { T p = x[i] * y[i]; acc = acc + p; }
  return acc;
}

// ============================================ other reduction_ functors
// Each is reduction_<tag, bop, neutral> folded left to right by fold_step:
//   sum_     plus        Zero     (boost/simd/reduction/functions/sum.hpp:26)
//   prod_    multiplies  One      (boost/simd/reduction/functions/prod.hpp:26)
//   maximum_ max_        Maxinit  (boost/simd/reduction/functions/maximum.hpp:30;
//            Maxinit = -FLT_MAX / -DBL_MAX, constant/constants/maxinit.hpp)
//   minimum_ min_        Valmax   (boost/simd/reduction/functions/minimum.hpp:29)
// with the scalar bops a0 > a1 ? a0 : a1 and a0 < a1 ? a0 : a1
// (arithmetic/functions/scalar/{max,min}.hpp:27).  asum1 = sum(abs(x)),
// asum2 = sum(sqr_abs(x)) (core/functions/expr/{asum1,asum2}.hpp:24),
// norm2 = sqrt(asum2(x)) (expr/norm2.hpp), mean = rec(T(size)) * sum(x)
// (expr/mean.hpp:22-23).
enum FoldOp { FOLD_SUM = 0, FOLD_PROD, FOLD_MAX, FOLD_MIN, FOLD_ASUM1, FOLD_ASUM2, FOLD_NORM2, FOLD_MEAN };
template <class T> inline T fold_neutral(int op) {
  switch (op) {
    case FOLD_PROD: return T(1);
    case FOLD_MAX: return -std::numeric_limits<T>::max();
    case FOLD_MIN: return std::numeric_limits<T>::max();
    default: return T(0);
  }
}
template <class T> inline T fold_pre(int op, T x) {
  switch (op) {
    case FOLD_ASUM1: return std::fabs(x);
    case FOLD_ASUM2: case FOLD_NORM2: { T a = std::fabs(x); return a * a; }
    default: return x;
  }
}
template <class T> inline T fold_bop(int op, T acc, T x) {
  switch (op) {
    case FOLD_PROD: return acc * x;
    case FOLD_MAX: return (acc > x) ? acc : x;
    case FOLD_MIN: return (acc < x) ? acc : x;
    default: return acc + x;
  }
}
// Fold over the middle index of a contiguous (lo, mid, hi) table: the
// reshaping NT2's reduction run_assign does before inner_fold (lo == 1) or
// outer_fold (dsl/functions/reduction/run.hpp:48-93);
// out[i + lo*h] = fold over m of x[i + lo*(m + mid*h)].
template <class T> inline void fold_lmh(int op, const T* x, int64_t lo, int64_t mid, int64_t hi, T* out) {
  for (int64_t h = 0; h < hi; ++h)
    for (int64_t i = 0; i < lo; ++i) {
      T acc = fold_neutral<T>(op);
      for (int64_t m = 0; m < mid; ++m) acc = fold_bop(op, acc, fold_pre(op, x[i + lo * (m + mid * h)]));
      if (op == FOLD_NORM2) acc = std::sqrt(acc);
      if (op == FOLD_MEAN) { T f = T(1) / T(mid); acc = f * acc; }
      out[i + lo * h] = acc;
    }
}
// Scans (cumulative_<tag, bop, neutral>, boost/simd/swar/functions/
// {cumsum,cumprod,cummax,cummin}.hpp: Zero, One, Minf, Inf) over the middle
// index, the scalar scan_step recurrence summary = bop(summary, in(i));
// out(i) = summary (core/functions/details/scan_step.hpp, scalar/
// {inner,outer}_scan.hpp).
enum ScanOp { SCAN_SUM = 0, SCAN_PROD, SCAN_MAX, SCAN_MIN };
template <class T> inline void scan_lmh(int op, const T* x, int64_t lo, int64_t mid, int64_t hi, T* y) {
  const T inf = std::numeric_limits<T>::infinity();
  for (int64_t h = 0; h < hi; ++h)
    for (int64_t i = 0; i < lo; ++i) {
      T acc = op == SCAN_PROD ? T(1) : op == SCAN_MAX ? -inf : op == SCAN_MIN ? inf : T(0);
      for (int64_t m = 0; m < mid; ++m) {
        const int64_t k = i + lo * (m + mid * h);
        T v = x[k];
        acc = op == SCAN_PROD ? acc * v : op == SCAN_MAX ? ((acc > v) ? acc : v)
            : op == SCAN_MIN ? ((acc < v) ? acc : v) : acc + v;
        y[k] = acc;
      }
    }
}

}  // namespace nt2ref
