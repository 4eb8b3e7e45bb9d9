// Device functors of the NT2 table-expression evaluator (gfx950).
//
// Each functor follows the algorithm and coefficient table of NT2's scalar
// functor (cited per function) so that GPU results stay within a few ULP of
// NT2's own evaluation; the arithmetic is chosen per call site by a policy:
//
//   Exact  — every operation rounded separately, IEEE division and square
//            root: NT2's scalar semantics (its fma/fnms/horner are a*b+c with
//            two roundings on x86).  Contraction is switched off from this
//            header on (pragma below, whatever the includer's -ffp-contract),
//            so a*b+c here really rounds twice and the result is
//            bit-identical to NT2's scalar path.
//   Fused  — v_fma_f32/f64 for every multiply-add and a refined hardware
//            reciprocal for float division: fewer VALU slots, same or better
//            accuracy, within a stated ULP bound of the Exact result.
//
// The large-argument trigonometric reduction is an integer Payne-Hanek over
// the bits of 2/pi; moderate arguments use a 3-part Cody-Waite in FP64 with
// FMA (exact first step), which is the only FP64 in the float functors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NT2H_FD __device__ __forceinline__
// host and device: the integer parts of the trigonometric reduction, which the
// exhaustive CPU checks run from this same source
#define NT2H_HD __host__ __device__ __forceinline__

// No contraction from here to the end of the including translation unit
// (clang has no push/pop for this on the AMDGPU target).
#pragma clang fp contract(off)

namespace nt2h {

// ------------------------------------------------------------ policies
struct Exact {
  static constexpr bool fused = false;
  NT2H_FD static float madd(float a, float b, float c) { return a * b + c; }
  NT2H_FD static double madd(double a, double b, double c) { return a * b + c; }
  NT2H_FD static float nms(float a, float b, float c) { return c - a * b; }
  NT2H_FD static double nms(double a, double b, double c) { return c - a * b; }
  NT2H_FD static float div(float a, float b) { return a / b; }
  NT2H_FD static double div(double a, double b) { return a / b; }
  NT2H_FD static float rcp(float a) { return 1.0f / a; }
  NT2H_FD static double rcp(double a) { return 1.0 / a; }
  NT2H_FD static float sqrt(float a) { return __builtin_sqrtf(a); }
  NT2H_FD static double sqrt(double a) { return __builtin_sqrt(a); }
};
struct Fused {
  static constexpr bool fused = true;
  NT2H_FD static float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  NT2H_FD static double madd(double a, double b, double c) { return __builtin_fma(a, b, c); }
  NT2H_FD static float nms(float a, float b, float c) { return __builtin_fmaf(-a, b, c); }
  NT2H_FD static double nms(double a, double b, double c) { return __builtin_fma(-a, b, c); }
  // a/b via v_rcp_f32 and one Newton correction of the quotient: <= 1 ulp
  // for b and a/b in the normal range (all call sites keep them there).
  // Where the correction is undefined (b or a/b infinite or zero) the
  // unrefined quotient is already exact and is returned instead.
  NT2H_FD static float div(float a, float b) {
    float r = __builtin_amdgcn_rcpf(b);
    float q = a * r;
    float e = __builtin_fmaf(-b, q, a);
    float q2 = __builtin_fmaf(e, r, q);
    return (q2 != q2) ? q : q2;
  }
  NT2H_FD static double div(double a, double b) { return a / b; }
  NT2H_FD static float rcp(float a) {
    float r = __builtin_amdgcn_rcpf(a);
    float e = __builtin_fmaf(-a, r, 1.0f);
    float r2 = __builtin_fmaf(e, r, r);
    return (r2 != r2) ? r : r2;
  }
  NT2H_FD static double rcp(double a) { return 1.0 / a; }
  NT2H_FD static float sqrt(float a) { return __builtin_sqrtf(a); }
  NT2H_FD static double sqrt(double a) { return __builtin_sqrt(a); }
};

// Static Horner, coefficients in decreasing powers:
// horner(x, c0, c1, ..., cn) = madd(x, ... madd(x, c0, c1) ..., cn)
template <class P, class T> NT2H_FD T horner(T, T c0) { return c0; }
template <class P, class T, class... R> NT2H_FD T horner(T x, T c0, T c1, R... rest) {
  return horner<P>(x, P::madd(x, c0, c1), T(rest)...);
}

NT2H_FD float fbits(uint32_t u) { return __uint_as_float(u); }
NT2H_FD uint32_t ubits(float f) { return __float_as_uint(f); }
NT2H_FD double dbits(uint64_t u) { return __longlong_as_double((long long)u); }
NT2H_FD uint64_t ubits(double d) { return (uint64_t)__double_as_longlong(d); }

constexpr float kInfF = __builtin_huge_valf();
constexpr double kInfD = __builtin_huge_val();

// ================================================================ exp
// NT2 float exp (f_expo_reduction.hpp:49-85, expo_base.hpp:54-69,
// expo_scale.hpp:27-35): k = round2even(x/ln2), 2-part Cody-Waite, degree-4
// minimax of (e^r-1-r)/r^2, 2^k built from exponent bits (so k <= -127 gives
// 0, never a subnormal, exactly as NT2).  round2even == rint in RN mode.
template <class P> NT2H_FD float exp_core(float a0) {
  float k = __builtin_rintf(a0 * 0x1.715476p+0f);
  float x = P::nms(k, 0x1.63p-1f, a0);
  x = P::nms(k, -0x1.bd0106p-13f, x);
  float y = horner<P>(x, 0x1.6ef19ep-10f, 0x1.131b16p-7f, 0x1.5552aep-5f, 0x1.55534ap-3f, 0x1.0p-1f);
  float c = 1.0f + P::madd(y, x * x, x);
  k = __builtin_fminf(__builtin_fmaxf(k, -127.0f), 127.0f);
  int32_t ik = ((int32_t)k + 127) << 23;
  return c * __int_as_float(ik);
}
template <class P> NT2H_FD float exp_f(float a0) {
  float r = exp_core<P>(a0);
  r = (a0 >= 0x1.61814ap+6f) ? kInfF : r;
  r = (a0 <= -0x1.61814ap+6f) ? 0.0f : r;
  return (a0 != a0) ? a0 : r;
}
// exp of an argument known to be finite and <= 0 (call sites: -r*T, -x^2),
// bit-identical to exp_core: k <= -127 (exactly when a0/ln2 rounds below
// -126.5) gives 0 as the exponent-bits scale does, every k in [-126, 0]
// scales exactly, so 2^k is applied with ldexp and one select.
template <class P> NT2H_FD float exp_neg_f(float a0) {
  float t = a0 * 0x1.715476p+0f;
  float k = __builtin_rintf(t);
  float x = P::nms(k, 0x1.63p-1f, a0);
  x = P::nms(k, -0x1.bd0106p-13f, x);
  float y = horner<P>(x, 0x1.6ef19ep-10f, 0x1.131b16p-7f, 0x1.5552aep-5f, 0x1.55534ap-3f, 0x1.0p-1f);
  float c = 1.0f + P::madd(y, x * x, x);
  return (t < -126.5f) ? 0.0f : __builtin_ldexpf(c, (int)k);
}
// exp2 / exp10 (f_expo_reduction.hpp:87-157)
template <class P> NT2H_FD float exp2_f(float a0) {
  float k = __builtin_rintf(a0);
  float x = a0 - k;
  float y = horner<P>(x, 0x1.45918800p-13f, 0x1.60243000p-10f, 0x1.3b284400p-7f, 0x1.c6add600p-5f,
                      0x1.ebfbe200p-3f);
  float c = 1.0f + P::madd(y, x * x, x * 0x1.62e430p-1f);
  k = __builtin_fminf(__builtin_fmaxf(k, -127.0f), 127.0f);
  float r = c * __int_as_float(((int32_t)k + 127) << 23);
  r = (a0 >= 127.0f) ? kInfF : r;
  r = (a0 <= -127.0f) ? 0.0f : r;
  return (a0 != a0) ? a0 : r;
}
template <class P> NT2H_FD float exp10_f(float a0) {
  float k = __builtin_rintf(0x1.a934f0p+1f * a0);
  float x = P::nms(k, 0x1.34p-2f, a0);
  x = x - k * 0x1.04d428p-12f;
  float c = 1.0f + horner<P>(x, 0x1.a9bfe2p-3f, 0x1.1559dep-1f, 0x1.2bd698p+0f, 0x1.046fb4p+1f,
                             0x1.53524cp+1f, 0x1.26bb1cp+1f) * x;
  k = __builtin_fminf(__builtin_fmaxf(k, -127.0f), 127.0f);
  float r = c * __int_as_float(((int32_t)k + 127) << 23);
  r = (a0 >= 0x1.31d8b2p+5f) ? kInfF : r;
  r = (a0 <= -0x1.2f3332p+5f) ? 0.0f : r;
  return (a0 != a0) ? a0 : r;
}
// NT2 double exp (d_expo_reduction.hpp:44-84, expo_base.hpp:34-50)
template <class P> NT2H_FD double exp_d(double a0) {
  double k = __builtin_rint(0x1.71547652b82fep+0 * a0);
  double hi = P::nms(k, 0x1.62e42fee00000p-1, a0);
  double lo = k * 0x1.a39ef35793c76p-33;
  double x = hi - lo;
  double t = x * x;
  double c = P::nms(t, horner<P>(t, 0x1.6376972bea4d0p-25, -0x1.bbd41c5d26bf1p-20, 0x1.1566aaf25de2cp-14,
                                 -0x1.6c16c16bebd93p-9, 0x1.555555555553ep-3), x);
  c = 1.0 - ((lo - P::div(x * c, 2.0 - c)) - hi);
  // below Maxlog k reaches 1024 (a0 in [1023.5 ln2, Maxlog)): NT2's scale
  // (expo_scale.hpp:27-35) then builds the exponent field 2047, infinity,
  // and returns inf where exp(a0) is still finite; the Fused policy scales
  // by ldexp and returns the finite value.  (The clamp only keeps the
  // integer conversion defined for arguments the selects below replace.)
  k = __builtin_fmin(__builtin_fmax(k, -1023.0), 1024.0);
  double r = P::fused ? __builtin_ldexp(c, (int)k) : c * __longlong_as_double(((long long)k + 1023) << 52);
  r = (a0 >= 0x1.62e42fefa39efp+9) ? kInfD : r;
  r = (a0 <= -0x1.6232bdd7abcd2p+9) ? 0.0 : r;
  return (a0 != a0) ? a0 : r;
}

// ================================================================ log
// NT2 float log (f_log.hpp:59-87, f_kernel.hpp:43-77): frexp mantissa in
// [sqrt(1/2), sqrt(2)), degree-8 Horner for g, recombination with the split
// ln2; subnormals prescaled by 2^23.  v_frexp_mant/exp give fast_frexp's
// mantissa and exponent for every normal input.
template <class P> NT2H_FD float log_core(float a0) {
  bool den = a0 < 0x1.0p-126f;
  float z = den ? a0 * 0x1.0p+23f : a0;
  float t = den ? -0x1.fe2804p+3f : 0.0f;
  float m = __builtin_amdgcn_frexp_mantf(z);
  int32_t e = __builtin_amdgcn_frexp_expf(z);
  bool lt = m < 0x1.6a09e6p-1f;
  float fe = lt ? (float)e + (-1.0f) : (float)e;
  float x = (lt ? m + m : m) - 1.0f;
  float x2 = x * x;
  float y = horner<P>(x, -0x1.4bbfe0p-4f, 0x1.0453f2p-3f, -0x1.09ad6ep-3f, 0x1.22dd00p-3f, -0x1.54d540p-3f,
                      0x1.99a146p-3f, -0x1.0000c8p-2f, 0x1.555552p-2f) * x * x2;
  y = P::madd(fe, -0x1.bd0106p-13f, y);
  y = P::madd(-0.5f, x2, y);
  return P::madd(0x1.63p-1f, fe, x + y + t);
}
template <class P> NT2H_FD float log_f(float a0) {
  float r = log_core<P>(a0);
  r = (a0 == kInfF) ? a0 : r;
  r = (a0 == 0.0f) ? -kInfF : r;
  return (a0 < 0.0f || a0 != a0) ? __builtin_nanf("") : r;
}
// NT2 double log (d_log.hpp:52-112): fdlibm-style s = f/(2+f) kernel; no
// subnormal prescale in the reference's double path (kept).
template <class P> NT2H_FD double log_d(double a0) {
  // Fused policy: subnormals prescaled by 2^54 (NT2's double path reads
  // their exponent field as if normal, which Exact keeps)
  const bool den = P::fused && a0 > 0.0 && a0 < 0x1.0p-1022;
  uint64_t b = ubits(den ? a0 * 0x1.0p+54 : a0);
  int64_t k = (int64_t)((b & 0x7ff0000000000000ull) >> 52) - (den ? 1022 + 54 : 1022);
  double x = dbits((b & ~0x7ff0000000000000ull) | 0x3fe0000000000000ull);
  bool lt = 0x1.6a09e667f3bcdp-1 > x;
  k = lt ? k - 1 : k;
  double f = (lt ? x + x : x + 0.0) - 1.0;
  double dk = (double)k;
  double s = P::div(f, 2.0 + f);
  double z = s * s;
  double w = z * z;
  double t1 = w * horner<P>(w, 0x1.39a09d078c69fp-3, 0x1.c71c51d8e78afp-3, 0x1.999999997fa04p-2);
  double t2 = z * horner<P>(w, 0x1.2f112df3e5244p-3, 0x1.7466496cb03dep-3, 0x1.2492494229359p-2,
                            0x1.5555555555593p-1);
  double R = t2 + t1;
  double hfsq = 0.5 * (f * f);
  double r = dk * 0x1.62e42fee00000p-1 - ((hfsq - (s * (hfsq + R) + dk * 0x1.a39ef35793c76p-33)) - f);
  r = (a0 == kInfD) ? a0 : r;
  r = (a0 == 0.0) ? -kInfD : r;
  return (a0 < 0.0 || a0 != a0) ? __builtin_nan("") : r;
}

// ============================================================ erf / erfc
// exp(-x^2).  Exact policy: NT2's exp(-sqr(x)), whose rounding of x^2 costs
// up to ~2|x| ulp(x^2) relative (about 70 ulp near x = 8).  Fused policy:
// x^2 = h + l exactly (FMA) and exp(-h-l) = exp(-h) (1 - l) to first order,
// which removes that error.
template <class P> NT2H_FD float exp_neg_of_square(float x) {
  float h = x * x;
  float e = exp_neg_f<P>(-h);
  if (P::fused) {  // (e == 0 covers x^2 = inf, where l would be NaN)
    float l = __builtin_fmaf(x, x, -h);
    e = (e != 0.0f) ? __builtin_fmaf(-l, e, e) : e;
  }
  return e;
}

// exp(-d^2/2), Fused policy with d^2 split exactly (see exp_neg_of_square).
// normcdf(d) = erfc(|t|)/2 with t = -d/sqrt2: forming t in float costs up to
// 2 t^2 ulp of relative error in exp(-t^2) (~160 ulp at d = -12.7), while
// exp(-d^2/2) from d itself does not.
template <class P> NT2H_FD float exp_neg_half_square(float d) {
  float h = d * d;
  float e = exp_neg_f<P>(-0.5f * h);
  if (P::fused) {
    float l = __builtin_fmaf(d, d, -h);
    e = (e != 0.0f) ? __builtin_fmaf(-0.5f * l, e, e) : e;
  }
  return e;
}

// NT2 float erfc (scalar/erfc.hpp:77-105, details/erf_kernel.hpp:29-89):
// z = x/(1+x); x < 2/3: (1-z)*P8(z); else exp(-x^2)*P10(z-0.4).  Both
// branches are evaluated and selected (waves mix them anyway).
// erfc_pos returns erfc(x) for x = |a0| >= 0 and, through e2, exp(-x^2)
// (0 where the small-x branch is taken and e2 was not needed).
template <class P> NT2H_FD float erfc_pos(float x, float* e2 = nullptr) {
  float z = P::div(x, 1.0f + x);
  float small = (1.0f - z) * horner<P>(z, 0x1.551caap+4f, 0x1.36ad00p+1f, -0x1.21b2acp+1f, 0x1.59fe76p-1f,
                                       -0x1.2d930ap+1f, -0x1.f92508p+0f, -0x1.422d30p+0f, -0x1.06dafcp-3f,
                                       0x1.fffffcp-1f);
  float zz = z - 0.4f;
  float ex = exp_neg_of_square<P>(x);
  if (e2) *e2 = ex;
  float big = ex * horner<P>(zz, -0x1.506c22p+1f, 0x1.acfc76p+2f, -0x1.9a350ap+2f, 0x1.a160c6p+1f,
                             -0x1.5d50cap+0f, 0x1.400de0p-3f, 0x1.22eb8ep-3f, 0x1.8994dcp-2f, 0x1.487050p-3f,
                             -0x1.2314c4p+0f, 0x1.141d16p-1f);
  float r = (x < 0x1.555556p-1f) ? small : big;
  return (x == kInfF) ? 0.0f : r;
}
// erfc(x) for x >= 0 given ex = exp(-x^2) computed elsewhere (a kernel that
// knows it from an identity); same branches and polynomials as erfc_pos.
template <class P> NT2H_FD float erfc_pos_e(float x, float ex) {
  float z = P::div(x, 1.0f + x);
  float small = (1.0f - z) * horner<P>(z, 0x1.551caap+4f, 0x1.36ad00p+1f, -0x1.21b2acp+1f, 0x1.59fe76p-1f,
                                       -0x1.2d930ap+1f, -0x1.f92508p+0f, -0x1.422d30p+0f, -0x1.06dafcp-3f,
                                       0x1.fffffcp-1f);
  float big = ex * horner<P>(z - 0.4f, -0x1.506c22p+1f, 0x1.acfc76p+2f, -0x1.9a350ap+2f, 0x1.a160c6p+1f,
                             -0x1.5d50cap+0f, 0x1.400de0p-3f, 0x1.22eb8ep-3f, 0x1.8994dcp-2f, 0x1.487050p-3f,
                             -0x1.2314c4p+0f, 0x1.141d16p-1f);
  float r = (x < 0x1.555556p-1f) ? small : big;
  return (x == kInfF) ? 0.0f : r;
}
template <class P> NT2H_FD float erfc_f(float a0) {
  float r = erfc_pos<P>(__builtin_fabsf(a0));
  r = (a0 < 0.0f) ? 2.0f - r : r;
  return (a0 != a0) ? a0 : r;
}
// NT2 float erf (scalar/erf.hpp:76-100)
template <class P> NT2H_FD float erf_f(float a0) {
  float x = __builtin_fabsf(a0);
  float small = a0 * horner<P>(x * x, -0x1.3f9076p-11f, 0x1.488d1ap-8f, -0x1.b6c3e8p-6f, 0x1.ce1e44p-4f,
                               -0x1.8126fcp-2f, 0x1.20dd74p+0f);
  float z = P::div(x, 1.0f + x) - 0.4f;
  float r2 = 1.0f - exp_neg_of_square<P>(x) * horner<P>(z, -0x1.506c22p+1f, 0x1.acfc76p+2f, -0x1.9a350ap+2f,
                                                       0x1.a160c6p+1f, -0x1.5d50cap+0f, 0x1.400de0p-3f,
                                                       0x1.22eb8ep-3f, 0x1.8994dcp-2f, 0x1.487050p-3f,
                                                       -0x1.2314c4p+0f, 0x1.141d16p-1f);
  r2 = (a0 < 0.0f) ? -r2 : r2;
  float r = (x < 0x1.555556p-1f) ? small : r2;
  r = (x == kInfF) ? __builtin_copysignf(1.0f, a0) : r;
  return (a0 != a0) ? a0 : r;
}
// NT2 double erfc / erf (scalar/erfc.hpp:39-74, scalar/erf.hpp:39-71,
// erf_kernel.hpp:91-183): four intervals split at 0.65, 2.2 and 6.
template <class P> NT2H_FD double erf1_num_d(double xx) {
  return horner<P>(xx, 0x1.10512d5b20332p-14, 0x1.3b7664358865ap-10, 0x1.4a59a4f02579cp-5,
                   0x1.16500f106c0a5p-3, 0x1.20dd750429b61p+0);
}
template <class P> NT2H_FD double erf1_den_d(double xx) {
  return horner<P>(xx, 0x1.7ea4332348252p-12, 0x1.166f75999dbd1p-7, 0x1.64536ca92ea2fp-4,
                   0x1.d0a84eb1ca867p-2, 0x1.0p+0);
}
template <class P> NT2H_FD double erfc2_num_d(double x) {
  return horner<P>(x, 0.0, 0x1.cf4cfe0aacbb4p-8, 0x1.2488a6b5cb5e5p-4, 0x1.53dd7a67c7e9fp-2,
                   0x1.c1986509e687bp-1, 0x1.54dfe9b258a60p+0, 0x1.ffffffbbb552bp-1);
}
template <class P> NT2H_FD double erfc2_den_d(double x) {
  return horner<P>(x, 0x1.9a996639b0d00p-7, 0x1.033c113a7deeep-3, 0x1.307622fcff772p-1,
                   0x1.9e677c2777c3cp+0, 0x1.53b1052dca8bdp+1, 0x1.3adeae79b9708p+1, 0x1.0p+0);
}
template <class P> NT2H_FD double erfc3_num_d(double x) {
  return horner<P>(x, 0.0, 0x1.71d0907ea7a92p-6, 0x1.42210f88b9d43p-3, 0x1.29be1cff90d94p-1,
                   0x1.44744306832aep+0, 0x1.9fa202deb88e5p+0, 0x1.fff5a9e697ae2p-1);
}
template <class P> NT2H_FD double erfc3_den_d(double x) {
  return horner<P>(x, 0x1.47bd61bbb3843p-5, 0x1.1d7ab774bb837p-2, 0x1.0cfd4cb6cde9fp+0,
                   0x1.315ffdfd5ce91p+1, 0x1.afd487397568fp+1, 0x1.602f24bf3fdb6p+1, 0x1.0p+0);
}
template <class P> NT2H_FD double erfc4_d(double rx) {
  return horner<P>(rx, -0x1.05fce04ec4ec5p+6, 0x1.5803d26c4ec4fp+6, -0x1.57e03041e9d8bp+5,
                   0x1.5ac7631f7ac4fp+2, 0x1.74fbabc514212p+1, 0x1.505a857e9ccc8p-4, -0x1.105e6b8ef1a63p+0,
                   0x1.4a451701654f1p-12, 0x1.b1494cac06d39p-2, 0x1.07e71e046a820p-22, -0x1.20dd753ae5dfdp-2,
                   0x1.60000e984b501p-36, 0x1.20dd750429a16p-1, -0x1.e4ad1ec7d0000p-56);
}
// erfc(x) for x = |a0| >= 0 (double); e2 receives exp(-x^2) when computed
// (1 for x <= 1e-10).  In a wave of mixed arguments each interval's branch runs
// under its own lane mask, yet every lane needs exactly one quotient: the
// intervals form their numerator and denominator (the far interval divides 1 by
// x ahead of its polynomial in 1/x) and one division serves them all, and
// exp(-x^2) is computed once.  Each lane performs NT2's operations on NT2's
// operands, so the values are those of the per-interval formulas.
template <class P> NT2H_FD double erfc_pos_d(double x, double* e2 = nullptr) {
  const double xx = x * x;
  const bool tiny = x <= 0.0000000001, small = x < 0.65, mid = x < 2.2, near = x < 6.0;
  double num, den;
  if (tiny) {
    num = x * 2.0;
    den = 0x1.c5bf891b4ef6ap+0;  // sqrt(pi)
  } else if (small) {
    num = erf1_num_d<P>(xx);
    den = erf1_den_d<P>(xx);
  } else if (mid) {
    num = erfc2_num_d<P>(x);
    den = erfc2_den_d<P>(x);
  } else if (near) {
    num = erfc3_num_d<P>(x);
    den = erfc3_den_d<P>(x);
  } else {
    num = 1.0;  // P::rcp(x)
    den = x;
  }
  double q = P::div(num, den);
  if (!near) q = erfc4_d<P>(q);
  double ex = 1.0;
  if (!small || (e2 != nullptr && !tiny)) ex = exp_d<P>(-xx);
  if (e2) *e2 = ex;
  return small ? 1.0 - (tiny ? q : x * q) : ex * q;
}
template <class P> NT2H_FD double erfc_d(double a0) {
  double z = erfc_pos_d<P>(__builtin_fabs(a0));
  z = (a0 < 0.0) ? 2.0 - z : z;
  return (a0 != a0) ? a0 : z;
}
// Same intervals as NT2's scalar erf, with one division and one exp(-x^2)
// serving every lane of a wave (as in erfc_pos_d).
template <class P> NT2H_FD double erf_d(double a0) {
  const double x = __builtin_fabs(a0);
  const double xx = x * x;
  const bool tiny = x <= 0.0000000001, small = x <= 0.65, mid = x <= 2.2, near = x <= 6.0;
  double num = 1.0, den = 1.0;
  if (tiny) {
    num = a0 * 2.0;
    den = 0x1.c5bf891b4ef6ap+0;  // sqrt(pi)
  } else if (small) {
    num = erf1_num_d<P>(xx);
    den = erf1_den_d<P>(xx);
  } else if (mid) {
    num = erfc2_num_d<P>(x);
    den = erfc2_den_d<P>(x);
  } else if (near) {
    num = erfc3_num_d<P>(x);
    den = erfc3_den_d<P>(x);
  }
  const double q = P::div(num, den);
  double ex = 1.0;
  if (!small && near) ex = exp_d<P>(-xx);
  double z = 1.0 - ex * q;
  z = a0 < 0.0 ? -z : z;
  // beyond 6 NT2 returns +1 whatever the sign (scalar/erf.hpp:69, reproduced
  // under Exact); the Fused policy returns sign(a0)
  const double far = P::fused ? __builtin_copysign(1.0, a0) : 1.0;
  const double r = tiny ? q : small ? a0 * q : near ? z : far;
  return (a0 != a0) ? a0 : r;
}

// ================================================ normcdf & friends
// generic/normcdf.hpp:34: Half * erfc(-Sqrt_2o_2 * a0)
template <class P> NT2H_FD float normcdf_f(float a0) {
  if (P::fused) {
    float t = -0x1.6a09e6p-1f * a0;
    float r = erfc_pos_e<P>(__builtin_fabsf(t), exp_neg_half_square<P>(a0));
    r = (t < 0.0f) ? 2.0f - r : r;
    return (a0 != a0) ? a0 : 0.5f * r;
  }
  return 0.5f * erfc_f<P>(-0x1.6a09e6p-1f * a0);
}
template <class P> NT2H_FD double normcdf_d(double a0) { return 0.5 * erfc_d<P>(-0x1.6a09e667f3bcdp-1 * a0); }
// generic/normpdf.hpp:31: exp(Mhalf*sqr(a0)) * Invsqrt_2pi
template <class P> NT2H_FD float normpdf_f(float a0) {
  if (P::fused) {  // exp(-x^2/2) with x^2 split as in exp_neg_of_square
    float h = a0 * a0;
    float l = __builtin_fmaf(a0, a0, -h);
    float e = exp_f<P>(-0.5f * h);
    e = (e != 0.0f) ? __builtin_fmaf(-0.5f * l, e, e) : e;
    return e * 0x1.988454p-2f;
  }
  return exp_f<P>(-0.5f * (a0 * a0)) * 0x1.988454p-2f;
}
template <class P> NT2H_FD double normpdf_d(double a0) {
  return exp_d<P>(-0.5 * (a0 * a0)) * 0x1.9884533d43651p-2;
}
// generic/fastnormcdf.hpp:39-54 (Abramowitz & Stegun 26.2.17)
template <class P> NT2H_FD float fastnormcdf_f(float a0) {
  float l = __builtin_fabsf(a0);
  float k = P::rcp(P::madd(0.2316419f, l, 1.0f));
  float w = horner<P>(k, 0x1.548cdep+0f, -0x1.d23dd4p+0f, 0x1.c80ef0p+0f, -0x1.6d1f0ep-2f, 0x1.470bf4p-2f);
  w *= k * 0x1.988454p-2f * exp_f<P>(-(l * l) * 0.5f);
  return a0 > 0.0f ? 1.0f - w : w;
}

// ================================================================ tanh
// NT2 tanh (scalar/tanh.hpp:45-54, details/tanh_kernel.hpp:25-80)
template <class P> NT2H_FD float tanh_f(float a0) {
  float x = __builtin_fabsf(a0);
  float x2 = x * x;
  float small = P::madd(horner<P>(x2, -0x1.75e1d4p-8f, 0x1.52269cp-6f, -0x1.b83c5ap-5f, 0x1.110726p-3f,
                                  -0x1.555532p-2f) * x2, a0, a0);
  float big = P::madd(-2.0f, P::rcp(1.0f + exp_f<P>(x + x)), 1.0f);
  big = __uint_as_float(__float_as_uint(big) ^ (__float_as_uint(a0) & 0x80000000u));
  return (x < 0.625f) ? small : big;
}
// Both branches end in a division (x^2 P/Q below 0.625, 1/(1+exp(2x)) above);
// each lane forms its operands under its branch and one division serves the
// wave, with NT2's operations on NT2's operands in every lane.
template <class P> NT2H_FD double tanh_d(double a0) {
  const double x = __builtin_fabs(a0);
  const bool small = x < 0.625;
  double num, den;
  if (small) {
    const double x2 = x * x;
    num = x2 * horner<P>(x2, -0x1.edc5baafd6f4bp-1, -0x1.8d26a0e26682dp+6, -0x1.93ac030580563p+10);
    den = horner<P>(x2, 0x1.0p+0, 0x1.c33f28a581b86p+6, 0x1.176fa0e5535fap+11, 0x1.2ec102442040cp+12);
  } else {
    num = 1.0;  // P::rcp(1 + exp(2x))
    den = 1.0 + exp_d<P>(x + x);
  }
  const double q = P::div(num, den);
  if (small) return P::madd(q, a0, a0);
  const double b = P::madd(-2.0, q, 1.0);
  return __longlong_as_double(__double_as_longlong(b) ^ (__double_as_longlong(a0) & (long long)0x8000000000000000ull));
}

// ========================================================== trigonometry
// 2/pi in 32-bit words, preceded by two zero words so that a window may
// start before the binary point.  These are the bits of 2/pi (0.A2F9836E...).
// (constexpr: one copy in device constant memory, readable from host code.)
constexpr uint32_t kTwoOverPi32[] = {
    0x00000000u, 0x00000000u, 0xA2F9836Eu, 0x4E441529u, 0xFC2757D1u, 0xF534DDC0u, 0xDB629599u, 0x3C439041u,
    0xFE5163ABu, 0xDEBBC561u, 0xB7246E3Au, 0x424DD2E0u, 0x06492EEAu, 0x09D1921Cu, 0xFE1DEB1Cu, 0xB129A73Eu,
    0xE88235F5u, 0x2EBB4484u, 0xE99C7026u, 0xB45F7E41u, 0x3991D639u, 0x835339F4u, 0x9C845F8Bu, 0xBDF9283Bu,
    0x1FF897FFu, 0xDE05980Fu, 0xEF2F118Bu, 0x5A0A6D1Fu, 0x6D367ECFu, 0x27CB09B7u, 0x4F463F66u, 0x9E5FEA2Du,
    0x7527BAC7u, 0xEBE5F17Bu, 0x3D0739F7u, 0x8A5292EAu, 0x6BFB5FB1u, 0x1F8D5D08u, 0x56033046u, 0xFC7B6BABu,
    0xF0CFBC20u, 0x9AF4361Du, 0xA9E39161u, 0x5EE61B08u, 0x6599855Fu, 0x14A06840u, 0x8DFFD880u, 0x4D732731u,
    0x06061556u, 0xCA73A8C9u};

// 32 bits of 2/pi starting at bit position pos (pos 0 = first fraction bit;
// negative positions read zeros).  pos + 64 stays inside the table for every
// finite double argument (max pos = 1023 - 52 + 3 = 974 < 46*32 - 64).
NT2H_HD uint32_t twoopi_window32(int pos) {
  int p = pos + 64;
  int w = p >> 5, s = p & 31;
  uint64_t hi = ((uint64_t)kTwoOverPi32[w] << 32) | kTwoOverPi32[w + 1];
  return (uint32_t)(s ? (hi << s) >> 32 : hi >> 32);
}

// Payne-Hanek for |x| = m * 2^e with m an integer of at most 53 bits.
// Forms x*(2/pi) mod 4 from a 160-bit window of 2/pi (bits e-2 .. e+157
// relative to the binary point), enough for |r| >= 2^-100 at 64+ correct
// bits.  Returns the quadrant (0..3) and r = (x mod pi/2) in [-pi/4, pi/4]
// as a double-double (rh + rl).
NT2H_FD int payne_hanek(uint64_t m, int e, double* rh, double* rl) {
  // window bit j (0 = most significant) has weight 2^(e-1-... ) : take
  // W = sum_j b_{e-1+j} 2^(159-j) (b_i = i-th fraction bit of 2/pi, 1-based),
  // then x*(2/pi) mod 4 = m * W * 2^-158 mod 4 + (terms below 2^-100).
  uint32_t w[5];
  int start = e - 2;  // 0-based position of the first bit of the window
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = twoopi_window32(start + 32 * i);
  // P = m * W as 7 32-bit limbs (little-endian), m split in two 32-bit halves
  uint32_t m0 = (uint32_t)m, m1 = (uint32_t)(m >> 32);
  uint32_t p[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    uint32_t mh = half ? m1 : m0;
    uint64_t carry = 0;
#pragma unroll
    for (int i = 4; i >= 0; --i) {  // w[4] is least significant
      int limb = (4 - i) + half;
      uint64_t t = (uint64_t)mh * w[i] + p[limb] + carry;
      p[limb] = (uint32_t)t;
      carry = t >> 32;
    }
    int limb = 5 + half;
    uint64_t t = (uint64_t)p[limb] + carry;
    p[limb] = (uint32_t)t;
    if (half == 0) p[6] += (uint32_t)(t >> 32);
  }
  // integer part sits above bit 158: quadrant = bits 158..159, fraction = bits 0..157
  uint32_t q = (p[4] >> 30) & 3u;
  // fraction as a 128-bit value f in [0,1): bits 157..30 of P (top of fraction)
  uint64_t fh = ((uint64_t)(p[4] & 0x3fffffffu) << 34) | ((uint64_t)p[3] << 2) | (p[2] >> 30);
  uint64_t fl = ((uint64_t)(p[2] & 0x3fffffffu) << 34) | ((uint64_t)p[1] << 2) | (p[0] >> 30);
  // fh:fl is the fraction times 2^128 (top 128 of the 158 fraction bits)
  bool neg = (fh >> 63) != 0;  // fraction >= 1/2: round to next quadrant
  if (neg) {                   // f - 1 as a magnitude: 2^128 - (fh:fl)
    fl = ~fl + 1;
    fh = ~fh + (fl == 0 ? 1 : 0);
    q = (q + 1) & 3u;
  }
  // normalise and convert to a double-double times pi/2
  int lz = fh ? __builtin_clzll(fh) : 64 + (fl ? __builtin_clzll(fl) : 64);
  if (lz >= 128) { *rh = 0.0; *rl = 0.0; return (int)q; }
  uint64_t nh, nl;
  if (lz == 0) { nh = fh; nl = fl; }
  else if (lz < 64) { nh = (fh << lz) | (fl >> (64 - lz)); nl = fl << lz; }
  else { nh = fl << (lz - 64); nl = 0; }
  // value = (nh + nl/2^64) * 2^(-64-lz)
  double hi = (double)(nh >> 11) * 0x1.0p-53;              // top 53 bits in [0.5,1)
  double lo = (double)(((nh & 0x7ffull) << 21) | (nl >> 43)) * 0x1.0p-85;
  double sc = __builtin_ldexp(1.0, -lz);
  hi *= sc; lo *= sc;
  // times pi/2 in double-double: (hi+lo)*(P1+P2)
  const double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c07p-54;
  double ph = hi * P1;
  double pl = __builtin_fma(hi, P1, -ph) + (hi * P2 + lo * P1);
  double sh = ph + pl;
  double sl = pl - (sh - ph);
  if (neg) { sh = -sh; sl = -sl; }
  *rh = sh; *rl = sl;
  return (int)q;
}

// Float argument reduction, Fused policy: |x| <= pi/4 unreduced; up to
// 2^6 pi a float 3-part Cody-Waite; below 2^23 an FP64 Cody-Waite with pi/2
// in three pieces (the first step is exact: k*P1 and x are multiples of
// 2^-32 and their difference is below 1); above 2^23 Payne-Hanek.
// x = |a0|, finite.
NT2H_FD int reduce_fused_f(float x, float* xr) {
  if (x <= 0x1.921fb6p-1f) { *xr = x; return 0; }
  if (x <= 0x1.921fb6p+7f) {
    // |x| <= 2^6 pi: float Cody-Waite with FMA; k < 2^7, so k*P1 and x are
    // multiples of 2^-23 and the first step x - k*P1 (< 1) is exact.
    float k = __builtin_rintf(x * 0x1.45f306p-1f);
    float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, x);
    r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    *xr = r;
    return ((int32_t)k) & 3;
  }
  if (x < 0x1.0p+23f) {
    double xd = (double)x;
    double k = __builtin_rint(xd * 0x1.45f306dc9c883p-1);
    double r = __builtin_fma(-k, 0x1.921fb54400000p+0, xd);
    r = __builtin_fma(-k, 0x1.0b4611a600000p-34, r);
    r = __builtin_fma(-k, 0x1.3198a2e037073p-69, r);
    *xr = (float)r;
    return (int)(((int64_t)k) & 3);
  }
  uint32_t b = __float_as_uint(x);
  double rh, rl;
  int q = payne_hanek((uint64_t)((b & 0x7fffffu) | 0x800000u), (int)(b >> 23) - 150, &rh, &rl);
  *xr = (float)rh;
  return q;
}

// fdlibm's double reduction, which NT2 runs above 2^20 pi/2 for doubles
// (trig_reduction.hpp:230-233 -> scalar/rem_pio2.hpp:415-428 ->
// __ieee754_rem_pio2 (484-622) -> __kernel_rem_pio2 with prec 2 (735-913);
// NT2 keeps the head y[0]), restated on integers held in registers, as reduce_fdlibm_f does
// for float but with fdlibm's 24-bit columns: the mantissa in three chunks
// t0 t1 t2 (24, 24 and 5 bits), 2/pi in 24-bit chunks P[], column sums
// q[i] = t0 P[jv+i] + t1 P[jv+i-1] + t2 P[jv+i-2] below 2^50 for the columns
// i <= J (J from 4), carried into 24-bit chunks B[i]; the same quadrant,
// complement and "more columns" steps; the integer part's column B[0] may
// hold up to 27 fraction bits, which fdlibm splits into two chunks.  The
// head is fdlibm's double sum, from the last column to the first, of
// F[u] 2^(q0-23-24u) with F[u] = sum_{k<=4} c_k B[u-k] (c = the 24-bit chunks
// of pi/2): every F[u] is an integer below 2^51, exact in double.  Columns
// are counted from the leading one, so every array index is a constant.
// Compared with a statement-by-statement restatement of fdlibm on sampled
// doubles of every binade and on the doubles closest to multiples of pi/2
// (tests/cpp/fdlibm_check.cpp).  x finite and > 2^20 pi/2.
constexpr uint32_t kPio2Chunks24[5] = {0xC90FDAu, 0xA22168u, 0xC234C4u, 0xC6628Bu, 0x80DC1Cu};
NT2H_HD uint64_t twoopi_chunk24(int m) { return m < 0 ? 0u : (uint64_t)(twoopi_window32(24 * m) >> 8); }
// One column sum t0 a + t1 b + t2 c of 24-bit integers, formed in double as
// fdlibm forms it (products below 2^48, the sum below 2^50: exact).  Formed
// as 64-bit integer products instead, gfx950 code from this compiler
// narrowed one to a 32-bit multiply-add whose operand had lost its 24-bit
// mask (wrong column sums on the GPU, right ones on the host).
NT2H_HD uint64_t column_d(uint64_t t0, uint64_t t1, uint64_t t2, uint32_t a, uint32_t b, uint32_t c) {
  const double s = __builtin_fma((double)t0, (double)a, __builtin_fma((double)t1, (double)b, (double)t2 * (double)c));
  return (uint64_t)s;
}
NT2H_HD int reduce_fdlibm_d(double x, double* xr) {
  constexpr int JK = 4, JMAX = 9;
  const uint64_t bx = __builtin_bit_cast(uint64_t, x);
  const int e0 = (int)((bx >> 52) & 0x7ff) - 1046;  // ilogb(x) - 23
  const uint64_t m = (bx & 0xfffffffffffffull) | 0x10000000000000ull;
  const uint64_t t0 = m >> 29, t1 = (m >> 5) & 0xffffffu, t2 = (m & 31u) << 19;
  const int jv = e0 > 3 ? (e0 - 3) / 24 : 0;
  const int q0 = e0 - 24 * (jv + 1);  // -27 .. 2
  // 2/pi chunks jv-2 .. jv+JMAX (P[k+2] = chunk jv+k); the column sums are
  // formed where they are carried
  uint32_t P[JMAX + 3];
#pragma unroll
  for (int k = 0; k < JMAX + 3; ++k) P[k] = (uint32_t)twoopi_chunk24(jv + k - 2);
  int J = JK, n = 0, ih = 0;
  uint32_t B[JMAX + 1];
  uint64_t zf;  // the fraction held in the integer part's column, in units of 2^q0 (q0 < 0)
  for (;;) {
    uint64_t c = 0;
#pragma unroll
    for (int i = JMAX; i >= 1; --i) {
      const uint64_t v = (i <= J) ? column_d(t0, t1, t2, P[i + 2], P[i + 1], P[i]) + c : 0u;
      B[i] = (uint32_t)(v & 0xffffffu);
      c = v >> 24;
    }
    const uint64_t top = column_d(t0, t1, t2, P[2], P[1], P[0]) + c;
    if (q0 >= 0) {
      n = (int)((top << q0) & 7u);
      zf = 0;
    } else {
      n = (int)((top >> -q0) & 7u);
      zf = top & ((1ull << -q0) - 1u);
    }
    ih = 0;
    if (q0 > 0) {
      const uint32_t t = B[1] >> (24 - q0);
      n += (int)t;
      B[1] -= t << (24 - q0);
      ih = (int)(B[1] >> (23 - q0));
    } else if (q0 == 0) {
      ih = (int)(B[1] >> 23);
    } else if (zf >= (1ull << (-q0 - 1))) {
      ih = 2;
    }
    if (ih > 0) {  // fraction >= 1/2: next quadrant, 1 - fraction
      n += 1;
      uint32_t borrow = 0;
#pragma unroll
      for (int i = JMAX; i >= 1; --i) {
        if (i <= J) {
          const uint32_t j = B[i];
          B[i] = borrow ? 0xffffffu - j : (j ? 0x1000000u - j : 0u);
          borrow |= (j != 0u) ? 1u : 0u;
        }
      }
      if (q0 > 0) B[1] &= 0xffffffu >> q0;
      if (ih == 2) zf = (1ull << -q0) - zf - borrow;
    }
    if (zf == 0) {
      uint32_t lead = 0;
#pragma unroll
      for (int i = 1; i <= JMAX - JK; ++i)
        if (i <= J - JK) lead |= B[i];
      if (lead == 0) {  // leading chunks vanish: take columns up to 4 past the first nonzero chunk
        int s = 0;
#pragma unroll
        for (int i = JMAX; i >= 1; --i)
          if (i <= J && B[i] != 0u) s = i;
        if (s == 0 || s > JMAX - JK) break;  // not reached by the doubles checked
        J = JK + s;
        continue;
      }
    }
    break;
  }
  // the integer part's column, split as fdlibm splits it above 24 bits
  const uint32_t Bt = (uint32_t)(zf >> 24), B0 = (uint32_t)(zf & 0xffffffu);
  double sum = 0.0;
#pragma unroll
  for (int u = JMAX; u >= -1; --u) {
    uint64_t F = 0;
#pragma unroll
    for (int k = 0; k <= 4; ++k) {
      const int d = u - k;
      if (d < -1) break;
      const uint32_t bd = d == -1 ? Bt : (d == 0 ? B0 : ((d <= J) ? B[d] : 0u));
      F += (uint64_t)kPio2Chunks24[k] * bd;
    }
    if (u > J) F = 0;
    sum = __builtin_fma((double)F, __builtin_ldexp(1.0, -24 * u), sum);
  }
  const double y0 = __builtin_ldexp(sum, q0 - 23);
  *xr = ih ? -y0 : y0;
  return n & 3;
}

// Double reduction: NT2's double ladder (straight, Cephes, medium up to
// 2^18 pi) then fdlibm's double medium up to 2^20 pi/2 (scalar/rem_pio2.hpp:
// 484-612, with its cancellation checks), then fdlibm's Payne-Hanek (Exact)
// or the integer Payne-Hanek (Fused: the correctly rounded argument).
template <class P> NT2H_FD int reduce_d(double x, double* xr) {
  if (!(x > 0x1.921fb54442d18p-1)) { *xr = x; return 0; }
  if (!(x > 0x1.921fb54442d18p+0)) {
    double t = x - 0x1.921fb54400000p+0;
    t = t - 0x1.0b4611a600000p-34;
    t = t - 0x1.3198a2e000000p-69;
    *xr = t;
    return 1;
  }
  if (!(x > 0x1.f6a7a2955385ep+5)) {
    double xi = __builtin_rint(x * 0x1.45f306dc9c883p-1);
    double t = P::nms(xi, 0x1.921fb54400000p+0, x);
    t = P::nms(xi, 0x1.0b4611a600000p-34, t);
    t = P::nms(xi, 0x1.3198a2e000000p-69, t);
    *xr = t;
    return ((int32_t)xi) & 3;
  }
  if (!(x > 0x1.921fb54442d18p+19)) {
    const double fn = __builtin_rint(x * 0x1.45f306dc9c883p-1);
    double r = x - fn * 0x1.921fb54400000p+0;
    double w = fn * 0x1.0b4611a626331p-34;
    double t2 = r;
    w = fn * 0x1.0b4611a600000p-34;
    r = t2 - w;
    w = fn * 0x1.3198a2e037073p-69 - ((t2 - r) - w);
    t2 = r;
    w = fn * 0x1.3198a2e000000p-69;
    r = t2 - w;
    w = fn * 0x1.b839a252049c1p-104 - ((t2 - r) - w);
    *xr = r - w;
    return ((int32_t)fn) & 3;
  }
  uint64_t bx = ubits(x);
  int32_t ix = (int32_t)(bx >> 32) & 0x7fffffff;
  if (ix <= 0x413921fb) {  // fdlibm medium, n < 2^20
    double t = x;
    int32_t n = (int32_t)(t * 0x1.45f306dc9c883p-1 + 0.5);
    double fn = (double)n;
    double r = t - fn * 0x1.921fb54400000p+0;
    double w = fn * 0x1.0b4611a626331p-34;
    int32_t j = ix >> 20;
    double y0 = r - w;
    int32_t i = j - (int32_t)((ubits(y0) >> 52) & 0x7ff);
    if (i > 16) {
      t = r; w = fn * 0x1.0b4611a600000p-34; r = t - w; w = fn * 0x1.3198a2e037073p-69 - ((t - r) - w);
      y0 = r - w;
      i = j - (int32_t)((ubits(y0) >> 52) & 0x7ff);
      if (i > 49) {
        t = r; w = fn * 0x1.3198a2e000000p-69; r = t - w; w = fn * 0x1.b839a252049c1p-104 - ((t - r) - w);
        y0 = r - w;
      }
    }
    *xr = y0;
    return n & 3;
  }
  if constexpr (!P::fused) return reduce_fdlibm_d(x, xr);  // NT2's result bit for bit
  double rh, rl;
  int q = payne_hanek((bx & 0xfffffffffffffull) | 0x10000000000000ull, (int)((bx >> 52) & 0x7ff) - 1075, &rh, &rl);
  *xr = rh;
  return q;
}
// The float band (2^6 pi, 2^18 pi] of the Exact reduction, where NT2
// converts the float to double and runs its double medium reduction (five
// products, nine differences; fdlibm's double medium for the last float,
// 0x1.921fb6p+19).  Two FMAs with pi/2 in two double pieces (the first
// exact: k*P1 and x are multiples of 2^-52 and their difference is below 2)
// give, after rounding to float, NT2's xr and quadrant on every float of the
// band — and the correctly rounded xr with the nearest quadrant on every
// float above, up to 2^28 (both checked exhaustively,
// tests/cpp/trig_band_check.cpp).  Outside the band the result is
// meaningless (callers select).
constexpr float kExactBandLoF = 0x1.921fb6p+7f;   // exclusive
constexpr float kExactBandHiF = 0x1.921fb6p+19f;  // inclusive: fdlibm's float reduction above
NT2H_FD bool in_exact_band_f(float x) { return x > kExactBandLoF && x <= kExactBandHiF; }
NT2H_FD int reduce_exact_f_band(float x, float* xr) {
  const double xd = (double)x;
  const double k = __builtin_rint(xd * 0x1.45f306dc9c883p-1);
  double r = __builtin_fma(-k, 0x1.921fb54442d18p+0, xd);
  r = __builtin_fma(-k, 0x1.1a62633145c07p-54, r);
  *xr = (float)r;
  return ((int32_t)k) & 3;
}
// Float argument reduction, Exact policy: NT2's scalar ladder for mode big_
// (trig_reduction.hpp:94-247) — straight (<= pi/2), Cephes 3-part (<= 20 pi),
// fdlibm-medium in float (<= 2^6 pi) — then the band above, up to 2^18 pi
// (x <= 2^18 pi here).
NT2H_FD int reduce_exact_f_near(float x, float* xr) {
  if (!(x > 0x1.921fb6p-1f)) { *xr = x; return 0; }
  if (!(x > 0x1.921fb6p+0f)) {
    float t = x - 0x1.921f00p+0f;
    t = t - 0x1.6a8800p-17f;
    t = t - 0x1.0b4600p-34f;
    *xr = t;
    return 1;
  }
  if (!(x > 0x1.f6a7a2p+5f)) {  // 20*pi
    float xi = __builtin_rintf(x * 0x1.45f306p-1f);
    float t = x - xi * 0x1.921f00p+0f;
    t = t - xi * 0x1.6a8800p-17f;
    t = t - xi * 0x1.0b4600p-34f;
    *xr = t;
    return ((int32_t)xi) & 3;
  }
  if (!(x > 0x1.921fb6p+7f)) {  // 2^6 pi: rem_pio2_medium<float>
    const float fn = __builtin_rintf(x * 0x1.45f308p-1f);
    float r = x - fn * 0x1.921f00p+0f;
    float w = fn * 0x1.6a8886p-17f;
    float t2 = r;
    w = fn * 0x1.6a8800p-17f;
    r = t2 - w;
    w = fn * 0x1.0b4610p-34f - ((t2 - r) - w);
    t2 = r;
    w = fn * 0x1.0b4600p-34f;
    r = t2 - w;
    w = fn * 0x1.1a6264p-54f - ((t2 - r) - w);
    *xr = r - w;
    return ((int32_t)fn) & 3;
  }
  return reduce_exact_f_band(x, xr);
}
// fdlibm's float reduction, which NT2 runs above 2^18 pi
// (trig_reduction.hpp:227-232 -> __ieee754_rem_pio2f -> __kernel_rem_pio2f
// with prec 2, scalar/rem_pio2.hpp:83-233; NT2 keeps the head y[0]),
// restated on integers held in registers.  The method is Payne-Hanek as
// fdlibm organises it for float: the mantissa in three bytes t0 t1 t2, 2/pi
// in bytes P[], column sums q[i] = t0 P[jv+i] + t1 P[jv+i-1] + t2 P[jv+i-2]
// of weight 2^(q0-8i) for the columns i <= J, carried into bytes; the
// integer part gives the quadrant, a fraction >= 1/2 is complemented (next
// quadrant), and while the fraction's leading bytes vanish more columns are
// taken (J grows from 9 to at most 13 over the floats).  The head is
// fdlibm's float sum, from the last column to the first, of
// fq[u] = F[u] 2^(q0-7-8u) with F[u] = sum_{k<=9} c_k B[u-k] (c = the bytes
// of pi/2, B = the fraction's bytes, B[0] the column of the integer part):
// every F[u] is an integer below 2^19, so fq[u] is exact and the sum rounds
// once per column as fdlibm's does.
//
// Here the bytes sit big-endian in four words, byte i of the frame being
// column i, so that four columns at once are one product: the column sums of
// bytes 4a .. 4a+3 are t0 W(jv+4a) + t1 W(jv+4a-1) + t2 W(jv+4a-2), W(p) the
// 32-bit word of 2/pi bytes P[p .. p+3] (columns beyond J masked off), and
// the carries run over four 64-bit words instead of thirteen bytes; the
// complement is a negation of the fraction field, and F[u] are 4-byte dot
// products with constant words.  Bit-identical to a statement-by-statement
// restatement of fdlibm on every float above 2^18 pi
// (tests/cpp/fdlibm_check.cpp).  x finite and > 2^18 pi.
constexpr uint32_t kPio2Bytes[10] = {0xC9u, 0x0Fu, 0xDAu, 0xA2u, 0x21u, 0x68u, 0xC2u, 0x34u, 0xC4u, 0xC6u};
constexpr float kFdlibmLoF = 0x1.921fb6p+19f;  // exclusive: NT2 reduces in double up to here
NT2H_HD uint32_t dot4_u8(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_udot4(a, b, 0u, false);
#else
  return (a & 255u) * (b & 255u) + ((a >> 8) & 255u) * ((b >> 8) & 255u) + ((a >> 16) & 255u) * ((b >> 16) & 255u) +
         (a >> 24) * (b >> 24);
#endif
}
// big-endian word of the constants that meet bytes 4a .. 4a+3 of the frame
// in F[u]
constexpr uint32_t pio2_word(int u, int a) {
  uint32_t wv = 0;
  for (int j = 0; j < 4; ++j) {
    const int k = u - 4 * a - j;
    wv |= ((k >= 0 && k <= 9) ? kPio2Bytes[k] : 0u) << (24 - 8 * j);
  }
  return wv;
}
NT2H_HD int reduce_fdlibm_f(float x, float* xr) {
  constexpr int JMAX = 13;
  const uint32_t b = __builtin_bit_cast(uint32_t, x);
  const int e0 = (int)(b >> 23) - 134;  // ilogb(x) - 7, >= 12 here
  const uint32_t m = (b & 0x7fffffu) | 0x800000u;
  const uint64_t t0 = m >> 16, t1 = (m >> 8) & 255u, t2 = m & 255u;
  const int jv = (e0 - 3) >> 3;
  const int q0 = e0 - 8 * (jv + 1);  // -5 .. 2
  // bytes jv-2 .. jv+17 of 2/pi as big-endian words (byte -1 reads zero),
  // then W(jv + 4a - j) for the four frame words and j = 0, 1, 2
  uint32_t w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = twoopi_window32(8 * (jv - 2) + 32 * i);
  int J = 9, n = 0, ih = 0;
  uint64_t hi, lo;  // the fraction's bytes B[1..7] (56 bits) and B[8..13] (top 48 bits)
  uint32_t zf;      // the fraction held in the integer part's column, in units of 2^q0 (q0 < 0)
  for (;;) {
    // column sums of the four frame words (last first, carried upwards),
    // columns beyond J masked off; W(jv+4a-j) = bytes jv+4a-j .. of 2/pi
    uint32_t wd[4];
    uint64_t v = 0;
#pragma unroll
    for (int a = 3; a >= 0; --a) {
      uint32_t mask = 0xffffffffu;
      if (a == 2) mask = J >= 11 ? 0xffffffffu : (J == 10 ? 0xffffff00u : 0xffff0000u);
      if (a == 3) mask = J >= 13 ? 0xffff0000u : (J == 12 ? 0xff000000u : 0u);
      const uint32_t W2 = w[a] & mask;
      const uint32_t W1 = ((w[a] << 8) | (w[a + 1] >> 24)) & mask;
      const uint32_t W0 = ((w[a] << 16) | (w[a + 1] >> 16)) & mask;
      v = t0 * W0 + t1 * W1 + t2 * W2 + (v >> 32);
      wd[a] = (uint32_t)v;
    }
    const uint64_t v0 = v;
    const uint32_t top = (uint32_t)(v0 >> 24);  // column 0 with everything carried into it
    hi = ((v0 & 0xffffffull) << 32) | wd[1];
    lo = ((uint64_t)wd[2] << 32) | wd[3];
    if (q0 >= 0) {
      n = (int)((top << q0) & 7u);
      zf = 0;
    } else {
      n = (int)((top >> -q0) & 7u);
      zf = top & ((1u << -q0) - 1u);
    }
    ih = 0;
    if (q0 > 0) {  // the field's top q0 bits belong to the quadrant
      n += (int)(hi >> (56 - q0));
      hi &= 0x00ffffffffffffffull >> q0;
      ih = (int)(hi >> (55 - q0));
    } else if (q0 == 0) {
      ih = (int)(hi >> 55);
    } else if (zf >= (1u << (-q0 - 1))) {
      ih = 2;
    }
    if (ih > 0) {  // fraction >= 1/2: next quadrant, 1 - fraction (fdlibm's bytewise complement)
      n += 1;
      const uint32_t borrow = (hi | lo) != 0 ? 1u : 0u;
      const uint64_t nlo = (uint64_t)0 - lo;
      hi = (~hi + (lo == 0 ? 1u : 0u)) & (0x00ffffffffffffffull >> (q0 > 0 ? q0 : 0));
      lo = nlo;
      if (ih == 2) zf = (1u << -q0) - zf - borrow;
    }
    if (zf == 0 && (J == 9 || (hi >> (56 - 8 * (J - 9))) == 0)) {
      // leading bytes vanish: take columns up to 9 past the first nonzero byte
      int s;
      if (hi != 0) s = (__builtin_clzll(hi) - 8) / 8 + 1;
      else if (lo != 0) s = __builtin_clzll(lo) / 8 + 8;
      else s = 0;
      if (s == 0 || s > JMAX - 9) break;  // no float gets here (checked on every float)
      J = 9 + s;
      continue;
    }
    break;
  }
  // F[u] as 4-byte dot products of the frame words with constant words
  const uint32_t Bw[4] = {(zf << 24) | (uint32_t)(hi >> 32), (uint32_t)hi, (uint32_t)(lo >> 32), (uint32_t)lo};
  float sum = 0.0f;
#pragma unroll
  for (int u = JMAX; u >= 0; --u) {
    uint32_t F = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const uint32_t cw = pio2_word(u, a);
      if (cw != 0u) F += dot4_u8(Bw[a], cw);
    }
    if (u > J) F = 0;
    sum = __builtin_fmaf((float)F, __builtin_ldexpf(1.0f, -8 * u), sum);
  }
  const float y0 = __builtin_ldexpf(sum, q0 - 7);
  *xr = ih ? -y0 : y0;
  return n & 3;
}

// The same result for x in (2^18 pi, 2^28] from the double reduced argument
// of the two-FMA band reduction, when it is decidable from there; returns
// false otherwise (about 4% of these floats), and the caller runs
// reduce_fdlibm_f.  fdlibm's head is a float sum, last column first, of
// nonnegative exact terms fq[u]; with s* the fraction's leading column,
// y0 = fq[s*] + (fq[s*+1] + s_2) rounded twice, s_2 = RN(fq[s*+2] + s_3) and
// s_3 the sum of the remaining columns, which differs from their exact total
// T_3 by at most 2^-24 (T_3 + 7 T_4) (one rounding at T_3, at most seven at
// columns below T_4).  The two-FMA reduction gives |r| = x mod pi/2 to about
// 2^-53 relative; the fraction g = |r| 2/pi gives the four leading bytes
// (hence fq[s*] .. fq[s*+3] exactly) and T_2 = |r| - fq[s*] - fq[s*+1]
// is exact in double.  Unless T_2 lies within that bound of a float
// rounding midpoint (or the bytes sit within 2^-14 of a boundary),
// s_2 = RN(T_2), and the last two sums are float additions.  Bit-identical to
// reduce_fdlibm_f wherever it returns true (tests/cpp/fdlibm_check.cpp walks
// every float).
constexpr float kFdlibmFastHiF = 0x1.0p+28f;  // inclusive
NT2H_HD bool reduce_fdlibm_f_fast(float x, float* xr, int* quadrant) {
  const double xd = (double)x;
  const double k = __builtin_rint(xd * 0x1.45f306dc9c883p-1);
  double r = __builtin_fma(-k, 0x1.921fb54442d18p+0, xd);
  r = __builtin_fma(-k, 0x1.1a62633145c07p-54, r);
  const double a = __builtin_fabs(r);
  const uint32_t b = __builtin_bit_cast(uint32_t, x);
  const int e0 = (int)(b >> 23) - 134;
  const int q0 = e0 - 8 * (((e0 - 3) >> 3) + 1);
  // g = |r| 2/pi; G = g 2^-q0: column 0 is its integer part, column s its
  // s-th fraction byte
  const double g = a * 0x1.45f306dc9c883p-1;
  const int eg = (int)((__builtin_bit_cast(uint64_t, g) >> 52) & 0x7ff) - 1023 - q0;
  const int s = eg >= 0 ? 0 : (7 - eg) >> 3;  // leading column
  const double t = __builtin_ldexp(g, 8 * (s + 3) - q0);  // the four leading bytes, < 2^32
  const double ti = __builtin_floor(t);
  const double tf = t - ti;
  const uint32_t I = (uint32_t)ti;
  const uint32_t B0 = I >> 24, B1 = (I >> 16) & 255u, B2 = (I >> 8) & 255u, B3 = I & 255u;
  const uint32_t F0 = kPio2Bytes[0] * B0, F1 = kPio2Bytes[0] * B1 + kPio2Bytes[1] * B0;
  const uint32_t F2 = kPio2Bytes[0] * B2 + kPio2Bytes[1] * B1 + kPio2Bytes[2] * B0;
  const uint32_t F3 = kPio2Bytes[0] * B3 + kPio2Bytes[1] * B2 + kPio2Bytes[2] * B1 + kPio2Bytes[3] * B0;
  // in units of 2^(q0-7-8s), where fq[s+j] = F_j 2^-8j
  const double A = __builtin_ldexp(a, 7 - q0 + 8 * s);
  const double T2 = __builtin_fma(-(double)F1, 0x1.0p-8, A - (double)F0);
  const double T3 = __builtin_fma(-(double)F2, 0x1.0p-16, T2);
  const double T4 = __builtin_fma(-(double)F3, 0x1.0p-24, T3);
  // |s_3 - T_3|: one rounding of s_3 and at most seven of columns at most T_4
  const double eps = __builtin_fma(__builtin_fma(7.0, T4, T3), 0x1.002p-24, A * 0x1.0p-50);
  // distance of T2 from the float rounding midpoints of its binade
  const int e2 = (int)((__builtin_bit_cast(uint64_t, T2) >> 52) & 0x7ff) - 1023;
  const double u2 = __builtin_ldexp(T2, 23 - e2);  // T2 in float ulps of its binade
  const double dist = __builtin_fabs((u2 - __builtin_floor(u2)) - 0.5);
  const bool ok = B0 != 0u && tf > 0x1.0p-14 && tf < 1.0 - 0x1.0p-14 && dist > __builtin_ldexp(eps, 23 - e2);
  const float s2 = (float)T2;
  const float s1 = (float)F1 * 0x1.0p-8f + s2;
  const float y0 = __builtin_ldexpf((float)F0 + s1, q0 - 7 - 8 * s);
  *xr = r < 0.0 ? -y0 : y0;
  *quadrant = ((int32_t)k) & 3;
  return ok;
}

NT2H_FD int reduce_exact_f(float x, float* xr) {
  if (!(x > kFdlibmLoF)) return reduce_exact_f_near(x, xr);
  int q;
  if (x <= kFdlibmFastHiF && reduce_fdlibm_f_fast(x, xr, &q)) return q;
  return reduce_fdlibm_f(x, xr);
}

// f_trig_evaluation.hpp:40-48
template <class P> NT2H_FD float cos_eval_f(float z) {
  float y = horner<P>(z, 0x1.99eb9cp-16f, -0x1.6c0c32p-10f, 0x1.55554ap-5f);
  return 1.0f + P::madd(z, -0.5f, y * (z * z));
}
template <class P> NT2H_FD float sin_eval_f(float z, float x) {
  float y1 = horner<P>(z, -0x1.9943f2p-13f, 0x1.11073ap-7f, -0x1.555544p-3f);
  return P::madd(y1 * z, x, x);
}
template <class P> NT2H_FD int reduce_f(float x, float* xr) {
  return P::fused ? reduce_fused_f(x, xr) : reduce_exact_f(x, xr);
}
// trig_base.hpp:131-161: cos / sin of a0 from the quadrant n and reduced
// argument xr of |a0| (inf and NaN give NaN)
// (both polynomials are evaluated and one selected: no divergent branch)
template <class P> NT2H_FD float cos_from(float a0, int n, float xr) {
  bool bad = !(__builtin_fabsf(a0) <= 0x1.fffffep+127f);
  float z = xr * xr;
  float se = sin_eval_f<P>(z, xr), ce = cos_eval_f<P>(z);
  float r = (n & 1) ? se : ce;
  uint32_t sign = (uint32_t)(((n & 1) ^ ((n & 2) >> 1))) << 31;
  r = __uint_as_float(__float_as_uint(r) ^ sign);
  return bad ? __builtin_nanf("") : r;
}
template <class P> NT2H_FD float sin_from(float a0, int n, float xr) {
  bool bad = !(__builtin_fabsf(a0) <= 0x1.fffffep+127f);
  float z = xr * xr;
  float se = sin_eval_f<P>(z, xr), ce = cos_eval_f<P>(z);
  float r = (n & 1) ? ce : se;
  uint32_t sign = (__float_as_uint(a0) & 0x80000000u) ^ ((uint32_t)(n & 2) << 30);
  r = __uint_as_float(__float_as_uint(r) ^ sign);
  return bad ? __builtin_nanf("") : r;
}
template <class P> NT2H_FD float cos_f(float a0) {
  float x = __builtin_fabsf(a0), xr;
  bool bad = !(x <= 0x1.fffffep+127f);
  int n = reduce_f<P>(bad ? 0.0f : x, &xr);
  return cos_from<P>(a0, n, xr);
}
template <class P> NT2H_FD float sin_f(float a0) {
  float x = __builtin_fabsf(a0), xr;
  bool bad = !(x <= 0x1.fffffep+127f);
  int n = reduce_f<P>(bad ? 0.0f : x, &xr);
  return sin_from<P>(a0, n, xr);
}
// sin and cos of one argument sharing the reduction (Fused kernels); inf
// and NaN give NaN, as sin_f and cos_f do.
template <class P> NT2H_FD void sincos_f(float a0, float* s, float* c) {
  float x = __builtin_fabsf(a0), xr;
  bool bad = !(x <= 0x1.fffffep+127f);
  int n = reduce_f<P>(bad ? 0.0f : x, &xr);
  float z = xr * xr;
  float se = sin_eval_f<P>(z, xr), ce = cos_eval_f<P>(z);
  bool sw = (n & 1) != 0;
  float sv = sw ? ce : se, cv = sw ? se : ce;
  uint32_t ssign = (__float_as_uint(a0) & 0x80000000u) ^ ((uint32_t)(n & 2) << 30);
  uint32_t csign = (uint32_t)(((n & 1) ^ ((n & 2) >> 1))) << 31;
  *s = bad ? __builtin_nanf("") : __uint_as_float(__float_as_uint(sv) ^ ssign);
  *c = bad ? __builtin_nanf("") : __uint_as_float(__float_as_uint(cv) ^ csign);
}

template <class P> NT2H_FD double cos_eval_d(double z) {
  double y = horner<P>(z, 0x1.8ff831ad9b219p-37, -0x1.1eea7c1e514d4p-29, 0x1.27e4f8e06d9a5p-22,
                       -0x1.a01a019ddbcd9p-16, 0x1.6c16c16c15d47p-10, -0x1.5555555555551p-5, 0x1.0p-1);
  return 1.0 - y * z;
}
template <class P> NT2H_FD double sin_eval_d(double z, double x) {
  double y1 = horner<P>(z, 0x1.5d8fd1fcf0ec1p-33, -0x1.ae5e5a9291691p-26, 0x1.71de3567d4896p-19,
                        -0x1.a01a019bfdf03p-13, 0x1.111111110f7d0p-7, -0x1.5555555555548p-3);
  return P::madd(y1 * z, x, x);
}
template <class P> NT2H_FD double cos_d(double a0) {
  double x = __builtin_fabs(a0), xr;
  bool bad = !(x <= 0x1.fffffffffffffp+1023);
  int n = reduce_d<P>(bad ? 0.0 : x, &xr);
  double z = xr * xr;
  double r = (n & 1) ? sin_eval_d<P>(z, xr) : cos_eval_d<P>(z);
  uint64_t sign = (uint64_t)(((n & 1) ^ ((n & 2) >> 1))) << 63;
  r = dbits(ubits(r) ^ sign);
  return bad ? __builtin_nan("") : r;
}
template <class P> NT2H_FD double sin_d(double a0) {
  double x = __builtin_fabs(a0), xr;
  bool bad = !(x <= 0x1.fffffffffffffp+1023);
  int n = reduce_d<P>(bad ? 0.0 : x, &xr);
  double z = xr * xr;
  double r = (n & 1) ? cos_eval_d<P>(z) : sin_eval_d<P>(z, xr);
  uint64_t sign = (ubits(a0) & 0x8000000000000000ull) ^ ((uint64_t)(n & 2) << 62);
  r = dbits(ubits(r) ^ sign);
  return bad ? __builtin_nan("") : r;
}

}  // namespace nt2h
