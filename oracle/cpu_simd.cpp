// Vectorised CPU evaluation of the Black-Scholes forward + adjoint — TEST
// INFRASTRUCTURE ONLY: the cpu_baseline leg of bench.py times it; nothing in
// the product loads it.
//
// NT2's SIMD path evaluates each functor on a whole pack with both sides of
// every branch computed and blended (e.g. modules/core/euler/include/nt2/
// euler/functions/simd/common/erfc.hpp:110-140), under the shared-memory
// grain policy (arch/shared_memory/include/nt2/core/functions/
// shared_memory/transform.hpp:39-49).  The functions below restate the
// scalar oracle (nt2ref.hpp) in that branch-free form so that the compiler
// vectorises the option loop (#pragma omp simd, AVX2); every operation is the
// same IEEE operation in the same order, so the results are bit-identical to
// the scalar oracle (checked in tests/test_oracle.py).
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>

// One clone per instruction set, chosen at load time by the host CPU: AVX-512
// (16 lanes), AVX2 (8 lanes), or the x86-64 baseline.
#define NT2REF_CLONES __attribute__((target_clones("avx512f", "avx2", "default")))

namespace {

inline float fb(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
inline uint32_t bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

// round2even (Boost.SIMD): |a| + 2^23 - 2^23 below 2^23, sign restored
inline float round2even(float a0) {
  const float t2n = 8388608.0f;
  float v = std::fabs(a0);
  float d = (v + t2n) - t2n;
  d = (v < t2n) ? d : v;
  return a0 < 0.0f ? -d : d;
}

// exp: f_expo_reduction.hpp:49-85, expo_base.hpp:54-69, expo_scale.hpp:27-35
inline float exp_bf(float a0) {
  const float maxlog = fb(0x42b0c0a5u), minlog = fb(0xc2b0c0a5u);
  // clamp first so that the integer steps stay defined; the specials are
  // blended in at the end
  float a = (a0 > minlog) ? a0 : minlog;
  a = (a < maxlog) ? a : maxlog;
  float k = round2even(fb(0x3fb8aa3bu) * a);
  float p = k * fb(0x3f318000u);
  float x = a - p;
  p = k * fb(0xb95e8083u);
  x = x - p;
  float y = fb(0x3ab778cf);
  y = x * y + fb(0x3c098d8b);
  y = x * y + fb(0x3d2aa957);
  y = x * y + fb(0x3e2aa9a5);
  y = x * y + fb(0x3f000000);
  float c = 1.0f + (y * (x * x) + x);
  int32_t ik = (int32_t)k + 127;
  float r = c * fb((uint32_t)ik << 23);
  r = (a0 >= maxlog) ? fb(0x7f800000u) : r;
  r = (a0 <= minlog) ? 0.0f : r;
  return (a0 != a0) ? a0 : r;
}

// log: f_log.hpp:59-87 + f_kernel.hpp:43-77
inline float log_bf(float a0) {
  const bool sub = std::fabs(a0) < fb(0x00800000u);
  float z = sub ? a0 * fb(0x4b000000u) : a0;
  float t = sub ? fb(0xc17f1402u) : 0.0f;
  uint32_t b = bf(z);
  int32_t e = (int32_t)((b & 0x7f800000u) >> 23) - 126;
  float x = fb((b & ~0x7f800000u) | 0x3f000000u);
  bool lt = x < fb(0x3f3504f3u);
  float fe = lt ? (float)e + (-1.0f) : (float)e;
  x = (lt ? x + x : x) - 1.0f;
  float x2 = x * x;
  float h = fb(0xbda5dff0u);  // the coefficients of the scalar oracle's kLogF
  h = x * h + fb(0x3e0229f9u);
  h = x * h + fb(0xbe04d6b7u);
  h = x * h + fb(0x3e116e80u);
  h = x * h + fb(0xbe2a6aa0u);
  h = x * h + fb(0x3e4cd0a3u);
  h = x * h + fb(0xbe800064u);
  h = x * h + fb(0x3eaaaaa9u);
  float y = h * x * x2;
  y = fe * fb(0xb95e8083u) + y;
  y = -0.5f * x2 + y;
  float r = fb(0x3f318000u) * fe + (x + y + t);
  r = (a0 == fb(0x7f800000u)) ? a0 : r;
  r = (a0 == 0.0f) ? fb(0xff800000u) : r;
  return ((a0 != a0) | (a0 < 0.0f)) ? fb(0x7fc00000u) : r;
}

// erfc: scalar/erfc.hpp:77-105 with both branches evaluated and blended
inline float erfc_bf(float a0) {
  float x = std::fabs(a0);
  float z = x / (1.0f + x);
  // (1 - z) * P8(z) below 2/3
  float s = fb(0x41aa8e55);
  s = z * s + fb(0x401b5680);
  s = z * s + fb(0xc010d956);
  s = z * s + fb(0x3f2cff3b);
  s = z * s + fb(0xc016c985);
  s = z * s + fb(0xbffc9284);
  s = z * s + fb(0xbfa11698);
  s = z * s + fb(0xbe036d7e);
  s = z * s + fb(0x3f7ffffe);
  float r_small = (1.0f - z) * s;
  // exp(-x^2) * P11(z - 0.4) above
  float zz = z - 0.4f;
  float q = fb(0xc0283611);
  q = zz * q + fb(0x40d67e3b);
  q = zz * q + fb(0xc0cd1a85);
  q = zz * q + fb(0x4050b063);
  q = zz * q + fb(0xbfaea865);
  q = zz * q + fb(0x3e2006f0);
  q = zz * q + fb(0x3e1175c7);
  q = zz * q + fb(0x3ec4ca6e);
  q = zz * q + fb(0x3e243828);
  q = zz * q + fb(0xbf918a62);
  q = zz * q + fb(0x3f0a0e8b);
  float r_big = exp_bf(-(x * x)) * q;
  float r1 = (x < fb(0x3f2aaaabu)) ? r_small : ((x == fb(0x7f800000u)) ? 0.0f : r_big);
  return (a0 < 0.0f) ? 2.0f - r1 : r1;
}
inline float normcdf_bf(float a0) { return 0.5f * erfc_bf(-fb(0x3F3504F3u) * a0); }
inline float normpdf_bf(float a0) { return exp_bf(-0.5f * (a0 * a0)) * fb(0x3ecc422au); }

// The option loop of one grain: the oracle's bs_fwd_adj<float> statement by
// statement.
NT2REF_CLONES void bs_grain(const float* __restrict S, const float* __restrict K, const float* __restrict Tm,
              const float* __restrict r, const float* __restrict sig, const float* __restrict w,
              float* __restrict price, float* __restrict dS, float* __restrict dK, float* __restrict dT,
              float* __restrict dr, float* __restrict dsig, int64_t n) {
#pragma omp simd
  for (int64_t i = 0; i < n; ++i) {
    float sqT = std::sqrt(Tm[i]);
    float sst = sig[i] * sqT;
    float d1 = (log_bf(S[i] / K[i]) + (r[i] + (sig[i] * sig[i]) * 0.5f) * Tm[i]) / sst;
    float d2 = d1 - sst;
    float ert = exp_bf(-r[i] * Tm[i]);
    float kert = K[i] * ert;
    float n1 = normcdf_bf(d1);
    float n2 = normcdf_bf(d2);
    float phi = normpdf_bf(d1);
    float kn2 = kert * n2;
    price[i] = S[i] * n1 - kn2;
    dS[i] = w[i] * n1;
    dK[i] = -(w[i] * (ert * n2));
    float sphi = S[i] * phi;
    dT[i] = w[i] * (sphi * sig[i] / (2.0f * sqT) + r[i] * kn2);
    dr[i] = w[i] * (Tm[i] * kn2);
    dsig[i] = w[i] * (sphi * sqT);
  }
}

}  // namespace

extern "C" {

// The instruction set the clones above run with on this host (the loader
// picks the first supported of avx512f, avx2, baseline).
const char* nt2ref_cpu_simd_isa() {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f")) return "avx512f";
  if (__builtin_cpu_supports("avx2")) return "avx2";
  return "x86-64";
}

// The branch-free functors over an array (op ids as the oracle's: 0 exp,
// 1 log, 4 erfc, 5 normcdf, 9 normpdf), for checking them against the
// scalar restatement.
NT2REF_CLONES void nt2ref_cpu_simd_map_f32(int op, const float* x, float* y, int64_t n) {
#pragma omp simd
  for (int64_t i = 0; i < n; ++i) {
    float v = x[i];
    y[i] = op == 0 ? exp_bf(v) : op == 1 ? log_bf(v) : op == 4 ? erfc_bf(v) : op == 5 ? normcdf_bf(v) : normpdf_bf(v);
  }
}

// NT2's shared-memory policy: grain = L2 bytes / sizeof(float), parallel only
// above 8 grains, static blocks of one grain each.
void nt2ref_cpu_simd_bs_fwd_adj_f32(const float* S, const float* K, const float* Tm, const float* r,
                                    const float* sig, const float* w, float* price, float* dS, float* dK,
                                    float* dT, float* dr, float* dsig, int64_t n, int64_t l2_bytes) {
  int64_t grain = l2_bytes / (int64_t)sizeof(float);
  if (grain < 1) grain = 1;
  if (n <= 8 * grain) {
    bs_grain(S, K, Tm, r, sig, w, price, dS, dK, dT, dr, dsig, n);
    return;
  }
  const int64_t nblk = (n + grain - 1) / grain;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nblk; ++b) {
    const int64_t o = b * grain, c = std::min(grain, n - o);
    bs_grain(S + o, K + o, Tm + o, r + o, sig + o, w + o, price + o, dS + o, dK + o, dT + o, dr + o, dsig + o, c);
  }
}

}  // extern "C"
