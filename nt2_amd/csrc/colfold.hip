// Column fold of a 2-D table expression with its adjoint, one pass over X.
//
//   z_j = sum_i exp(-X_ij^2 / (2 s_j^2)) * cos(k_j X_ij + p_j)
//
// is NT2's sum(exp(-sqr(X)/(2*sqr(sx(s)))) .* cos(...), 1): a reduction
// along dimension 1 that NT2 lowers to inner_fold over (rows, cols)
// (core/reduction/include/nt2/dsl/functions/reduction/run.hpp:48-81) and,
// multithreaded, to one worker per column group
// (arch/shared_memory/include/nt2/core/functions/shared_memory/inner_fold.hpp:53-57).
//
// Here one workgroup owns one column: its four waves stream the column
// with 128-bit loads (column-major, so the reads are contiguous), every lane
// evaluates exp, sin and cos once per element and feeds four FP64
// accumulators (a float pairwise sum per group of four elements, then one
// double add) — z and the three gradient sums — so the backward sweep
// costs no second read of X and stores nothing per element.  Gradients
// with cotangent gz_j:
//   dz/ds_j = gz_j sum e c X^2 / s_j^3,  dz/dk_j = -gz_j sum e sin(.) X,
//   dz/dp_j = -gz_j sum e sin(.)
// The kernel is VALU-bound, so the arithmetic per element is kept short:
// FP64 adds issue at the FP32 rate on this part (two instructions per
// accumulator and group instead of a float TwoSum), and the trigonometric
// argument k_j X + p_j (|.| < 2^6 pi for moderate k_j, X and p_j) takes a
// branch-free float Cody-Waite with a wave-uniform fallback to the general
// reduction.
#include <hip/hip_runtime.h>

#include "nt2h_common.hpp"
#include "../../include/nt2hip/math.hpp"

namespace nt2h {

// sin and cos (Fused policy) of a signed argument: |a| <= 2^6 pi by NT2's
// float 3-part Cody-Waite on a itself (k = rint(a 2/pi) < 2^7, so the first
// step is exact; the quadrant k mod 4 selects and signs the two
// polynomials); larger or non-finite arguments take sincos_f.  Equal to
// sincos_f except that sin(-0) is +0.
__device__ __forceinline__ void sincos_med(float a, float* s, float* c) {
  float k = __builtin_rintf(a * 0x1.45f306p-1f);
  float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, a);
  r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
  r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
  const uint32_t ik = (uint32_t)(int32_t)k;
  float z = r * r;
  float se = sin_eval_f<Fused>(z, r), ce = cos_eval_f<Fused>(z);
  bool sw = (ik & 1u) != 0;
  float sv = sw ? ce : se, cv = sw ? se : ce;
  sv = __uint_as_float(__float_as_uint(sv) ^ ((ik << 30) & 0x80000000u));        // quadrants 2, 3
  cv = __uint_as_float(__float_as_uint(cv) ^ (((ik + 1u) << 30) & 0x80000000u));  // quadrants 1, 2
  if (!(__builtin_fabsf(a) <= 0x1.921fb6p+7f)) sincos_f<Fused>(a, &sv, &cv);
  *s = sv;
  *c = cv;
}
__device__ __forceinline__ float cos_med(float a) {
  float k = __builtin_rintf(a * 0x1.45f306p-1f);
  float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, a);
  r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
  r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
  const uint32_t ik = (uint32_t)(int32_t)k;
  float z = r * r;
  float v = (ik & 1u) ? sin_eval_f<Fused>(z, r) : cos_eval_f<Fused>(z);
  v = __uint_as_float(__float_as_uint(v) ^ (((ik + 1u) << 30) & 0x80000000u));
  if (!(__builtin_fabsf(a) <= 0x1.921fb6p+7f)) v = cos_f<Fused>(a);
  return v;
}

// sin(pi f) and cos(pi f) for |f| <= 1/2: minimax polynomials in z = f^2
// (fitted for this kernel by nt2_amd/csrc/sincospi_coefficients.py; largest
// error 0.83 and 0.46 x 2^-24 before the float evaluation's own rounding).
__device__ __forceinline__ float sinpi_poly(float f, float z) {
  float p = __builtin_fmaf(z, 0x1.3c4b2ap-4f, -0x1.3232fap-1f);
  p = __builtin_fmaf(z, p, 0x1.46676ep+1f);
  p = __builtin_fmaf(z, p, -0x1.4abbb6p+2f);
  p = __builtin_fmaf(z, p, 0x1.921fb4p+1f);
  return f * p;
}
__device__ __forceinline__ float cospi_poly(float z) {
  float p = __builtin_fmaf(z, -0x1.901bd4p-6f, 0x1.e12f86p-3f);
  p = __builtin_fmaf(z, p, -0x1.55d0b8p+0f);
  p = __builtin_fmaf(z, p, 0x1.03c1e6p+2f);
  p = __builtin_fmaf(z, p, -0x1.3bd3ccp+2f);
  return __builtin_fmaf(z, p, 1.0f);
}

// Per-column constants.  FUSED: 1/(2 s^2), k, p for NT2's exp and sin/cos
// algorithms.  FAST: exp(-x^2/(2 s^2)) = 2^(x^2 ce) with ce = -log2(e)/(2 s^2),
// and k x + p = pi (kh x + ph) with kh = k/pi, ph = p/pi (each formed in
// double and rounded once).
struct ColConst {
  float m2, k, p;   // FUSED
  float ce, kh, ph; // FAST
};

template <bool GRAD, bool FAST>
struct ColAcc {
  static constexpr int NV = GRAD ? 4 : 1;
  double a[NV];
  __device__ __forceinline__ ColAcc() {
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] = 0.0;
  }
  // FAST: hardware exp2 (v_exp_f32) of x^2 ce, and sin/cos of pi y with
  // y = kh x + ph reduced to f = y - rint(y) in [-1/2, 1/2] (exact): the
  // parity of rint(y) flips the sign of both, returned as the sign mask m
  // (applied once, to the exp factor every term carries).  Rounding y costs
  // about as much as rounding the float argument k x + p itself (|.| 2^-24
  // of the argument); beyond |y| = 2^20 the general reduction of the float
  // argument runs (and m = 0).
  __device__ __forceinline__ void fast_sincos(float xv, const ColConst& cc, float* sn, float* cs, uint32_t* m) {
    float y = __builtin_fmaf(cc.kh, xv, cc.ph);
    float q = __builtin_rintf(y);
    float f = y - q;
    float z = f * f;
    float sv = sinpi_poly(f, z), cv = cospi_poly(z);
    uint32_t mv = ((uint32_t)(int32_t)q) << 31;
    if (!(__builtin_fabsf(y) <= 0x1.0p20f)) {
      sincos_f<Fused>(__builtin_fmaf(cc.k, xv, cc.p), &sv, &cv);
      mv = 0u;
    }
    *sn = sv;
    *cs = cv;
    *m = mv;
  }
  __device__ __forceinline__ float fast_cos(float xv, const ColConst& cc, uint32_t* m) {
    float y = __builtin_fmaf(cc.kh, xv, cc.ph);
    float q = __builtin_rintf(y);
    float f = y - q;
    float cv = cospi_poly(f * f);
    uint32_t mv = ((uint32_t)(int32_t)q) << 31;
    if (!(__builtin_fabsf(y) <= 0x1.0p20f)) {
      cv = cos_f<Fused>(__builtin_fmaf(cc.k, xv, cc.p));
      mv = 0u;
    }
    *m = mv;
    return cv;
  }
  // the NV terms of one element
  __device__ __forceinline__ void terms(float xv, const ColConst& cc, float (&t)[NV]) {
    float x2 = xv * xv;
    float ex = FAST ? __builtin_amdgcn_exp2f(x2 * cc.ce) : exp_neg_f<Fused>(-x2 * cc.m2);
    uint32_t m = 0u;
    if (GRAD) {
      float sn, cs;
      if (FAST) fast_sincos(xv, cc, &sn, &cs, &m);
      else sincos_med(__builtin_fmaf(cc.k, xv, cc.p), &sn, &cs);
      if (FAST) ex = __uint_as_float(__float_as_uint(ex) ^ m);
      float ec = ex * cs, es = ex * sn;
      t[0] = ec;
      // (ec x) x, not ec x^2: x^2 overflows for |x| > 2^64 while ex is 0
      t[NV > 1 ? 1 : 0] = (ec * xv) * xv;
      t[NV > 2 ? 2 : 0] = es * xv;
      t[NV > 3 ? 3 : 0] = es;
    } else {
      float cv = FAST ? fast_cos(xv, cc, &m) : cos_med(__builtin_fmaf(cc.k, xv, cc.p));
      if (FAST) ex = __uint_as_float(__float_as_uint(ex) ^ m);
      t[0] = ex * cv;
    }
  }
  // FAST: four elements folded into running float sums p[], one FMA per
  // term: z += ex c, ds += (ex x^2) c, dk += (ex x) s, dp += ex s, where ex
  // is exp(-x^2/(2 s^2)) carrying the sign of the reduced argument's parity.
  // One test of the four reduced arguments sends them to the general path
  // (a NaN y passes the test: its terms are NaN either way).
  __device__ __forceinline__ void fast_fold4(const f32x4& v, const ColConst& cc, float (&p)[NV]) {
    const float x[4] = {v.x, v.y, v.z, v.w};
    float y[4], sv[4], cv[4];
    uint32_t m[4];  // rint(y) (its parity flips both signs)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      y[e] = __builtin_fmaf(cc.kh, x[e], cc.ph);
      float q = __builtin_rintf(y[e]);
      float f = y[e] - q;
      float z = f * f;
      sv[e] = GRAD ? sinpi_poly(f, z) : 0.0f;
      cv[e] = cospi_poly(z);
      m[e] = (uint32_t)(int32_t)q;
    }
    float ymax = __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(y[0]), __builtin_fabsf(y[1])),
                                 __builtin_fmaxf(__builtin_fabsf(y[2]), __builtin_fabsf(y[3])));
    if (!(ymax <= 0x1.0p20f)) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!(__builtin_fabsf(y[e]) <= 0x1.0p20f)) {
          const float a = __builtin_fmaf(cc.k, x[e], cc.p);
          if (GRAD) sincos_f<Fused>(a, &sv[e], &cv[e]);
          else cv[e] = cos_f<Fused>(a);
          m[e] = 0u;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float x2 = x[e] * x[e];
      // flipping the sign bit is adding 2^31: one shift-and-add
      const float ex = __uint_as_float(__float_as_uint(__builtin_amdgcn_exp2f(x2 * cc.ce)) + (m[e] << 31));
      p[0] = __builtin_fmaf(ex, cv[e], p[0]);
      if (GRAD) {
        // ex x^2 as (ex x) x: x^2 overflows for |x| > 2^64 while ex is 0
        const float exx = ex * x[e];
        p[NV > 1 ? 1 : 0] = __builtin_fmaf(exx * x[e], cv[e], p[NV > 1 ? 1 : 0]);
        p[NV > 2 ? 2 : 0] = __builtin_fmaf(exx, sv[e], p[NV > 2 ? 2 : 0]);
        p[NV > 3 ? 3 : 0] = __builtin_fmaf(ex, sv[e], p[NV > 3 ? 3 : 0]);
      }
    }
  }
  // pairwise float sums of the terms of four consecutive elements
  __device__ __forceinline__ void sums4(const f32x4& v, const ColConst& cc, float (&p)[NV]) {
    if (FAST) {
#pragma unroll
      for (int q = 0; q < NV; ++q) p[q] = 0.0f;
      fast_fold4(v, cc, p);
      return;
    }
    float t[4][NV];
    terms(v.x, cc, t[0]);
    terms(v.y, cc, t[1]);
    terms(v.z, cc, t[2]);
    terms(v.w, cc, t[3]);
#pragma unroll
    for (int q = 0; q < NV; ++q) p[q] = (t[0][q] + t[1][q]) + (t[2][q] + t[3][q]);
  }
  // four consecutive elements of the column: a float pairwise sum, then one
  // double add per accumulator
  __device__ __forceinline__ void add4(const f32x4& v, const ColConst& cc) {
    float p[NV];
    sums4(v, cc, p);
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] += (double)p[q];
  }
  // two vectors (eight elements, two rounds apart): pairwise to eight, then
  // one double add per accumulator
  __device__ __forceinline__ void add8(const f32x4& v0, const f32x4& v1, const ColConst& cc) {
    float p[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) p[q] = 0.0f;
    fast_fold4(v0, cc, p);
    fast_fold4(v1, cc, p);
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] += (double)p[q];
  }
  __device__ __forceinline__ void add1(float xv, const ColConst& cc) {
    float t[NV];
    terms(xv, cc, t);
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] += (double)t[q];
  }
};

template <bool GRAD, bool VEC, bool FAST>
__global__ __launch_bounds__(kBlock) void colfold_kernel(const float* __restrict__ X, int64_t rows, int64_t ld,
                                                         const float* __restrict__ sv, const float* __restrict__ kv,
                                                         const float* __restrict__ pv, const float* __restrict__ gz,
                                                         float* z, float* dz_ds, float* dz_dk, float* dz_dp) {
  const int64_t j = blockIdx.x;
  const float sj = sv[j], kj = kv[j], pj = pv[j];
  ColConst cc;
  cc.m2 = Fused::rcp(2.0f * (sj * sj));  // 1 / (2 s^2)
  cc.k = kj;
  cc.p = pj;
  if (FAST) {
    const double s2 = (double)sj * (double)sj;
    cc.ce = (float)(-1.4426950408889634 / (2.0 * s2));
    cc.kh = (float)((double)kj * 0.3183098861837907);
    cc.ph = (float)((double)pj * 0.3183098861837907);
  }
  const float* col = X + j * ld;
  ColAcc<GRAD, FAST> acc;
  int64_t done = 0;
  if (VEC) {
    // full rounds (every lane holds a vector) with a workgroup-uniform trip
    // count; the next vector is loaded before the current one is
    // evaluated, so that its latency hides behind the arithmetic; then one
    // partial round
    const int64_t nv = rows / 4;
    const int64_t full = nv / kBlock;
    const f32x4* c4 = reinterpret_cast<const f32x4*>(col) + threadIdx.x;
    int64_t r = 0;
    if (FAST && full >= 2) {
      // two rounds per step: eight elements per double add
      f32x4 v0 = __builtin_nontemporal_load(c4), v1 = __builtin_nontemporal_load(c4 + kBlock);
      for (r = 2; r + 1 < full; r += 2) {
        f32x4 n0 = __builtin_nontemporal_load(c4 + r * kBlock);
        f32x4 n1 = __builtin_nontemporal_load(c4 + (r + 1) * kBlock);
        acc.add8(v0, v1, cc);
        v0 = n0;
        v1 = n1;
      }
      acc.add8(v0, v1, cc);
    } else if (full > 0) {
      f32x4 v = __builtin_nontemporal_load(c4);
      for (r = 1; r < full; ++r) {
        f32x4 nx = __builtin_nontemporal_load(c4 + r * kBlock);
        acc.add4(v, cc);
        v = nx;
      }
      acc.add4(v, cc);
    }
    for (; r < full; ++r) acc.add4(__builtin_nontemporal_load(c4 + r * kBlock), cc);  // odd last round
    if (full * kBlock + threadIdx.x < nv) acc.add4(__builtin_nontemporal_load(c4 + full * kBlock), cc);
    done = nv * 4;
  }
  for (int64_t i = done + threadIdx.x; i < rows; i += kBlock) acc.add1(col[i], cc);

  // fixed-order workgroup reduction in double
  constexpr int NV = ColAcc<GRAD, FAST>::NV;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) acc.a[q] += __shfl_xor(acc.a[q], m, 64);
  }
  __shared__ double lds[NV * kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[q * kWaves + wave] = acc.a[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      tot[q] = lds[q * kWaves];
      for (int w = 1; w < kWaves; ++w) tot[q] += lds[q * kWaves + w];
    }
    z[j] = (float)tot[0];
    if (GRAD) {
      float g = gz ? gz[j] : 1.0f;
      double s3 = (double)sj * (double)sj * (double)sj;
      dz_ds[j] = (float)((double)g * tot[NV > 1 ? 1 : 0] / s3);
      dz_dk[j] = (float)(-(double)g * tot[NV > 2 ? 2 : 0]);
      dz_dp[j] = (float)(-(double)g * tot[NV > 3 ? 3 : 0]);
    }
  }
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_colfold_gc_f32_prec(int precision, const float* X, int64_t rows, int64_t cols, int64_t ld,
                             const float* s_, const float* k_, const float* p_, const float* gz, float* z,
                             float* dz_ds, float* dz_dk, float* dz_dp, nt2h_stream_t s) {
  if (precision != NT2H_FUSED && precision != NT2H_FAST) return NT2H_ERR_INVALID_ARG;
  if (rows < 0 || cols < 0) return NT2H_ERR_INVALID_ARG;
  if (ld < rows || ld < 1) return NT2H_ERR_INVALID_ARG;
  if (cols == 0) return NT2H_OK;
  if (!s_ || !k_ || !p_ || !z) return NT2H_ERR_NULL_POINTER;
  if (rows > 0 && !X) return NT2H_ERR_NULL_POINTER;
  const int ng = (dz_ds != nullptr) + (dz_dk != nullptr) + (dz_dp != nullptr);
  if (ng != 0 && ng != 3) return NT2H_ERR_NULL_POINTER;  // all three gradients or none
  if (cols > 0x7fffffffLL) return NT2H_ERR_TOO_LARGE;
  const bool grad = (ng == 3);
  const bool vec = X && aligned16(X) && (ld % 4 == 0);
  const bool fast = precision == NT2H_FAST;
  dim3 g((unsigned)cols), b(kBlock);
  hipStream_t st = as_hip(s);
#define NT2H_COLFOLD_LAUNCH(G, V, F)                                                                              \
  hipLaunchKernelGGL((colfold_kernel<G, V, F>), g, b, 0, st, X, rows, ld, s_, k_, p_, gz, z, dz_ds, dz_dk, dz_dp)
  if (fast) {
    if (grad) { if (vec) NT2H_COLFOLD_LAUNCH(true, true, true); else NT2H_COLFOLD_LAUNCH(true, false, true); }
    else { if (vec) NT2H_COLFOLD_LAUNCH(false, true, true); else NT2H_COLFOLD_LAUNCH(false, false, true); }
  } else {
    if (grad) { if (vec) NT2H_COLFOLD_LAUNCH(true, true, false); else NT2H_COLFOLD_LAUNCH(true, false, false); }
    else { if (vec) NT2H_COLFOLD_LAUNCH(false, true, false); else NT2H_COLFOLD_LAUNCH(false, false, false); }
  }
#undef NT2H_COLFOLD_LAUNCH
  NT2H_HIP(hipGetLastError());
  return NT2H_OK;
}

int nt2h_colfold_gc_f32(const float* X, int64_t rows, int64_t cols, int64_t ld, const float* s_,
                        const float* k_, const float* p_, const float* gz, float* z, float* dz_ds,
                        float* dz_dk, float* dz_dp, nt2h_stream_t s) {
  return nt2h_colfold_gc_f32_prec(NT2H_FAST, X, rows, cols, ld, s_, k_, p_, gz, z, dz_ds, dz_dk, dz_dp, s);
}

}  // extern "C"
