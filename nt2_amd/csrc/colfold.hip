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

template <bool GRAD>
struct ColAcc {
  static constexpr int NV = GRAD ? 4 : 1;
  double a[NV];
  __device__ __forceinline__ ColAcc() {
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] = 0.0;
  }
  // the NV terms of one element
  __device__ __forceinline__ void terms(float xv, float m2, float kj, float pj, float (&t)[NV]) {
    float x2 = xv * xv;
    float ex = exp_neg_f<Fused>(-x2 * m2);
    float arg = __builtin_fmaf(kj, xv, pj);
    if (GRAD) {
      float sn, cs;
      sincos_med(arg, &sn, &cs);
      float ec = ex * cs, es = ex * sn;
      t[0] = ec;
      t[NV > 1 ? 1 : 0] = ec * x2;
      t[NV > 2 ? 2 : 0] = es * xv;
      t[NV > 3 ? 3 : 0] = es;
    } else {
      t[0] = ex * cos_med(arg);
    }
  }
  // four consecutive elements of the column: a float pairwise sum, then one
  // double add per accumulator
  __device__ __forceinline__ void add4(const f32x4& v, float m2, float kj, float pj) {
    float t[4][NV];
    terms(v.x, m2, kj, pj, t[0]);
    terms(v.y, m2, kj, pj, t[1]);
    terms(v.z, m2, kj, pj, t[2]);
    terms(v.w, m2, kj, pj, t[3]);
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] += (double)((t[0][q] + t[1][q]) + (t[2][q] + t[3][q]));
  }
  __device__ __forceinline__ void add1(float xv, float m2, float kj, float pj) {
    float t[NV];
    terms(xv, m2, kj, pj, t);
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q] += (double)t[q];
  }
};

template <bool GRAD, bool VEC>
__global__ __launch_bounds__(kBlock) void colfold_kernel(const float* __restrict__ X, int64_t rows, int64_t ld,
                                                         const float* __restrict__ sv, const float* __restrict__ kv,
                                                         const float* __restrict__ pv, const float* __restrict__ gz,
                                                         float* z, float* dz_ds, float* dz_dk, float* dz_dp) {
  const int64_t j = blockIdx.x;
  const float sj = sv[j], kj = kv[j], pj = pv[j];
  const float m2 = Fused::rcp(2.0f * (sj * sj));  // 1 / (2 s^2)
  const float* col = X + j * ld;
  ColAcc<GRAD> acc;
  int64_t done = 0;
  if (VEC) {
    // full rounds (every lane holds a vector) with a workgroup-uniform trip
    // count; the next vector is loaded before the current one is
    // evaluated, so that its latency hides behind the arithmetic; then one
    // partial round
    const int64_t nv = rows / 4;
    const int64_t full = nv / kBlock;
    const f32x4* c4 = reinterpret_cast<const f32x4*>(col) + threadIdx.x;
    if (full > 0) {
      f32x4 v = __builtin_nontemporal_load(c4);
      for (int64_t r = 1; r < full; ++r) {
        f32x4 nx = __builtin_nontemporal_load(c4 + r * kBlock);
        acc.add4(v, m2, kj, pj);
        v = nx;
      }
      acc.add4(v, m2, kj, pj);
    }
    if (full * kBlock + threadIdx.x < nv) acc.add4(__builtin_nontemporal_load(c4 + full * kBlock), m2, kj, pj);
    done = nv * 4;
  }
  for (int64_t i = done + threadIdx.x; i < rows; i += kBlock) acc.add1(col[i], m2, kj, pj);

  // fixed-order workgroup reduction in double
  constexpr int NV = ColAcc<GRAD>::NV;
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

int nt2h_colfold_gc_f32(const float* X, int64_t rows, int64_t cols, int64_t ld, const float* s_,
                        const float* k_, const float* p_, const float* gz, float* z, float* dz_ds,
                        float* dz_dk, float* dz_dp, nt2h_stream_t s) {
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
  dim3 g((unsigned)cols), b(kBlock);
  hipStream_t st = as_hip(s);
  if (grad) {
    if (vec) hipLaunchKernelGGL((colfold_kernel<true, true>), g, b, 0, st, X, rows, ld, s_, k_, p_, gz, z, dz_ds, dz_dk, dz_dp);
    else hipLaunchKernelGGL((colfold_kernel<true, false>), g, b, 0, st, X, rows, ld, s_, k_, p_, gz, z, dz_ds, dz_dk, dz_dp);
  } else {
    if (vec) hipLaunchKernelGGL((colfold_kernel<false, true>), g, b, 0, st, X, rows, ld, s_, k_, p_, gz, z, dz_ds, dz_dk, dz_dp);
    else hipLaunchKernelGGL((colfold_kernel<false, false>), g, b, 0, st, X, rows, ld, s_, k_, p_, gz, z, dz_ds, dz_dk, dz_dp);
  }
  NT2H_HIP(hipGetLastError());
  return NT2H_OK;
}

}  // extern "C"
