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
// evaluates exp, sin and cos once per element and feeds four compensated
// accumulators — z and the three gradient sums — so the backward sweep
// costs no second read of X and stores nothing per element.  Gradients
// with cotangent gz_j:
//   dz/ds_j = gz_j sum e c X^2 / s_j^3,  dz/dk_j = -gz_j sum e sin(.) X,
//   dz/dp_j = -gz_j sum e sin(.)
#include <hip/hip_runtime.h>

#include "nt2h_common.hpp"
#include "../../include/nt2hip/math.hpp"

namespace nt2h {

template <bool GRAD>
struct ColAcc {
  Comp a[GRAD ? 4 : 1];
  __device__ __forceinline__ ColAcc() {
#pragma unroll
    for (int q = 0; q < (GRAD ? 4 : 1); ++q) a[q] = Comp{0.0f, 0.0f};
  }
  // four consecutive elements of the column: pairwise per group of four,
  // then one compensated add per accumulator
  __device__ __forceinline__ void add4(const float (&x)[4], float m2, float kj, float pj) {
    float t[GRAD ? 4 : 1][4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float xv = x[e];
      float ex = exp_neg_f<Fused>(-(xv * xv) * m2);
      float arg = __builtin_fmaf(kj, xv, pj);
      if (GRAD) {
        float sn, cs;
        sincos_f<Fused>(arg, &sn, &cs);
        float ec = ex * cs, es = ex * sn;
        t[0][e] = ec;
        t[1][e] = ec * (xv * xv);
        t[2][e] = es * xv;
        t[3][e] = es;
      } else {
        t[0][e] = ex * cos_f<Fused>(arg);
      }
    }
#pragma unroll
    for (int q = 0; q < (GRAD ? 4 : 1); ++q) a[q].add((t[q][0] + t[q][1]) + (t[q][2] + t[q][3]));
  }
  __device__ __forceinline__ void add1(float xv, float m2, float kj, float pj) {
    float ex = exp_neg_f<Fused>(-(xv * xv) * m2);
    float arg = __builtin_fmaf(kj, xv, pj);
    if (GRAD) {
      float sn, cs;
      sincos_f<Fused>(arg, &sn, &cs);
      float ec = ex * cs, es = ex * sn;
      a[0].add(ec);
      a[1].add(ec * (xv * xv));
      a[2].add(es * xv);
      a[3].add(es);
    } else {
      a[0].add(ex * cos_f<Fused>(arg));
    }
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
    // the next vector is loaded before the current one is evaluated, so
    // that its latency hides behind the arithmetic
    const int64_t nv = rows / 4;
    const f32x4* c4 = reinterpret_cast<const f32x4*>(col);
    // (unconditionally: the last iteration reloads the last vector, unused)
    int64_t i = threadIdx.x;
    if (i < nv) {
      f32x4 v = __builtin_nontemporal_load(c4 + i);
      for (; i < nv; i += kBlock) {
        const int64_t inext = (i + kBlock < nv) ? i + kBlock : nv - 1;
        f32x4 nx = __builtin_nontemporal_load(c4 + inext);
        float x[4] = {v.x, v.y, v.z, v.w};
        acc.add4(x, m2, kj, pj);
        v = nx;
      }
    }
    done = nv * 4;
  }
  for (int64_t i = done + threadIdx.x; i < rows; i += kBlock) acc.add1(col[i], m2, kj, pj);

  constexpr int NV = GRAD ? 4 : 1;
  __shared__ float lds[NV * 2 * kWaves];
  double tot[NV] = {};
  block_reduce_comp<NV>(acc.a, lds, tot);
  if (threadIdx.x == 0) {
    z[j] = (float)tot[0];
    if (GRAD) {
      float g = gz ? gz[j] : 1.0f;
      double s3 = (double)sj * (double)sj * (double)sj;
      dz_ds[j] = (float)((double)g * tot[1] / s3);
      dz_dk[j] = (float)(-(double)g * tot[2]);
      dz_dp[j] = (float)(-(double)g * tot[3]);
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
