// Whole-table folds (NT2 inner_fold_ with plus / Zero: sum, globalsum, dot)
// as one single-pass HIP kernel each.
//
// Each lane accumulates a grid-strided slice with 128-bit loads (two
// vectors per lane in flight, up to 64K workgroups so that the hardware
// keeps launching waves while others wait on memory), the four waves of a
// workgroup combine through shuffles and LDS, and the workgroup partials
// are combined in FP64 in a fixed order: by a two-level ticket (groups of
// 64 workgroups, then the groups) in the same launch for small grids, and
// for large grids by a second, small launch (fold_partials_kernel) — no
// float atomics, the same bits on every run for a given size.
// NT2's counterpart: scalar fold_step (core/reduction/include/nt2/core/
// functions/details/fold_step.hpp:31-43), SIMD inner_fold (simd/
// inner_fold.hpp:44-67) and the OpenMP block fold merged under omp
// critical (arch/openmp/include/nt2/sdk/openmp/spawner/fold.hpp:46-79).
#include <hip/hip_runtime.h>

#include "nt2h_common.hpp"

namespace nt2h {

// Per-lane accumulator for one summation mode.
template <int MODE> struct LaneAcc;

template <> struct LaneAcc<NT2H_SUM_NAIVE> {
  float s = 0.0f;
  __device__ __forceinline__ void add4(float a, float b, float c, float d) { s += a; s += b; s += c; s += d; }
  __device__ __forceinline__ void add1(float a) { s += a; }
  __device__ __forceinline__ void add_err(float) {}
  __device__ __forceinline__ Comp finish() const { return Comp{s, 0.0f}; }
};

template <> struct LaneAcc<NT2H_SUM_KAHAN> {
  Comp c{0.0f, 0.0f};
  __device__ __forceinline__ void add4(float a, float b, float x, float d) { c.add(a); c.add(b); c.add(x); c.add(d); }
  __device__ __forceinline__ void add1(float a) { c.add(a); }
  // exact rounding error of a product, folded into the compensation (Dot2)
  __device__ __forceinline__ void add_err(float e) { c.c += e; }
  __device__ __forceinline__ Comp finish() const { return c; }
};

// Binary-counter cascade: level k holds the pairwise sum of 2^k groups of
// four; a new group carries up like an increment.  Static indices only, so
// the levels stay in VGPRs.
template <> struct LaneAcc<NT2H_SUM_PAIRWISE> {
  static constexpr int L = 20;
  float lev[L];
  uint32_t cnt = 0;
  float tail = 0.0f;
  __device__ __forceinline__ LaneAcc() {
#pragma unroll
    for (int k = 0; k < L; ++k) lev[k] = 0.0f;
  }
  __device__ __forceinline__ void push(float v) {
    bool carry = true;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      bool bit = (cnt >> k) & 1u;
      float merged = lev[k] + v;
      if (carry) {
        if (bit) v = merged;
        else { lev[k] = v; carry = false; }
      }
    }
    ++cnt;
  }
  __device__ __forceinline__ void add4(float a, float b, float c, float d) { push((a + b) + (c + d)); }
  __device__ __forceinline__ void add1(float a) { tail += a; }
  __device__ __forceinline__ void add_err(float) {}
  __device__ __forceinline__ Comp finish() const {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < L; ++k)
      if ((cnt >> k) & 1u) s = s + lev[k];
    return Comp{s + tail, 0.0f};
  }
};

// One 16-byte step of a lane: four elements (DOT: four exact products
// whose rounding errors enter the compensation).
template <int MODE, bool DOT>
__device__ __forceinline__ void fold_step4(LaneAcc<MODE>& acc, const f32x4& a, const f32x4& b) {
  if (DOT) {
    float p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z, p3 = a.w * b.w;
    acc.add4(p0, p1, p2, p3);
    if (MODE == NT2H_SUM_KAHAN) {
      float e = (__builtin_fmaf(a.x, b.x, -p0) + __builtin_fmaf(a.y, b.y, -p1)) +
                (__builtin_fmaf(a.z, b.z, -p2) + __builtin_fmaf(a.w, b.w, -p3));
      acc.add_err(e);
    }
  } else {
    acc.add4(a.x, a.y, a.z, a.w);
  }
}

// TWO: store the block partial for fold_partials_kernel instead of taking
// the ticket (large grids).
template <int MODE, bool DOT, bool VEC, bool TWO>
__global__ __launch_bounds__(kBlock) void fold_f32_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                          int64_t n, Workspace ws, double* out) {
  LaneAcc<MODE> acc;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t done = 0;
  if (VEC) {
    const int64_t nv = n / 4;
    const f32x4* x4 = reinterpret_cast<const f32x4*>(x);
    const f32x4* y4 = reinterpret_cast<const f32x4*>(y);
    int64_t i = gid;
    // two steps at a time, both loads issued before either is folded
    for (; i + stride < nv; i += 2 * stride) {
      f32x4 a0 = __builtin_nontemporal_load(x4 + i), a1 = __builtin_nontemporal_load(x4 + i + stride);
      f32x4 b0 = a0, b1 = a1;
      if (DOT) {
        b0 = __builtin_nontemporal_load(y4 + i);
        b1 = __builtin_nontemporal_load(y4 + i + stride);
      }
      fold_step4<MODE, DOT>(acc, a0, b0);
      fold_step4<MODE, DOT>(acc, a1, b1);
    }
    for (; i < nv; i += stride) {
      f32x4 a = __builtin_nontemporal_load(x4 + i);
      f32x4 b = DOT ? __builtin_nontemporal_load(y4 + i) : a;
      fold_step4<MODE, DOT>(acc, a, b);
    }
    done = nv * 4;
  }
  for (int64_t i = done + gid; i < n; i += stride) {
    float a = x[i];
    if (DOT) {
      float b = y[i];
      float p = a * b;
      acc.add1(p);
      if (MODE == NT2H_SUM_KAHAN) acc.add_err(__builtin_fmaf(a, b, -p));
    } else {
      acc.add1(a);
    }
  }
  __shared__ float lds[2 * kWaves];
  Comp c[1] = {acc.finish()};
  double bsum[1] = {0.0};
  block_reduce_comp<1>(c, lds, bsum);
  if (TWO) {
    if (threadIdx.x == 0) ws.bpartials[blockIdx.x] = bsum[0];
    return;
  }
  double res[1] = {0.0};
  if (grid_combine<1>(bsum, ws, res)) out[0] = res[0];
}

// Double tables: plain FP64 lane sums (16-byte loads when aligned, two
// vectors in flight), the same block tree and combine as the float kernel.
template <bool DOT, bool VEC, bool TWO>
__global__ __launch_bounds__(kBlock) void fold_f64_kernel(const double* __restrict__ x,
                                                          const double* __restrict__ y, int64_t n,
                                                          Workspace ws, double* out) {
  double s = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t done = 0;
  if (VEC) {
    const int64_t nv = n / 2;
    const f64x2* x2 = reinterpret_cast<const f64x2*>(x);
    const f64x2* y2 = reinterpret_cast<const f64x2*>(y);
    int64_t i = gid;
    for (; i + stride < nv; i += 2 * stride) {  // two vectors in flight
      f64x2 a0 = __builtin_nontemporal_load(x2 + i), a1 = __builtin_nontemporal_load(x2 + i + stride);
      if (DOT) {
        f64x2 b0 = __builtin_nontemporal_load(y2 + i), b1 = __builtin_nontemporal_load(y2 + i + stride);
        s += a0.x * b0.x;
        s += a0.y * b0.y;
        s += a1.x * b1.x;
        s += a1.y * b1.y;
      } else {
        s += a0.x;
        s += a0.y;
        s += a1.x;
        s += a1.y;
      }
    }
    for (; i < nv; i += stride) {
      f64x2 a = __builtin_nontemporal_load(x2 + i);
      if (DOT) {
        f64x2 b = __builtin_nontemporal_load(y2 + i);
        s += a.x * b.x;
        s += a.y * b.y;
      } else {
        s += a.x;
        s += a.y;
      }
    }
    done = nv * 2;
  }
  for (int64_t i = done + gid; i < n; i += stride) {
    double a = __builtin_nontemporal_load(x + i);
    if (DOT) s += a * __builtin_nontemporal_load(y + i);
    else s += a;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += shfl_xor_d(s, m);
  __shared__ double lds[kWaves];
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  double bsum[1] = {0.0};
  if (threadIdx.x == 0) {
    double t = lds[0];
    for (int w = 1; w < kWaves; ++w) t += lds[w];
    bsum[0] = t;
  }
  if (TWO) {
    if (threadIdx.x == 0) ws.bpartials[blockIdx.x] = bsum[0];
    return;
  }
  double res[1] = {0.0};
  if (grid_combine<1>(bsum, ws, res)) out[0] = res[0];
}

// Grids of at least kTwoLaunchBlocks workgroups finish in a second launch
// (fold_partials_kernel); smaller ones, where a second launch would cost
// more than the tickets, combine in the same launch.
constexpr unsigned kTwoLaunchBlocks = 4096;

template <int MODE, bool DOT>
static void launch_f32(bool vec, unsigned g, const float* x, const float* y, int64_t n, const Workspace& ws,
                       double* dst, hipStream_t st) {
  if (g >= kTwoLaunchBlocks) {
    if (vec)
      hipLaunchKernelGGL((fold_f32_kernel<MODE, DOT, true, true>), dim3(g), dim3(kBlock), 0, st, x, y, n, ws, dst);
    else
      hipLaunchKernelGGL((fold_f32_kernel<MODE, DOT, false, true>), dim3(g), dim3(kBlock), 0, st, x, y, n, ws, dst);
    hipLaunchKernelGGL((fold_partials_kernel<1>), dim3(partials_blocks(g)), dim3(kBlock), 0, st, (int)g, ws, dst);
    return;
  }
  if (vec)
    hipLaunchKernelGGL((fold_f32_kernel<MODE, DOT, true, false>), dim3(g), dim3(kBlock), 0, st, x, y, n, ws, dst);
  else
    hipLaunchKernelGGL((fold_f32_kernel<MODE, DOT, false, false>), dim3(g), dim3(kBlock), 0, st, x, y, n, ws, dst);
}

static int fold_f32(const float* x, const float* y, int64_t n, int mode, double* out, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  if (mode != NT2H_SUM_NAIVE && mode != NT2H_SUM_KAHAN && mode != NT2H_SUM_PAIRWISE) return NT2H_ERR_INVALID_ARG;
  if (!out) return NT2H_ERR_NULL_POINTER;
  if (y) NT2H_TRY(check_ptrs(n, x, y));
  else NT2H_TRY(check_ptrs(n, x));
  bool host = false;
  NT2H_TRY(is_host_pointer(out, &host, s));
  Workspace ws;
  NT2H_TRY(get_workspace(s, &ws));
  double* dst = host ? ws.result : out;
  bool vec = y ? aligned16(x, y) : aligned16(x);
  unsigned g = fold_blocks(vec ? (n + 3) / 4 : n);
  hipStream_t st = as_hip(s);
  const bool dot = (y != nullptr);
  switch (mode) {
    case NT2H_SUM_NAIVE:
      dot ? launch_f32<NT2H_SUM_NAIVE, true>(vec, g, x, y, n, ws, dst, st)
          : launch_f32<NT2H_SUM_NAIVE, false>(vec, g, x, y, n, ws, dst, st);
      break;
    case NT2H_SUM_KAHAN:
      dot ? launch_f32<NT2H_SUM_KAHAN, true>(vec, g, x, y, n, ws, dst, st)
          : launch_f32<NT2H_SUM_KAHAN, false>(vec, g, x, y, n, ws, dst, st);
      break;
    default:
      dot ? launch_f32<NT2H_SUM_PAIRWISE, true>(vec, g, x, y, n, ws, dst, st)
          : launch_f32<NT2H_SUM_PAIRWISE, false>(vec, g, x, y, n, ws, dst, st);
  }
  NT2H_HIP(hipGetLastError());
  return deliver(s, out, ws.result, 1, host);
}

static int fold_f64(const double* x, const double* y, int64_t n, double* out, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  if (!out) return NT2H_ERR_NULL_POINTER;
  if (y) NT2H_TRY(check_ptrs(n, x, y));
  else NT2H_TRY(check_ptrs(n, x));
  bool host = false;
  NT2H_TRY(is_host_pointer(out, &host, s));
  Workspace ws;
  NT2H_TRY(get_workspace(s, &ws));
  double* dst = host ? ws.result : out;
  // 16-byte loads for sums; the dot reads 8-byte elements of its two
  // streams (measured 10% faster for it than 16-byte pairs)
  const bool vec = !y && aligned16(x);
  const unsigned g = fold_blocks(vec ? (n + 1) / 2 : n);
  const bool two = g >= kTwoLaunchBlocks;
  hipStream_t st = as_hip(s);
#define NT2H_F64_FOLD(DOT, VEC, TWO) \
  hipLaunchKernelGGL((fold_f64_kernel<DOT, VEC, TWO>), dim3(g), dim3(kBlock), 0, st, x, y, n, ws, dst)
  if (y) {
    if (two) NT2H_F64_FOLD(true, false, true);
    else NT2H_F64_FOLD(true, false, false);
  } else {
    if (vec) { if (two) NT2H_F64_FOLD(false, true, true); else NT2H_F64_FOLD(false, true, false); }
    else { if (two) NT2H_F64_FOLD(false, false, true); else NT2H_F64_FOLD(false, false, false); }
  }
#undef NT2H_F64_FOLD
  if (two) hipLaunchKernelGGL((fold_partials_kernel<1>), dim3(partials_blocks(g)), dim3(kBlock), 0, st, (int)g, ws, dst);
  NT2H_HIP(hipGetLastError());
  return deliver(s, out, ws.result, 1, host);
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_sum_f32(const float* x, int64_t n, int mode, double* out, nt2h_stream_t s) {
  return fold_f32(x, nullptr, n, mode, out, s);
}
int nt2h_dot_f32(const float* x, const float* y, int64_t n, int mode, double* out, nt2h_stream_t s) {
  if (n > 0 && !y) return NT2H_ERR_NULL_POINTER;
  if (n == 0 && !y) y = x;  // empty dot: any non-null stand-in
  return fold_f32(x, y ? y : x, n, mode, out, s);
}
int nt2h_sum_f64(const double* x, int64_t n, double* out, nt2h_stream_t s) { return fold_f64(x, nullptr, n, out, s); }
int nt2h_dot_f64(const double* x, const double* y, int64_t n, double* out, nt2h_stream_t s) {
  if (n > 0 && !y) return NT2H_ERR_NULL_POINTER;
  return fold_f64(x, y ? y : x, n, out, s);
}

}  // extern "C"
