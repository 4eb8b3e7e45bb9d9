// Black-Scholes forward + adjoint in one pass (the fused kernel whose
// options/s and HBM fraction are the headline measurement).
//
// Forward (per option i): the call price of the table expression
//   d1 = (log(S/K) + (r + sig^2/2) T) / (sig sqrt T),  d2 = d1 - sig sqrt T
//   price = S N(d1) - K e^{-rT} N(d2),  N(d) = normcdf(d) = erfc(-d/sqrt2)/2
// (NT2 functors: f_log.hpp, f_expo_reduction.hpp, erfc.hpp, normcdf.hpp;
// demo/blackscholes/nt2/blackandscholes_nt2.cpp:53-60 evaluates the same
// tree, with a drift-only d1, as four loop nests with three temporaries).
//
// Reverse sweep with cotangent w_i of price_i, hand-derived and simplified
// with S phi(d1) = K e^{-rT} phi(d2) (the d1/d2 terms cancel):
//   dS = w N(d1)                 dK = -w e^{-rT} N(d2)
//   dT = w (S phi(d1) sig / (2 sqrt T) + r K e^{-rT} N(d2))
//   dr = w T K e^{-rT} N(d2)     dsig = w S phi(d1) sqrt T
// Every forward intermediate is recomputed in registers in the same
// thread; nothing is stored between the sweeps.  phi(d1) reuses the
// exp(-d1^2/2) that the erfc of N(d1) evaluates anyway.
//
// Traffic: 6 input and 6 output streams, 48 B per float option; the two
// totals sum_i dr_i and sum_i dsig_i (dP/dr, dP/dsigma for scalar r and
// sigma) are compensated per lane (TwoSum), merged per wave and workgroup,
// and combined across workgroups in FP64 by the last workgroup.
#include <hip/hip_runtime.h>

#include <mutex>

#include "nt2h_common.hpp"
#include "../../include/nt2hip/math.hpp"

namespace nt2h {

constexpr float kInvSqrt2 = 0x1.6a09e6p-1f;   // sqrt(1/2)
constexpr float kInvSqrt2Pi = 0x1.988454p-2f; // 1/sqrt(2 pi)

struct BSAdjOut {
  float price, dS, dK, dT, dr, dsig;
};

// One option, float, Fused arithmetic.  Beyond the functors' own fusion:
// sqrt(T) is v_sqrt_f32 (1 ulp), 1/(sig sqrt T) is formed once (refined)
// and serves d1 and dT, and exp(-d2^2/2) is exp(-d1^2/2) * (S/K) * e^{rT}
// (the identity behind the adjoint's simplification), so the two erfc
// evaluations share one exp.
__device__ __forceinline__ BSAdjOut bs_fwd_adj_one(float S, float K, float Tm, float r, float sig, float w) {
  using P = Fused;
  float sqT = __builtin_amdgcn_sqrtf(Tm);
  float sst = sig * sqT;
  float rsst = P::rcp(sst);
  float sk = P::div(S, K);
  float lsk = log_f<P>(sk);
  float d1 = P::madd(P::madd(sig, sig * 0.5f, r), Tm, lsk) * rsst;
  float d2 = d1 - sst;
  float ert = exp_f<P>(-r * Tm);
  float kert = K * ert;
  // N(d1) and phi(d1) share exp(-d1^2/2), formed from d1 with d1^2 split
  // exactly (not from t1 = -d1/sqrt2, whose rounding would cost ~2 t1^2 ulp)
  float t1 = -kInvSqrt2 * d1;
  float e1 = exp_neg_half_square<P>(d1);
  float r1 = erfc_pos_e<P>(__builtin_fabsf(t1), e1);
  float n1 = 0.5f * ((t1 < 0.0f) ? 2.0f - r1 : r1);
  float t2 = -kInvSqrt2 * d2;
  float e2 = e1 * (sk * P::rcp(ert));
  // exp(-d2^2/2) <= 1.  Beyond (inf, or NaN from 0 * inf) a factor of the
  // identity left the float range: e^{-rT} underflowed to 0 (then kert is 0
  // and N(d2) only has to be finite), or S/K or (S/K) e^{rT} overflowed (then
  // K e^{-rT} < S 2^-128, and N(d2) enters the price below float
  // resolution).  0 gives N(d2) in {0, 1} there instead of a NaN.
  e2 = (e2 <= 2.0f) ? e2 : 0.0f;
  float r2 = erfc_pos_e<P>(__builtin_fabsf(t2), e2);
  float n2 = 0.5f * ((t2 < 0.0f) ? 2.0f - r2 : r2);
  float phi1 = e1 * kInvSqrt2Pi;
  float kn2 = kert * n2;
  float sphi = S * phi1;
  BSAdjOut o;
  o.price = P::madd(S, n1, -kn2);
  o.dS = w * n1;
  o.dK = -(w * (ert * n2));
  // S phi sig / (2 sqrt T) = S phi sig^2 / (2 sig sqrt T), and 0 wherever
  // S phi(d1) is (phi underflowed, S = 0), also when sig or sqrt T is 0
  // (rsst = inf) or sig^2 overflows
  const bool nophi = sphi == 0.0f;
  o.dT = w * P::madd(nophi ? 0.0f : sphi * (0.5f * sig * sig), nophi ? 0.0f : rsst, r * kn2);
  o.dr = w * (Tm * kn2);
  o.dsig = w * (sphi * sqT);
  return o;
}

// One option, double, Fused (FMA) arithmetic with IEEE division: the plain
// FP64 evaluation of the same tree, for the precision/speed comparison.
__device__ __forceinline__ void bs_fwd_adj_one(double S, double K, double Tm, double r, double sig, double w,
                                               double* o) {
  using P = Fused;
  double sqT = __builtin_sqrt(Tm);
  double sst = sig * sqT;
  double lsk = log_d<P>(S / K);
  double d1 = P::madd(P::madd(sig, sig * 0.5, r), Tm, lsk) / sst;
  double d2 = d1 - sst;
  double ert = exp_d<P>(-r * Tm);
  double kert = K * ert;
  double t1 = -0x1.6a09e667f3bcdp-1 * d1;
  double e1;
  double r1 = erfc_pos_d<P>(__builtin_fabs(t1), &e1);
  double n1 = 0.5 * ((t1 < 0.0) ? 2.0 - r1 : r1);
  double t2 = -0x1.6a09e667f3bcdp-1 * d2;
  double r2 = erfc_pos_d<P>(__builtin_fabs(t2));
  double n2 = 0.5 * ((t2 < 0.0) ? 2.0 - r2 : r2);
  double phi1 = e1 * 0x1.9884533d43651p-2;
  double kn2 = kert * n2;
  double sphi = S * phi1;
  o[0] = P::madd(S, n1, -kn2);
  o[1] = w * n1;
  o[2] = -(w * (ert * n2));
  o[3] = w * P::madd(sphi * sig, 0.5 / sqT, r * kn2);
  o[4] = w * (Tm * kn2);
  o[5] = w * (sphi * sqT);
}

struct BSAdjArgs {
  const float* in[6];
  float* out[6];
};
struct BSAdjArgsD {
  const double* in[6];
  double* out[6];
};

// Four options (one 16-byte vector per stream) of the float kernel.
__device__ __forceinline__ void bs_vec4(const BSAdjArgs& a, int64_t i, f32x4 (&o)[6]) {
  f32x4 v[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(a.in[k]) + i);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    BSAdjOut r = bs_fwd_adj_one(v[0][e], v[1][e], v[2][e], v[3][e], v[4][e], v[5][e]);
    o[0][e] = r.price;
    o[1][e] = r.dS;
    o[2][e] = r.dK;
    o[3][e] = r.dT;
    o[4][e] = r.dr;
    o[5][e] = r.dsig;
  }
}
__device__ __forceinline__ void bs_store4(const BSAdjArgs& a, int64_t i, const f32x4 (&o)[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) __builtin_nontemporal_store(o[k], reinterpret_cast<f32x4*>(a.out[k]) + i);
}

template <bool VEC, bool TOTALS>
__global__ __launch_bounds__(kBlock) void bs_fwd_adj_f32_kernel(BSAdjArgs a, int64_t n, Workspace ws,
                                                                double* totals) {
  Comp acc[2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t done = 0;
  if (VEC) {
    const int64_t nv = n / 4;
    for (int64_t i = gid; i < nv; i += stride) {
      f32x4 o[6];
      bs_vec4(a, i, o);
      if (TOTALS) {  // every option through the compensated add (free: HBM-bound)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          acc[0].add(o[4][e]);
          acc[1].add(o[5][e]);
        }
      }
      bs_store4(a, i, o);
    }
    done = nv * 4;
  }
  for (int64_t i = done + gid; i < n; i += stride) {
    BSAdjOut r = bs_fwd_adj_one(a.in[0][i], a.in[1][i], a.in[2][i], a.in[3][i], a.in[4][i], a.in[5][i]);
    a.out[0][i] = r.price;
    a.out[1][i] = r.dS;
    a.out[2][i] = r.dK;
    a.out[3][i] = r.dT;
    a.out[4][i] = r.dr;
    a.out[5][i] = r.dsig;
    if (TOTALS) {
      acc[0].add(r.dr);
      acc[1].add(r.dsig);
    }
  }
  if (TOTALS) {
    __shared__ float lds[2 * 2 * kWaves];
    double bsum[2] = {0.0, 0.0};
    block_reduce_comp<2>(acc, lds, bsum);
    double res[2] = {0.0, 0.0};
    if (grid_combine<2>(bsum, ws, res)) {
      totals[0] = res[0];
      totals[1] = res[1];
    }
  }
}

// totals of an empty portfolio
__global__ void zero_totals_kernel(double* t) {
  if (threadIdx.x < 2) t[threadIdx.x] = 0.0;
}
constexpr unsigned kAdj64Blocks = 8192;  // workgroups of the FP64 adjoint kernel, at most
template <bool TOTALS>
__global__ __launch_bounds__(kBlock, 4) void bs_fwd_adj_f64_kernel(BSAdjArgsD a, int64_t n, Workspace ws,
                                                                double* totals) {
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    double o[6];
    bs_fwd_adj_one(__builtin_nontemporal_load(a.in[0] + i), __builtin_nontemporal_load(a.in[1] + i),
                   __builtin_nontemporal_load(a.in[2] + i), __builtin_nontemporal_load(a.in[3] + i),
                   __builtin_nontemporal_load(a.in[4] + i), __builtin_nontemporal_load(a.in[5] + i), o);
#pragma unroll
    for (int k = 0; k < 6; ++k) __builtin_nontemporal_store(o[k], a.out[k] + i);
    if (TOTALS) {
      acc[0] += o[4];
      acc[1] += o[5];
    }
  }
  if (TOTALS) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      acc[0] += shfl_xor_d(acc[0], m);
      acc[1] += shfl_xor_d(acc[1], m);
    }
    __shared__ double lds[2 * kWaves];
    if ((threadIdx.x & 63) == 0) {
      lds[threadIdx.x >> 6] = acc[0];
      lds[kWaves + (threadIdx.x >> 6)] = acc[1];
    }
    __syncthreads();
    double bsum[2] = {0.0, 0.0};
    if (threadIdx.x == 0) {
      double t0 = lds[0], t1 = lds[kWaves];
      for (int w = 1; w < kWaves; ++w) {
        t0 += lds[w];
        t1 += lds[kWaves + w];
      }
      bsum[0] = t0;
      bsum[1] = t1;
    }
    double res[2] = {0.0, 0.0};
    if (grid_combine<2>(bsum, ws, res)) {
      totals[0] = res[0];
      totals[1] = res[1];
    }
  }
}

// ------------------------------------------------ allocation-free variant
// Scratch of nt2h_bs_fwd_adj_f32_noalloc: the reduction workspace of at most
// kNaBlocks workgroups (the grid of 2^24 options, one 16-byte vector per lane
// per stream) in one static device array, so that the call allocates nothing.
//   [0, 256 KiB)         block partials, 2 x kNaBlocks doubles
//   [256 KiB, +4 KiB)    group partials, 2 x kNaGroups doubles
//   then                 kNaGroups + 1 ticket counters, kCounterStride apart
//   then                 2 doubles (result slot, unused: totals go to device memory)
// Each device holds its own copy of the array until the process ends.  Host
// code never uses the array's name as an address: the per-device address
// comes from hipGetSymbolAddress (noalloc_setup).
constexpr int kNaBlocks = 16384;
constexpr int kNaGroups = kNaBlocks / kGroup;
constexpr int64_t kNaMaxOptions = (int64_t)kNaBlocks * kBlock * 4;  // 2^24
constexpr size_t kNaPartialsOff = 0;
constexpr size_t kNaGroupOff = kNaPartialsOff + 2 * sizeof(double) * kNaBlocks;
constexpr size_t kNaCountersOff = kNaGroupOff + 2 * sizeof(double) * kNaGroups;
constexpr size_t kNaCountersBytes = sizeof(unsigned) * (kNaGroups + 1) * kCounterStride;
constexpr size_t kNaResultOff = (kNaCountersOff + kNaCountersBytes + 255) / 256 * 256;
constexpr size_t kNaBytes = (kNaResultOff + 2 * sizeof(double) + 255) / 256 * 256;
static_assert(kNaCountersOff % 256 == 0, "counters start on a 256-byte boundary");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
void* g_na_ptr[kNaMaxDevices] = {};
}  // namespace

// Once per device, with that device current: the device address of
// `scratch` on it, and its ticket counters cleared (every later launch
// returns the counters it used to zero).  Called by nt2h_init and, for a
// device nt2h_init never saw, by the entry point's first call there.
int noalloc_setup(int dev, void** out) {
  if (dev < 0 || dev >= kNaMaxDevices) return NT2H_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    NT2H_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NT2H_HIP(hipMemset(static_cast<unsigned char*>(p) + kNaCountersOff, 0, kNaCountersBytes));
    NT2H_HIP(hipDeviceSynchronize());
    g_na_ptr[dev] = p;
  }
  if (out) *out = g_na_ptr[dev];
  return NT2H_OK;
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_bs_fwd_adj_f32(const float* S, const float* K, const float* T, const float* r, const float* sig,
                        const float* w, float* price, float* dS, float* dK, float* dT, float* dr, float* dsig,
                        double* totals, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  NT2H_TRY(check_ptrs(n, S, K, T, r, sig, w, price, dS, dK, dT, dr, dsig));
  Workspace ws{};
  bool host = false;
  double* dst = nullptr;
  if (totals) {
    NT2H_TRY(is_host_pointer(totals, &host, s));
    NT2H_TRY(get_workspace(s, &ws));
    dst = host ? ws.result : totals;
  }
  if (n == 0) {
    if (totals) {
      if (host) { totals[0] = 0.0; totals[1] = 0.0; return NT2H_OK; }
      // a kernel, not a copy from the host: asynchronous and capture-safe
      hipLaunchKernelGGL(zero_totals_kernel, dim3(1), dim3(64), 0, as_hip(s), totals);
      NT2H_HIP(hipGetLastError());
    }
    return NT2H_OK;
  }
  BSAdjArgs a{{S, K, T, r, sig, w}, {price, dS, dK, dT, dr, dsig}};
  bool vec = aligned16(S, K, T, r, sig, w, price, dS, dK, dT, dr, dsig);
  unsigned g = fold_blocks(vec ? (n + 3) / 4 : n);
  hipStream_t st = as_hip(s);
  if (vec) {
    if (totals) hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<true, true>), dim3(g), dim3(kBlock), 0, st, a, n, ws, dst);
    else hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<true, false>), dim3(g), dim3(kBlock), 0, st, a, n, ws, dst);
  } else {
    if (totals) hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<false, true>), dim3(g), dim3(kBlock), 0, st, a, n, ws, dst);
    else hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<false, false>), dim3(g), dim3(kBlock), 0, st, a, n, ws, dst);
  }
  NT2H_HIP(hipGetLastError());
  return totals ? deliver(s, totals, ws.result, 2, host) : NT2H_OK;
}

int nt2h_bs_fwd_adj_f64(const double* S, const double* K, const double* T, const double* r,
                        const double* sig, const double* w, double* price, double* dS, double* dK, double* dT,
                        double* dr, double* dsig, double* totals, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  NT2H_TRY(check_ptrs(n, S, K, T, r, sig, w, price, dS, dK, dT, dr, dsig));
  Workspace ws{};
  bool host = false;
  double* dst = nullptr;
  if (totals) {
    NT2H_TRY(is_host_pointer(totals, &host, s));
    NT2H_TRY(get_workspace(s, &ws));
    dst = host ? ws.result : totals;
  }
  if (n == 0) {
    if (totals) {
      if (host) { totals[0] = 0.0; totals[1] = 0.0; return NT2H_OK; }
      // a kernel, not a copy from the host: asynchronous and capture-safe
      hipLaunchKernelGGL(zero_totals_kernel, dim3(1), dim3(64), 0, as_hip(s), totals);
      NT2H_HIP(hipGetLastError());
    }
    return NT2H_OK;
  }
  BSAdjArgsD a{{S, K, T, r, sig, w}, {price, dS, dK, dT, dr, dsig}};
  // at most 8192 grid-striding workgroups (measured at 2^24 options: 441 us
  // with one option per lane on 65536 workgroups, 365 us on 16384 to 4096)
  unsigned g = fold_blocks(n);
  if (g > kAdj64Blocks) g = kAdj64Blocks;
  if (totals) hipLaunchKernelGGL((bs_fwd_adj_f64_kernel<true>), dim3(g), dim3(kBlock), 0, as_hip(s), a, n, ws, dst);
  else hipLaunchKernelGGL((bs_fwd_adj_f64_kernel<false>), dim3(g), dim3(kBlock), 0, as_hip(s), a, n, ws, dst);
  NT2H_HIP(hipGetLastError());
  return totals ? deliver(s, totals, ws.result, 2, host) : NT2H_OK;
}

int nt2h_bs_fwd_adj_f32_noalloc(const float* S, const float* K, const float* T, const float* r,
                                const float* sig, const float* w, float* price, float* dS, float* dK, float* dT,
                                float* dr, float* dsig, double* totals, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  NT2H_TRY(check_ptrs(n, S, K, T, r, sig, w, price, dS, dK, dT, dr, dsig));
  if (n > kNaMaxOptions) return NT2H_ERR_TOO_LARGE;
  if (totals) {
    bool host = false;
    NT2H_TRY(is_host_pointer(totals, &host, s));
    if (host) return NT2H_ERR_INVALID_ARG;  // the result is written by the kernel: device memory only
  }
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  void* base = nullptr;
  {
    // set-up is not capturable (a memset and a device synchronisation)
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (dev >= 0 && dev < kNaMaxDevices) base = g_na_ptr[dev];
  }
  if (!base) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(as_hip(s), &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
      return NT2H_ERR_CAPTURE;
    NT2H_TRY(noalloc_setup(dev, &base));
  }
  if (n == 0) {
    if (totals) {
      hipLaunchKernelGGL(zero_totals_kernel, dim3(1), dim3(64), 0, as_hip(s), totals);
      NT2H_HIP(hipGetLastError());
    }
    return NT2H_OK;
  }
  unsigned char* b = static_cast<unsigned char*>(base);
  Workspace ws{};
  ws.partials = reinterpret_cast<double*>(b + kNaPartialsOff);
  ws.gpartials = reinterpret_cast<double*>(b + kNaGroupOff);
  ws.counters = reinterpret_cast<unsigned*>(b + kNaCountersOff);
  ws.result = reinterpret_cast<double*>(b + kNaResultOff);
  ws.bpartials = nullptr;
  BSAdjArgs a{{S, K, T, r, sig, w}, {price, dS, dK, dT, dr, dsig}};
  bool vec = aligned16(S, K, T, r, sig, w, price, dS, dK, dT, dr, dsig);
  unsigned g = fold_blocks(vec ? (n + 3) / 4 : n);
  if (g > (unsigned)kNaBlocks) g = kNaBlocks;  // the scalar body grid-strides; the vector grid fits already
  hipStream_t st = as_hip(s);
  if (vec) {
    if (totals) hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<true, true>), dim3(g), dim3(kBlock), 0, st, a, n, ws, totals);
    else hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<true, false>), dim3(g), dim3(kBlock), 0, st, a, n, ws, totals);
  } else {
    if (totals) hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<false, true>), dim3(g), dim3(kBlock), 0, st, a, n, ws, totals);
    else hipLaunchKernelGGL((bs_fwd_adj_f32_kernel<false, false>), dim3(g), dim3(kBlock), 0, st, a, n, ws, totals);
  }
  NT2H_HIP(hipGetLastError());
  return NT2H_OK;
}

}  // extern "C"
