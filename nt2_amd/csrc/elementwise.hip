// Elementwise loop nests (NT2 transform_) as single HIP kernels on gfx950.
//
// One launch evaluates one fused loop nest: every input stream is read once
// with 128-bit loads (f32x4 / double2 per lane), the expression tree is
// evaluated in registers, and every output stream is written once.  A
// grid-stride loop over 16-byte vectors keeps all 256 CUs busy; a scalar
// tail (or a scalar kernel when a pointer is not 16-byte aligned) covers the
// rest.  NT2's counterpart is the transform loop nest (core/functions/
// transform.hpp:55-69, scalar body expr/transform.hpp:51-59, SIMD body
// simd/transform.hpp:40-57, OpenMP spawner openmp/spawner/transform.hpp:36-73).
#include <hip/hip_runtime.h>

#include "nt2h_common.hpp"
#include "../../include/nt2hip/math.hpp"

namespace nt2h {

template <class V, class T> __device__ __forceinline__ T vget(const V& v, int i) {
  return reinterpret_cast<const T*>(&v)[i];
}
template <class V, class T> __device__ __forceinline__ void vset(V& v, int i, T x) {
  reinterpret_cast<T*>(&v)[i] = x;
}

template <class T, int NI, int NO> struct Streams {
  const T* in[NI];
  T* out[NO];
};

// out_k[i] = F(in_0[i], ..., in_{NI-1}[i])_k for i in [0, n).  VEC16 selects
// the 16-byte vector body (all pointers 16-byte aligned) or the scalar one.
template <class T, int NI, int NO, bool VEC16, class F>
__global__ __launch_bounds__(kBlock) void ew_kernel(Streams<T, NI, NO> st, int64_t n, F f) {
  using VT = typename Vec16<T>::type;
  constexpr int W = VEC16 ? Vec16<T>::n : 1;
  const int64_t nv = n / W;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (VEC16) {
    for (; i < nv; i += stride) {
      VT a[NI];
#pragma unroll
      for (int k = 0; k < NI; ++k) a[k] = __builtin_nontemporal_load(reinterpret_cast<const VT*>(st.in[k]) + i);
      VT o[NO];
#pragma unroll
      for (int e = 0; e < W; ++e) {
        T xin[NI], xout[NO];
#pragma unroll
        for (int k = 0; k < NI; ++k) xin[k] = vget<VT, T>(a[k], e);
        f(xin, xout);
#pragma unroll
        for (int k = 0; k < NO; ++k) vset<VT, T>(o[k], e, xout[k]);
      }
#pragma unroll
      for (int k = 0; k < NO; ++k) __builtin_nontemporal_store(o[k], reinterpret_cast<VT*>(st.out[k]) + i);
    }
    // tail elements [nv*W, n)
    int64_t t = nv * W + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) {
      T xin[NI], xout[NO];
#pragma unroll
      for (int k = 0; k < NI; ++k) xin[k] = st.in[k][t];
      f(xin, xout);
#pragma unroll
      for (int k = 0; k < NO; ++k) st.out[k][t] = xout[k];
    }
  } else {
    for (; i < n; i += stride) {
      T xin[NI], xout[NO];
#pragma unroll
      for (int k = 0; k < NI; ++k) xin[k] = st.in[k][i];
      f(xin, xout);
#pragma unroll
      for (int k = 0; k < NO; ++k) st.out[k][i] = xout[k];
    }
  }
}

// Grid: one 16-byte vector per lane (measured 8-20% faster at 2^28
// elements than 65536 grid-striding workgroups for the light functors and
// saxpy); the erfc family, whose long bodies favour fewer, longer-lived
// workgroups, passes max_blocks = 65536.
constexpr int64_t kEwOnePerLane = int64_t(1) << 30;

template <class T, int NI, int NO, class F>
static int launch_ew(const Streams<T, NI, NO>& st, int64_t n, F f, nt2h_stream_t s,
                     int64_t max_blocks = kEwOnePerLane) {
  if (n == 0) return NT2H_OK;
  bool al = true;
  for (int k = 0; k < NI; ++k) al = al && aligned16(st.in[k]);
  for (int k = 0; k < NO; ++k) al = al && aligned16(st.out[k]);
  if (al) {
    unsigned g = stream_grid(n, Vec16<T>::n, max_blocks);
    hipLaunchKernelGGL((ew_kernel<T, NI, NO, true, F>), dim3(g), dim3(kBlock), 0, as_hip(s), st, n, f);
  } else {
    unsigned g = stream_grid(n, 1, max_blocks);
    hipLaunchKernelGGL((ew_kernel<T, NI, NO, false, F>), dim3(g), dim3(kBlock), 0, as_hip(s), st, n, f);
  }
  NT2H_HIP(hipGetLastError());
  return NT2H_OK;
}

// ------------------------------------------------------------- saxpy
// demo/saxpy/nt2/saxpy_nt2.cpp:41: Y = alpha*X + Y with two roundings (this
// file is compiled with -ffp-contract=off, so no FMA is formed).
struct SaxpyOp {
  float a;
  __device__ __forceinline__ void operator()(const float (&in)[2], float (&out)[1]) const {
    float p = a * in[0];
    out[0] = p + in[1];
  }
};

// ---------------------------------------------------------------- maps
#define NT2H_MAP_OP(NAME, FN_F, FN_D)                                                          \
  template <class P> struct NAME {                                                             \
    __device__ __forceinline__ void operator()(const float (&in)[1], float (&out)[1]) const { \
      out[0] = FN_F<P>(in[0]);                                                                 \
    }                                                                                          \
    __device__ __forceinline__ void operator()(const double (&in)[1], double (&out)[1]) const {\
      out[0] = FN_D<P>(in[0]);                                                                 \
    }                                                                                          \
  };
NT2H_MAP_OP(OpExp, exp_f, exp_d)
NT2H_MAP_OP(OpLog, log_f, log_d)
NT2H_MAP_OP(OpErf, erf_f, erf_d)
NT2H_MAP_OP(OpErfc, erfc_f, erfc_d)
NT2H_MAP_OP(OpNormcdf, normcdf_f, normcdf_d)
NT2H_MAP_OP(OpNormpdf, normpdf_f, normpdf_d)
NT2H_MAP_OP(OpTanh, tanh_f, tanh_d)
NT2H_MAP_OP(OpSin, sin_f, sin_d)
NT2H_MAP_OP(OpCos, cos_f, cos_d)
#undef NT2H_MAP_OP
template <class P> struct OpSqrt {
  __device__ __forceinline__ void operator()(const float (&in)[1], float (&out)[1]) const { out[0] = P::sqrt(in[0]); }
  __device__ __forceinline__ void operator()(const double (&in)[1], double (&out)[1]) const { out[0] = P::sqrt(in[0]); }
};
template <class P> struct OpFastNormcdf {
  __device__ __forceinline__ void operator()(const float (&in)[1], float (&out)[1]) const { out[0] = fastnormcdf_f<P>(in[0]); }
};
template <class P> struct OpExp2 {
  __device__ __forceinline__ void operator()(const float (&in)[1], float (&out)[1]) const { out[0] = exp2_f<P>(in[0]); }
};
template <class P> struct OpExp10 {
  __device__ __forceinline__ void operator()(const float (&in)[1], float (&out)[1]) const { out[0] = exp10_f<P>(in[0]); }
};

// EXACT float sin / cos over 16-byte vectors with the arguments above
// 2^18 pi compacted.  There NT2 runs fdlibm's float Payne-Hanek; run in place
// it would occupy a whole wave for each of the lane's four elements whenever
// any lane needs it.  Instead each wave reduces the band (2^6 pi, 2^18 pi]
// on every lane without branching, runs the ladder of NT2's reduction only
// for lanes holding smaller arguments, and gathers the far ones (ballot +
// prefix count) into an LDS list that all 64 lanes reduce with
// reduce_fdlibm_f_fast.  The arguments it leaves undecided (about 1.5% of
// those up to 2^28) and those beyond 2^28 are queued per wave in LDS with
// their element index; their vector is stored with a provisional
// value, and whenever 64 are queued (and once at the end) the wave reduces
// them with reduce_fdlibm_f on all 64 lanes and stores each result over the
// provisional one, after waiting for its own earlier stores to complete.
// The arithmetic per element is that of sin_f / cos_f, so the results are.
// Lists and queues are per wave, so the waves synchronise on their own (no
// workgroup barrier: a wave with many far arguments does not hold up the
// others).  Each wave runs several iterations so that the queue mostly
// fills before it is drained.
constexpr unsigned kTrigExactGrid = 16384;
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Reduces and stores the first cnt (<= 64) queued arguments of wave w.
template <bool COS>
__device__ __forceinline__ void trig_exact_drain(const float (&dqx)[kWaves][128], const int64_t (&dqi)[kWaves][128],
                                                 int w, int lane, int cnt, float* __restrict__ y) {
  // the provisional values were stored by this wave's earlier vector stores:
  // let them complete before the same addresses are written again
  __builtin_amdgcn_s_waitcnt(0);
  if (lane < cnt) {
    const float a0 = dqx[w][lane];
    float r;
    const int qq = reduce_fdlibm_f(__builtin_fabsf(a0), &r);
    y[dqi[w][lane]] = COS ? cos_from<Exact>(a0, qq, r) : sin_from<Exact>(a0, qq, r);
  }
}
template <bool COS>
__global__ __launch_bounds__(kBlock) void trig_exact_kernel(const float* __restrict__ x, float* __restrict__ y,
                                                            int64_t n) {
  __shared__ float phx[kWaves][4 * 64];
  __shared__ float phr[kWaves][4 * 64];
  __shared__ int phq[kWaves][4 * 64];
  __shared__ float dqx[kWaves][128];    // queued arguments (signed)
  __shared__ int64_t dqi[kWaves][128];  // and their element indices
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));  // lanes under this one
  const int64_t nv = n / 4;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int nq = 0;  // queued entries of this wave (wave-uniform)
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < nv; base += stride) {
    const int64_t i = base + threadIdx.x;
    const bool act = i < nv;
    f32x4 v = act ? __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(x) + i) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // gather the arguments above 2^18 pi
    int slot[4];
    int total = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float ax = __builtin_fabsf(v[e]);
      const bool far = ax > kFdlibmLoF && ax <= 0x1.fffffep+127f;
      const uint64_t m = __ballot(far);
      slot[e] = far ? total + __popcll(m & below) : -1;
      if (far) phx[w][slot[e]] = ax;
      total += __popcll(m);
    }
    wave_sync();
    for (int j = lane; j < total; j += 64) {
      const float ax = phx[w][j];
      float r = 0.0f;
      int qq = -1;  // queued
      if (ax <= kFdlibmFastHiF && !reduce_fdlibm_f_fast(ax, &r, &qq)) qq = -1;
      phq[w][j] = qq;
      phr[w][j] = r;
    }
    wave_sync();
    f32x4 o;
    bool later[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float ax = __builtin_fabsf(v[e]);
      const bool bad = !(ax <= 0x1.fffffep+127f);
      float xr;
      // the band on every lane without a branch; NT2's ladder only where a
      // lane holds a smaller argument
      int q = reduce_exact_f_band(ax, &xr);
      if (slot[e] >= 0) {
        xr = phr[w][slot[e]];
        q = phq[w][slot[e]];
      } else if (!in_exact_band_f(ax)) {
        q = reduce_exact_f_near(bad ? 0.0f : ax, &xr);
      }
      later[e] = q < 0;
      if (later[e]) q = 0;  // provisional
      o[e] = COS ? cos_from<Exact>(v[e], q, xr) : sin_from<Exact>(v[e], q, xr);
    }
    if (act) __builtin_nontemporal_store(o, reinterpret_cast<f32x4*>(y) + i);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint64_t m = __ballot(later[e]);
      if (m == 0) continue;  // wave-uniform
      if (later[e]) {
        const int sl = nq + __popcll(m & below);
        dqx[w][sl] = v[e];
        dqi[w][sl] = 4 * i + e;
      }
      nq += __popcll(m);
      if (nq >= 64) {  // drain 64, move the rest (at most 63) to the front
        wave_sync();
        trig_exact_drain<COS>(dqx, dqi, w, lane, 64, y);
        const bool mv = lane < nq - 64;
        const float mx = mv ? dqx[w][64 + lane] : 0.0f;
        const int64_t mi = mv ? dqi[w][64 + lane] : 0;
        wave_sync();
        if (mv) {
          dqx[w][lane] = mx;
          dqi[w][lane] = mi;
        }
        nq -= 64;
      }
    }
    wave_sync();  // the lists are rewritten by the next iteration
  }
  if (nq > 0) {
    wave_sync();
    trig_exact_drain<COS>(dqx, dqi, w, lane, nq, y);
  }
  const int64_t t = 4 * nv + (int64_t)blockIdx.x * kBlock + threadIdx.x;  // tail [4 nv, n)
  if (t < n) y[t] = COS ? cos_f<Exact>(x[t]) : sin_f<Exact>(x[t]);
}

template <class T, class P>
static int map_dispatch(int op, const T* x, T* y, int64_t n, nt2h_stream_t s) {
  if constexpr (sizeof(T) == 4 && !P::fused) {
    if ((op == NT2H_OP_SIN || op == NT2H_OP_COS) && n > 0 && aligned16(x) && aligned16(y)) {
      const unsigned g = stream_grid(n, 4, kTrigExactGrid);
      if (op == NT2H_OP_SIN)
        hipLaunchKernelGGL((trig_exact_kernel<false>), dim3(g), dim3(kBlock), 0, as_hip(s), x, y, n);
      else
        hipLaunchKernelGGL((trig_exact_kernel<true>), dim3(g), dim3(kBlock), 0, as_hip(s), x, y, n);
      NT2H_HIP(hipGetLastError());
      return NT2H_OK;
    }
  }
  Streams<T, 1, 1> st{{x}, {y}};
  switch (op) {
    case NT2H_OP_EXP: return launch_ew(st, n, OpExp<P>{}, s);
    case NT2H_OP_LOG: return launch_ew(st, n, OpLog<P>{}, s);
    case NT2H_OP_SQRT: return launch_ew(st, n, OpSqrt<P>{}, s);
    case NT2H_OP_ERF: return launch_ew(st, n, OpErf<P>{}, s, 65536);
    case NT2H_OP_ERFC: return launch_ew(st, n, OpErfc<P>{}, s, 65536);
    case NT2H_OP_NORMCDF: return launch_ew(st, n, OpNormcdf<P>{}, s, 65536);
    // double tanh, sin and cos: 65536 grid-striding workgroups (measured at
    // 2^28: sin 943 -> 807 us, tanh 899 -> 832 us; exp and log lose with it)
    case NT2H_OP_TANH: return launch_ew(st, n, OpTanh<P>{}, s, sizeof(T) == 8 ? 65536 : kEwOnePerLane);
    case NT2H_OP_SIN: return launch_ew(st, n, OpSin<P>{}, s, sizeof(T) == 8 ? 65536 : kEwOnePerLane);
    case NT2H_OP_COS: return launch_ew(st, n, OpCos<P>{}, s, sizeof(T) == 8 ? 65536 : kEwOnePerLane);
    case NT2H_OP_NORMPDF: return launch_ew(st, n, OpNormpdf<P>{}, s);
  }
  if constexpr (sizeof(T) == 4) {
    switch (op) {
      case NT2H_OP_FASTNORMCDF: return launch_ew(st, n, OpFastNormcdf<P>{}, s);
      case NT2H_OP_EXP2: return launch_ew(st, n, OpExp2<P>{}, s);
      case NT2H_OP_EXP10: return launch_ew(st, n, OpExp10<P>{}, s);
    }
  }
  return NT2H_ERR_BAD_OP;
}

static bool op_valid(int op, bool is_float) {
  if (op >= NT2H_OP_EXP && op <= NT2H_OP_NORMPDF) return true;
  return is_float && op >= NT2H_OP_FASTNORMCDF && op <= NT2H_OP_EXP10;
}

// ------------------------------------------------------- Black-Scholes
// normcdf(d) = 0.5*erfc(-sqrt(1/2)*d) and normcdf(-d) share one erfc core:
// erfc(-t) = 2 - erfc(t), the same |t| enters both (NT2 evaluates them
// separately, with identical results).
template <class P> __device__ __forceinline__ void ncdf_pair(float d, float& nd, float& nmd) {
  float t = -0x1.6a09e6p-1f * d;
  float r = P::fused ? erfc_pos_e<P>(__builtin_fabsf(t), exp_neg_half_square<P>(d)) : erfc_pos<P>(__builtin_fabsf(t));
  float rr = 2.0f - r;
  nd = 0.5f * ((t < 0.0f) ? rr : r);
  nmd = 0.5f * ((t > 0.0f) ? rr : r);
  if (d != d) { nd = d; nmd = d; }
}
// the same with exp(-d^2/2) given (Fused policy)
template <class P> __device__ __forceinline__ void ncdf_pair_e(float d, float e, float& nd, float& nmd) {
  float t = -0x1.6a09e6p-1f * d;
  float r = erfc_pos_e<P>(__builtin_fabsf(t), e);
  float rr = 2.0f - r;
  nd = 0.5f * ((t < 0.0f) ? rr : r);
  nmd = 0.5f * ((t > 0.0f) ? rr : r);
  if (d != d) { nd = d; nmd = d; }
}
template <class P> __device__ __forceinline__ void ncdf_pair(double d, double& nd, double& nmd) {
  double t = -0x1.6a09e667f3bcdp-1 * d;
  double r = erfc_pos_d<P>(__builtin_fabs(t));
  double rr = 2.0 - r;
  nd = 0.5 * ((t < 0.0) ? rr : r);
  nmd = 0.5 * ((t > 0.0) ? rr : r);
  if (d != d) { nd = d; nmd = d; }
}
template <class P> __device__ __forceinline__ float logT(float x) { return log_f<P>(x); }
template <class P> __device__ __forceinline__ double logT(double x) { return log_d<P>(x); }
template <class P> __device__ __forceinline__ float expT(float x) { return exp_f<P>(x); }
template <class P> __device__ __forceinline__ double expT(double x) { return exp_d<P>(x); }

template <class T, class P> struct BSFwdOp {
  __device__ __forceinline__ void operator()(const T (&in)[5], T (&out)[2]) const {
    const T S = in[0], K = in[1], Tm = in[2], r = in[3], sig = in[4];
    if constexpr (P::fused && sizeof(T) == 4) {
      // exp(-d2^2/2) = exp(-d1^2/2) (S/K) e^{rT} (d2 = d1 - sig sqrt T), as in
      // the adjoint kernel: one exp for both normcdf pairs, and more accurate
      // than exp(-d2^2/2) of the rounded d2.  Where a factor left the float
      // range (a product above 1: e^{-rT} underflowed or (S/K) e^{rT}
      // overflowed; or e^{-rT} itself overflowed, which makes (S/K) e^{rT}
      // vanish) exp(-d2^2/2) is formed from d2 instead (rare lanes)
      const float sst = sig * P::sqrt(Tm);
      const float sk = P::div(S, K);
      const float d1 = P::div(P::madd(P::madd(sig, sig * 0.5f, r), Tm, logT<P>(sk)), sst);
      const float d2 = d1 - sst;
      const float ert = expT<P>(-r * Tm);
      const float kert = K * ert;
      const float e1 = exp_neg_half_square<P>(d1);
      float e2 = e1 * (sk * P::rcp(ert));
      if (!(e2 <= 2.0f) || !(ert <= 0x1.fffffep+127f)) e2 = exp_neg_half_square<P>(d2);
      float n1, nm1, n2, nm2;
      ncdf_pair_e<P>(d1, e1, n1, nm1);
      ncdf_pair_e<P>(d2, e2, n2, nm2);
      out[0] = S * n1 - kert * n2;
      out[1] = kert * nm2 - S * nm1;
      return;
    }
    T sqT = P::sqrt(Tm);
    T sst = sig * sqT;
    T d1;
    if (P::fused) d1 = P::div(P::madd(P::madd(sig, sig * T(0.5), r), Tm, logT<P>(P::div(S, K))), sst);
    else d1 = P::div(logT<P>(P::div(S, K)) + (r + (sig * sig) * T(0.5)) * Tm, sst);
    T d2 = d1 - sst;
    T kert = K * expT<P>(-r * Tm);
    T n1, nm1, n2, nm2;
    ncdf_pair<P>(d1, n1, nm1);
    ncdf_pair<P>(d2, n2, nm2);
    out[0] = S * n1 - kert * n2;
    out[1] = kert * nm2 - S * nm1;
  }
};

// grid-striding workgroups for the Black-Scholes forward bodies that are
// long enough to prefer fewer, longer-lived workgroups
constexpr int64_t kBsLongBlocks = 32768;

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_saxpy_f32(float alpha, const float* x, float* y, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  NT2H_TRY(check_ptrs(n, x, y));
  Streams<float, 2, 1> st{{x, y}, {y}};
  return launch_ew(st, n, SaxpyOp{alpha}, s);
}

int nt2h_map_f32(int op, int precision, const float* x, float* y, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  if (!op_valid(op, true)) return NT2H_ERR_BAD_OP;
  if (precision != NT2H_EXACT && precision != NT2H_FUSED) return NT2H_ERR_INVALID_ARG;
  NT2H_TRY(check_ptrs(n, x, y));
  return precision == NT2H_EXACT ? map_dispatch<float, Exact>(op, x, y, n, s)
                                 : map_dispatch<float, Fused>(op, x, y, n, s);
}

int nt2h_map_f64(int op, int precision, const double* x, double* y, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  if (!op_valid(op, false)) return NT2H_ERR_BAD_OP;
  if (precision != NT2H_EXACT && precision != NT2H_FUSED) return NT2H_ERR_INVALID_ARG;
  NT2H_TRY(check_ptrs(n, x, y));
  return precision == NT2H_EXACT ? map_dispatch<double, Exact>(op, x, y, n, s)
                                 : map_dispatch<double, Fused>(op, x, y, n, s);
}

int nt2h_bs_fwd_f32(int precision, const float* S, const float* K, const float* T, const float* r,
                    const float* sig, float* call, float* put, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  if (precision != NT2H_EXACT && precision != NT2H_FUSED) return NT2H_ERR_INVALID_ARG;
  NT2H_TRY(check_ptrs(n, S, K, T, r, sig, call, put));
  Streams<float, 5, 2> st{{S, K, T, r, sig}, {call, put}};
  // measured at 2^26 options: FUSED one vector per lane (310 us; 339 with
  // 32768 grid-striding workgroups), EXACT and FP64 32768 workgroups
  // (EXACT 476 vs 497 us, FP64 FUSED 1062 vs 1141 us)
  return precision == NT2H_EXACT ? launch_ew(st, n, BSFwdOp<float, Exact>{}, s, kBsLongBlocks)
                                 : launch_ew(st, n, BSFwdOp<float, Fused>{}, s);
}

int nt2h_bs_fwd_f64(int precision, const double* S, const double* K, const double* T, const double* r,
                    const double* sig, double* call, double* put, int64_t n, nt2h_stream_t s) {
  NT2H_TRY(check_n(n));
  if (precision != NT2H_EXACT && precision != NT2H_FUSED) return NT2H_ERR_INVALID_ARG;
  NT2H_TRY(check_ptrs(n, S, K, T, r, sig, call, put));
  Streams<double, 5, 2> st{{S, K, T, r, sig}, {call, put}};
  return precision == NT2H_EXACT ? launch_ew(st, n, BSFwdOp<double, Exact>{}, s, kBsLongBlocks)
                                 : launch_ew(st, n, BSFwdOp<double, Fused>{}, s, kBsLongBlocks);
}

}  // extern "C"
