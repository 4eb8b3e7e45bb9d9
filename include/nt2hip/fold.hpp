// Device-side fold machinery shared by libnt2hip's kernels and the
// header-only table front end: compensated accumulators, the workgroup
// reduction, and the fence-free two-level cross-workgroup combine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nt2h {

// Native clang vectors: 128-bit loads/stores that the nontemporal builtins accept.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
// The 16-byte vector of an element type and its width.
template <class T> struct Vec16;
template <> struct Vec16<float> { using type = f32x4; static constexpr int n = 4; };
template <> struct Vec16<double> { using type = f64x2; static constexpr int n = 2; };

constexpr int kBlock = 256;          // 4 wave64 per workgroup
constexpr int kWaves = kBlock / 64;
constexpr int kMaxFoldBlocks = 65536;  // workgroups of one fold launch (partials slots)
constexpr int kGroup = 64;            // workgroups combined per first-level ticket
constexpr int kMaxGroups = kMaxFoldBlocks / kGroup;
constexpr int kMaxPartials = 4;       // accumulators per block in one fold
// ticket counters sit 256 bytes apart so that groups never contend for one
// cache line (atomics to one line serialise at the coherence point)
constexpr int kCounterStride = 64;

// Reduction workspace of one (device, stream): per-block partials
// (kMaxPartials x kMaxFoldBlocks doubles), per-group partials (kMaxPartials x
// kMaxGroups), kMaxGroups + 1 ticket counters (kCounterStride apart) that the
// kernels reset after use, a result slot for host-pointer outputs, and the
// block partials of folds that finish in a second launch.
struct Workspace {
  double* partials;
  double* gpartials;
  unsigned* counters;
  double* result;     // kMaxPartials doubles
  double* bpartials;  // block partials of two-launch folds (as partials); may be null
};
// Workgroups for a fold over `units` vectors, one vector per lane per step.
inline unsigned fold_blocks(int64_t units) {
  int64_t b = (units + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > kMaxFoldBlocks) b = kMaxFoldBlocks;
  return (unsigned)b;
}
// ---------------------------------------------------------------- device
// Error-free transformation: s + e == a + b exactly (Knuth TwoSum).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}
// Compensated accumulator: value = s + c.
struct Comp {
  float s, c;
  __device__ __forceinline__ void add(float x) {
    float t, e;
    two_sum(s, x, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void merge(float s2, float c2) {
    float t, e;
    two_sum(s, s2, t, e);
    s = t;
    c = c + c2 + e;
  }
};

// The value s + c of a compensated pair.  An error term is non-finite only
// when its sum overflowed or met an infinity or a NaN, and the running sum
// stays non-finite from then on, so c is finite whenever s is; a
// non-finite s is then the sum's own IEEE value (+-inf, or NaN for NaN data
// or +inf + -inf), which adding c (NaN) would turn into NaN.
template <class T> __device__ __forceinline__ T comp_value(T s, T c) { return __builtin_isfinite(s) ? s + c : s; }

__device__ __forceinline__ float shfl_xor_f(float v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// Block-wide fixed-order reduction of NV compensated accumulators; the
// result (s + c in double) is written in thread 0 only, so callers keep
// `out` initialised (reading it elsewhere would be undefined).  Uses
// lds[NV*2*kWaves].
template <int NV>
__device__ __forceinline__ void block_reduce_comp(Comp (&acc)[NV], float* lds, double (&out)[NV]) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      float s2 = shfl_xor_f(acc[v].s, m), c2 = shfl_xor_f(acc[v].c, m);
      acc[v].merge(s2, c2);
    }
  }
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      lds[(v * kWaves + wave) * 2 + 0] = acc[v].s;
      lds[(v * kWaves + wave) * 2 + 1] = acc[v].c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      Comp a{lds[(v * kWaves) * 2], lds[(v * kWaves) * 2 + 1]};
      for (int w = 1; w < kWaves; ++w) a.merge(lds[(v * kWaves + w) * 2], lds[(v * kWaves + w) * 2 + 1]);
      out[v] = comp_value((double)a.s, (double)a.c);
    }
  }
}

// Cross-block handoff without an agent-scope fence.  On gfx950 a release
// fence at agent scope writes back the whole L2 of the XCD, which costs
// more than the fold itself when every block issues it.  Instead each
// block publishes its partials with agent-scope atomic stores (written
// through to the point of coherence of all XCDs), waits for them to
// complete, and takes a ticket with an agent-scope atomic add; the block
// holding the last ticket reads the partials back with agent-scope atomic
// loads (which bypass any stale L2 copy).
__device__ __forceinline__ void publish(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double fetch(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two-level deterministic combine of per-block partials, after
// block_reduce_comp (or an equivalent) left the block's NV sums in thread 0.
// Only wave 0 takes part; the other waves leave at once.  Thread 0
// publishes the block's partials and takes a ticket on its group of kGroup
// blocks; the block holding the group's last ticket sums the group's
// partials in block order (one lane per block, fixed shuffle tree), publishes
// the group partial and takes a ticket on the top counter (slot ng, the first
// one no group of this launch uses, so a workspace needs only ng + 1
// counters); the last group
// sums the group partials in group order.  Every ticket counter is reset by
// the block that consumed it.  Returns true in lane 0 of the final block,
// with the totals in res[0..NV).
// `between` runs once in every thread: in waves 1.. before they leave, in
// wave 0 after the ticket atomic is issued and before its result is used —
// the place for a kernel's deferred output stores, so that the ticket's
// wait covers only the partials and not the bulk stores.
struct NoDeferred {
  __device__ __forceinline__ void operator()() const {}
};
template <int NV, class Between = NoDeferred>
__device__ __forceinline__ bool grid_combine(const double (&bsum)[NV], const Workspace& ws, double* res,
                                             Between between = Between()) {
  if (threadIdx.x >= 64) {
    between();
    return false;
  }
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned ng = (nb + kGroup - 1) / kGroup, g = b / kGroup;
  const unsigned gsize = (nb - g * kGroup) < (unsigned)kGroup ? (nb - g * kGroup) : (unsigned)kGroup;
  const int lane = threadIdx.x;
  unsigned t = 0;
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) publish(&ws.partials[v * nb + b], bsum[v]);
    __builtin_amdgcn_s_waitcnt(0);
    t = __hip_atomic_fetch_add(&ws.counters[g * kCounterStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  between();
  t = __shfl(t, 0, 64);
  if (t != gsize - 1) return false;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = ((unsigned)lane < gsize) ? fetch(&ws.partials[v * nb + g * kGroup + lane]) : 0.0;
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += shfl_xor_d(acc[v], m);
  }
  unsigned t2 = 0;
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) publish(&ws.gpartials[v * ng + g], acc[v]);
    __hip_atomic_store(&ws.counters[g * kCounterStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    t2 = __hip_atomic_fetch_add(&ws.counters[ng * kCounterStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  t2 = __shfl(t2, 0, 64);
  if (t2 != ng - 1) return false;
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  // lane l sums group partials l, l + 64, l + 128, ... in that order; four
  // of them are fetched at a time so that their round trips overlap
  for (unsigned i0 = lane; i0 < ng; i0 += 256) {
    double b[NV][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const unsigned i = i0 + 64u * j;
#pragma unroll
      for (int v = 0; v < NV; ++v) b[v][j] = (i < ng) ? fetch(&ws.gpartials[v * ng + i]) : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += b[v][j];
    }
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += shfl_xor_d(acc[v], m);
  }
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) res[v] = acc[v];
    __hip_atomic_store(&ws.counters[ng * kCounterStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return lane == 0;
}

// Second launch of a two-launch fold.  A single-launch fold keeps wave 0 of
// every workgroup alive for two memory round trips (publish, then ticket),
// which on a grid of tens of thousands of workgroups holds enough wave
// slots to cost ~10% of the bandwidth; the first launch therefore only
// stores each workgroup's NV partials to ws.bpartials[v * nb + b] with plain
// stores, and this kernel (a few workgroups) sums them: workgroup k takes a
// contiguous range of blocks, lane-strided, then a fixed shuffle tree and
// the waves in order, and the ticket combine finishes over the workgroups.
// The same order on every call.
template <int NV>
__global__ __launch_bounds__(kBlock) void fold_partials_kernel(int nb, Workspace ws, double* out) {
  const int per = (nb + (int)gridDim.x - 1) / (int)gridDim.x;
  const int b0 = (int)blockIdx.x * per;
  const int b1 = (b0 + per < nb) ? b0 + per : nb;
  double s[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v] = 0.0;
  for (int i = b0 + (int)threadIdx.x; i < b1; i += kBlock) {
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] += ws.bpartials[(int64_t)v * nb + i];
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] += shfl_xor_d(s[v], m);
  }
  __shared__ double lw[NV][kWaves];
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) lw[v][threadIdx.x >> 6] = s[v];
  }
  __syncthreads();
  double bsum[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    bsum[v] = 0.0;
    if (threadIdx.x == 0) {
      double t = lw[v][0];
      for (int w = 1; w < kWaves; ++w) t += lw[v][w];
      bsum[v] = t;
    }
  }
  double res[NV];
  if (grid_combine<NV>(bsum, ws, res)) {
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] = res[v];
  }
}
// Workgroups of the second launch for nb block partials.
inline unsigned partials_blocks(int64_t nb) {
  int64_t g = (nb + 1023) / 1024;
  if (g < 1) g = 1;
  if (g > 64) g = 64;
  return (unsigned)g;
}

}  // namespace nt2h
