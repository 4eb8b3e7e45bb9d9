// Shared host/device plumbing of libnt2hip: status codes, argument checks,
// per-(device, stream) reduction workspace, launch geometry and the wave64 /
// block reductions used by every fold kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nt2hip.h"

namespace nt2h {

// Native clang vectors: 128-bit loads/stores that the nontemporal builtins accept.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

constexpr int kBlock = 256;          // 4 wave64 per workgroup
constexpr int kWaves = kBlock / 64;
constexpr int kMaxReduceBlocks = 2048;  // 8 per CU on 256 CUs
constexpr int kMaxPartials = 4;       // accumulators per block in one fold

// ---------------------------------------------------------------- host
void set_hip_error(hipError_t e);
inline int hip_check(hipError_t e) {
  if (e == hipSuccess) return NT2H_OK;
  set_hip_error(e);
  return e == hipErrorOutOfMemory ? NT2H_ERR_OUT_OF_MEMORY : NT2H_ERR_HIP;
}
#define NT2H_TRY(call)                        \
  do {                                        \
    int _st = (call);                         \
    if (_st != NT2H_OK) return _st;           \
  } while (0)
#define NT2H_HIP(call) NT2H_TRY(::nt2h::hip_check(call))

inline hipStream_t as_hip(nt2h_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline int check_n(int64_t n) { return n < 0 ? NT2H_ERR_INVALID_ARG : NT2H_OK; }
template <class... Ps> inline int check_ptrs(int64_t n, Ps... ps) {
  if (n == 0) return NT2H_OK;
  const void* a[] = {static_cast<const void*>(ps)...};
  for (const void* p : a)
    if (!p) return NT2H_ERR_NULL_POINTER;
  return NT2H_OK;
}
template <class... Ps> inline bool aligned16(Ps... ps) {
  const void* a[] = {static_cast<const void*>(ps)...};
  for (const void* p : a)
    if (reinterpret_cast<uintptr_t>(p) & 15u) return false;
  return true;
}

// Grid for a streaming kernel over n elements handled vec at a time with a
// grid-stride loop: enough workgroups to fill 256 CUs several times over.
inline unsigned stream_grid(int64_t n, int vec, int64_t max_blocks = 8192) {
  int64_t units = (n + vec - 1) / vec;
  int64_t b = (units + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

// Reduction workspace of one (device, stream): kMaxReduceBlocks *
// kMaxPartials doubles of per-block partials, a completion counter that the
// last block resets, and a result slot for host-pointer outputs.
struct Workspace {
  double* partials;
  unsigned* counter;
  double* result;  // kMaxPartials doubles
};
int get_workspace(nt2h_stream_t s, Workspace* ws);
// Deliver nv doubles computed into dev (device memory) to out (device or
// host memory); for host memory the stream is synchronised.
int deliver(nt2h_stream_t s, double* out, const double* dev, int nv, bool out_is_host);
int is_host_pointer(const void* p, bool* host);

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

__device__ __forceinline__ float shfl_xor_f(float v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// Block-wide fixed-order reduction of NV compensated accumulators; the
// result (s + c in double) is valid in thread 0.  Uses lds[NV*2*kWaves].
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
      out[v] = (double)a.s + (double)a.c;
    }
  }
}

// Last-block combine of per-block partials (gridDim.x blocks, NV values
// each, partials[v * gridDim.x + b]) in FP64, in block order per thread
// stride and a fixed shuffle tree: deterministic for a given grid.  Call
// from every thread of the block after the block's own partial is written.
// Returns true in the one block that performed the combine (results in
// res[0..NV) written by thread 0); resets the counter for the next launch.
template <int NV>
__device__ __forceinline__ bool grid_combine(double* partials, unsigned* counter, double* res) {
  __shared__ bool am_last;
  __shared__ double lds[NV * kWaves];
  __threadfence();
  if (threadIdx.x == 0) {
    unsigned t = atomicAdd(counter, 1u);
    am_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += __builtin_nontemporal_load(&partials[v * gridDim.x + b]);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += shfl_xor_d(acc[v], m);
  }
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) lds[v * kWaves + wave] = acc[v];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double t = lds[v * kWaves];
      for (int w = 1; w < kWaves; ++w) t += lds[v * kWaves + w];
      res[v] = t;
    }
    *counter = 0u;
  }
  return true;
}

}  // namespace nt2h
