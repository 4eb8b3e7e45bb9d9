// Shared host/device plumbing of libnt2hip: status codes, argument checks,
// per-(device, stream) reduction workspace, launch geometry and the wave64 /
// block reductions used by every fold kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nt2hip.h"
#include "../../include/nt2hip/fold.hpp"

namespace nt2h {

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
inline unsigned stream_grid(int64_t n, int vec, int64_t max_blocks = 65536) {
  int64_t units = (n + vec - 1) / vec;
  int64_t b = (units + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

int get_workspace(nt2h_stream_t s, Workspace* ws);
int get_scratch(nt2h_stream_t s, size_t bytes, void** out);
// Per (device, stream) state of the single-pass look-back scans: at least
// `bytes`, zero when first returned and then owned by those kernels (never
// shared with the scratch above).  NT2H_ERR_CAPTURE if it would have to
// grow while the stream is capturing.
int get_lookback(nt2h_stream_t s, size_t bytes, void** out);
// Deliver nv doubles computed into dev (device memory) to out (device or
// host memory); for host memory the stream is synchronised.
int deliver(nt2h_stream_t s, double* out, const double* dev, int nv, bool out_is_host);
int is_host_pointer(const void* p, bool* host, nt2h_stream_t s);
// Device address of the allocation-free entry point's static scratch on
// device dev (which must be current), resolved and cleared once per device.
int noalloc_setup(int dev, void** out);

}  // namespace nt2h
