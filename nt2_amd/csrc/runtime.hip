// Runtime layer of libnt2hip: device selection, buffers, copies, status
// strings and the per-(device, stream) reduction workspace.
//
// Stands in for NT2's CUDA buffer/locality layer (modules/arch/cuda/include/
// nt2/sdk/memory/cuda/{allocator,buffer,pinned_allocator}.hpp,
// memory/functions/cuda/copy.hpp:52-65), with status codes in place of
// CUDA_ERROR's fprintf+exit (sdk/cuda/cuda.hpp:44-50).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "nt2h_common.hpp"

namespace nt2h {

static thread_local int g_last_hip_error = 0;
void set_hip_error(hipError_t e) { g_last_hip_error = (int)e; }

namespace {
std::mutex g_ws_mutex;
std::map<std::pair<int, hipStream_t>, Workspace> g_ws;

// HBM colouring of large allocations.  A fused loop nest reads and writes
// element i of all its tables at once; tables that start at the same offset
// modulo the allocator's 2 MiB alignment put those accesses on the same HBM
// channel and bank group (a 6-in/6-out stream pattern lost ~15% of its
// bandwidth that way).  Allocations of kColorMin bytes or more therefore
// start at successive 4 KiB colours: colour c = (address / 4 KiB) mod
// kColors, cycling over calls.  The returned pointer is 4 KiB aligned;
// nt2h_free maps it back to the allocation.
constexpr size_t kColorMin = size_t(1) << 20;
constexpr size_t kColorStep = 4096;
constexpr unsigned kColors = 16;
std::mutex g_alloc_mutex;
std::map<void*, void*> g_colored;  // returned pointer -> hipMalloc base
unsigned g_next_color = 0;
}  // namespace

// Allocation cannot be recorded into a graph: a stream that is capturing
// must already own its workspace and scratch.
static bool capturing(nt2h_stream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(as_hip(s), &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Device bytes of one reduction workspace (the allocations below).
static size_t workspace_bytes() {
  return (size_t)kMaxPartials * kMaxFoldBlocks * sizeof(double) * 2 + (size_t)kMaxPartials * kMaxGroups * sizeof(double) +
         (size_t)(kMaxGroups + 1) * kCounterStride * sizeof(unsigned) + kMaxPartials * sizeof(double);
}

int get_workspace(nt2h_stream_t s, Workspace* ws) {
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  auto key = std::make_pair(dev, as_hip(s));
  auto it = g_ws.find(key);
  if (it != g_ws.end()) {
    *ws = it->second;
    return NT2H_OK;
  }
  if (capturing(s)) return NT2H_ERR_CAPTURE;
  Workspace w{};
  NT2H_HIP(hipMalloc((void**)&w.partials, (size_t)kMaxPartials * kMaxFoldBlocks * sizeof(double)));
  NT2H_HIP(hipMalloc((void**)&w.gpartials, (size_t)kMaxPartials * kMaxGroups * sizeof(double)));
  NT2H_HIP(hipMalloc((void**)&w.counters, (size_t)(kMaxGroups + 1) * kCounterStride * sizeof(unsigned)));
  NT2H_HIP(hipMalloc((void**)&w.result, kMaxPartials * sizeof(double)));
  NT2H_HIP(hipMalloc((void**)&w.bpartials, (size_t)kMaxPartials * kMaxFoldBlocks * sizeof(double)));
  // zero the tickets once, synchronously; kernels reset them after each use
  NT2H_HIP(hipMemset(w.counters, 0, (size_t)(kMaxGroups + 1) * kCounterStride * sizeof(unsigned)));
  g_ws[key] = w;
  *ws = w;
  return NT2H_OK;
}

namespace {
// A grow-only device buffer of one (device, stream).  Growing never frees
// the buffer it replaces: a graph recorded on the stream keeps using the
// old one, and kernels of earlier calls may still be running.  Retired
// buffers are freed by nt2h_release_stream (at most twice the final size
// in all, since the lookback state doubles and the scratch grows to fit).
struct Grown {
  void* p = nullptr;
  size_t bytes = 0;
  std::vector<std::pair<void*, size_t>> retired;
};
std::mutex g_scratch_mutex;
std::map<std::pair<int, hipStream_t>, Grown> g_scratch;
std::mutex g_lb_mutex;
std::map<std::pair<int, hipStream_t>, Grown> g_lb;
}  // namespace

// Grow-only device scratch of one (device, stream) for multi-launch folds
// and scans (segment summaries).  Calls on one stream are ordered, so a
// buffer is never in use by two calls at once.
int get_scratch(nt2h_stream_t s, size_t bytes, void** out) {
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  Grown& sc = g_scratch[std::make_pair(dev, as_hip(s))];
  if (sc.bytes < bytes) {
    if (capturing(s)) return NT2H_ERR_CAPTURE;
    size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
    void* p = nullptr;
    NT2H_HIP(hipMalloc(&p, want));
    if (sc.p) sc.retired.emplace_back(sc.p, sc.bytes);
    sc.p = p;
    sc.bytes = want;
  }
  *out = sc.p;
  return NT2H_OK;
}

int get_lookback(nt2h_stream_t s, size_t bytes, void** out) {
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_lb_mutex);
  Grown& lb = g_lb[std::make_pair(dev, as_hip(s))];
  if (lb.bytes < bytes) {
    if (capturing(s)) return NT2H_ERR_CAPTURE;
    size_t want = (size_t)1 << 20;
    while (want < bytes) want <<= 1;
    void* p = nullptr;
    NT2H_HIP(hipMalloc(&p, want));
    // zeroed once, synchronously: the kernels keep it consistent (epoch
    // tags on the tile states, counters reset by the last workgroup)
    NT2H_HIP(hipMemset(p, 0, want));
    if (lb.p) lb.retired.emplace_back(lb.p, lb.bytes);
    lb.p = p;
    lb.bytes = want;
  }
  *out = lb.p;
  return NT2H_OK;
}

// Error flag of the look-back scans on this stream (ctrl[3] of its state:
// 1 = a bounded wait gave up, 2 = more tickets than tiles): read and
// cleared once the stream is idle.
static int lookback_status(nt2h_stream_t s) {
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_lb_mutex);
    auto it = g_lb.find(std::make_pair(dev, as_hip(s)));
    if (it == g_lb.end() || !it->second.p) return NT2H_OK;
    p = it->second.p;
  }
  unsigned flag = 0;
  unsigned* f = static_cast<unsigned*>(p) + 3;
  NT2H_HIP(hipMemcpyAsync(&flag, f, sizeof flag, hipMemcpyDeviceToHost, as_hip(s)));
  NT2H_HIP(hipStreamSynchronize(as_hip(s)));
  if (flag == 0) return NT2H_OK;
  const unsigned zero = 0;
  NT2H_HIP(hipMemcpyAsync(f, &zero, sizeof zero, hipMemcpyHostToDevice, as_hip(s)));
  NT2H_HIP(hipStreamSynchronize(as_hip(s)));
  return NT2H_ERR_WAIT;
}

int is_host_pointer(const void* p, bool* host, nt2h_stream_t s) {
  // Under stream capture the attribute query is refused (and would
  // invalidate the capture), and a host result would need a stream
  // synchronisation, which a graph cannot contain: result pointers passed
  // while the stream is capturing are device memory.
  if (capturing(s)) {
    *host = false;
    return NT2H_OK;
  }
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory: clear the sticky error
    *host = true;
    return NT2H_OK;
  }
  *host = !(attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged);
  return NT2H_OK;
}

int deliver(nt2h_stream_t s, double* out, const double* dev, int nv, bool out_is_host) {
  if (!out_is_host) return NT2H_OK;  // kernel wrote out directly
  NT2H_HIP(hipMemcpyAsync(out, dev, nv * sizeof(double), hipMemcpyDeviceToHost, as_hip(s)));
  NT2H_HIP(hipStreamSynchronize(as_hip(s)));
  return NT2H_OK;
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_init(int device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return NT2H_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) return NT2H_ERR_NO_DEVICE;
  NT2H_HIP(hipSetDevice(device));
  NT2H_HIP(hipFree(nullptr));  // create the context now rather than on first launch
  // the static scratch of the allocation-free entry point, resolved here so
  // that its callers never meet the set-up (a memset and a synchronisation)
  NT2H_TRY(noalloc_setup(device, nullptr));
  return NT2H_OK;
}

int nt2h_device_count(int* count) {
  if (!count) return NT2H_ERR_NULL_POINTER;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
  }
  return NT2H_OK;
}

const char* nt2h_strerror(int code) {
  switch (code) {
    case NT2H_OK: return "success";
    case NT2H_ERR_INVALID_ARG: return "invalid argument";
    case NT2H_ERR_NULL_POINTER: return "null pointer";
    case NT2H_ERR_BAD_OP: return "unsupported functor id";
    case NT2H_ERR_NO_DEVICE: return "no HIP device";
    case NT2H_ERR_HIP: return "HIP runtime error";
    case NT2H_ERR_OUT_OF_MEMORY: return "out of device memory";
    case NT2H_ERR_TOO_LARGE: return "size too large";
    case NT2H_ERR_CAPTURE: return "workspace allocation needed while the stream is capturing";
    case NT2H_ERR_WAIT: return "a bounded wait in a kernel gave up: that call's results are invalid";
  }
  return "unknown status";
}

int nt2h_last_hip_error(void) { return g_last_hip_error; }

int nt2h_malloc(void** ptr, size_t bytes) {
  if (!ptr) return NT2H_ERR_NULL_POINTER;
  *ptr = nullptr;
  if (bytes == 0) return NT2H_OK;
  if (bytes < kColorMin) {
    NT2H_HIP(hipMalloc(ptr, bytes));
    return NT2H_OK;
  }
  if (bytes > SIZE_MAX - kColorStep * kColors) return NT2H_ERR_TOO_LARGE;
  void* base = nullptr;
  NT2H_HIP(hipMalloc(&base, bytes + kColorStep * kColors));
  std::lock_guard<std::mutex> lock(g_alloc_mutex);
  const unsigned want = g_next_color++ % kColors;
  const uintptr_t b = (uintptr_t)base;
  const uintptr_t first = (b + kColorStep - 1) / kColorStep * kColorStep;  // 4 KiB aligned start
  const unsigned have = (unsigned)((first / kColorStep) % kColors);
  void* p = (void*)(first + (size_t)((want + kColors - have) % kColors) * kColorStep);
  g_colored[p] = base;
  *ptr = p;
  return NT2H_OK;
}

int nt2h_free(void* ptr) {
  if (!ptr) return NT2H_OK;
  void* base = ptr;
  {
    std::lock_guard<std::mutex> lock(g_alloc_mutex);
    auto it = g_colored.find(ptr);
    if (it != g_colored.end()) {
      base = it->second;
      g_colored.erase(it);
    }
  }
  NT2H_HIP(hipFree(base));
  return NT2H_OK;
}

int nt2h_host_alloc(void** ptr, size_t bytes) {
  if (!ptr) return NT2H_ERR_NULL_POINTER;
  *ptr = nullptr;
  if (bytes == 0) return NT2H_OK;
  NT2H_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return NT2H_OK;
}

int nt2h_host_free(void* ptr) {
  if (!ptr) return NT2H_OK;
  NT2H_HIP(hipHostFree(ptr));
  return NT2H_OK;
}

int nt2h_memcpy(void* dst, const void* src, size_t bytes, int kind, nt2h_stream_t s) {
  if (bytes == 0) return NT2H_OK;
  if (!dst || !src) return NT2H_ERR_NULL_POINTER;
  hipMemcpyKind k;
  switch (kind) {
    case NT2H_H2H: k = hipMemcpyHostToHost; break;
    case NT2H_H2D: k = hipMemcpyHostToDevice; break;
    case NT2H_D2H: k = hipMemcpyDeviceToHost; break;
    case NT2H_D2D: k = hipMemcpyDeviceToDevice; break;
    case NT2H_DEFAULT: k = hipMemcpyDefault; break;
    default: return NT2H_ERR_INVALID_ARG;
  }
  NT2H_HIP(hipMemcpyAsync(dst, src, bytes, k, as_hip(s)));
  return NT2H_OK;
}

int nt2h_sync(nt2h_stream_t s) {
  NT2H_HIP(hipStreamSynchronize(as_hip(s)));
  return lookback_status(s);
}

int nt2h_stream_buffer(nt2h_stream_t s, int kind, size_t bytes, void** out) {
  if (!out) return NT2H_ERR_NULL_POINTER;
  *out = nullptr;
  if (kind == 0) return get_scratch(s, bytes, out);
  if (kind == 1) return get_lookback(s, bytes, out);
  return NT2H_ERR_INVALID_ARG;
}

int nt2h_state_bytes(size_t* bytes) {
  if (!bytes) return NT2H_ERR_NULL_POINTER;
  size_t n = 0;
  {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    n += g_ws.size() * workspace_bytes();
  }
  for (auto* m : {&g_scratch, &g_lb}) {
    std::lock_guard<std::mutex> lock(m == &g_scratch ? g_scratch_mutex : g_lb_mutex);
    for (auto& kv : *m) {
      n += kv.second.bytes;
      for (auto& r : kv.second.retired) n += r.second;
    }
  }
  *bytes = n;
  return NT2H_OK;
}

int nt2h_release_stream(nt2h_stream_t s) {
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  NT2H_HIP(hipStreamSynchronize(as_hip(s)));
  const auto key = std::make_pair(dev, as_hip(s));
  int st = NT2H_OK;
  auto keep = [&](hipError_t e) {
    if (e != hipSuccess && st == NT2H_OK) st = hip_check(e);
  };
  {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto it = g_ws.find(key);
    if (it != g_ws.end()) {
      Workspace& w = it->second;
      for (void* p : {(void*)w.partials, (void*)w.gpartials, (void*)w.counters, (void*)w.result, (void*)w.bpartials})
        if (p) keep(hipFree(p));
      g_ws.erase(it);
    }
  }
  for (auto* m : {&g_scratch, &g_lb}) {
    std::lock_guard<std::mutex> lock(m == &g_scratch ? g_scratch_mutex : g_lb_mutex);
    auto it = m->find(key);
    if (it != m->end()) {
      if (it->second.p) keep(hipFree(it->second.p));
      for (auto& r : it->second.retired) keep(hipFree(r.first));
      m->erase(it);
    }
  }
  return st;
}

}  // extern "C"
