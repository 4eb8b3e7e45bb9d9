// nt2h_scan_f32 / nt2h_scan_f64: NT2's cumsum, cumprod, cummax and cummin
// along one dimension of a contiguous table, (lo, mid, hi) reshaping
// (core/swar/include/nt2/core/functions/scalar/{inner,outer}_scan.hpp).  The
// kernels and their launch plans are in include/nt2hip/scan_kernels.hpp,
// shared with the C++ front end, which instantiates them over expressions.
#include <hip/hip_runtime.h>

#include <atomic>

#include "nt2h_common.hpp"
#include "../../include/nt2hip/scan_kernels.hpp"

namespace nt2h {

namespace {
// Wait settings of the look-back scans (read at each launch): the poll
// limit (0: the kernel's default) and, for tests of the error path only, an
// epoch mask that makes every predecessor's state read as "not yet", so
// that each waiting tile gives up after spin_cap polls.
std::atomic<unsigned> g_lb_spin_cap{0};
std::atomic<unsigned> g_lb_epoch_mask{0};
}  // namespace

// The library's environment of a launch plan: the per-(device, stream)
// scratch and look-back state, and the thread's last HIP error.
struct ScanEnv {
  nt2h_stream_t st;
  int scratch(size_t bytes, void** p) { return get_scratch(st, bytes, p); }
  int lookback(size_t bytes, void** p) { return get_lookback(st, bytes, p); }
  void lookback_wait(unsigned* cap, unsigned* mask) {
    *cap = g_lb_spin_cap.load();
    *mask = g_lb_epoch_mask.load();
  }
  int hip(hipError_t e) { return hip_check(e); }
};

template <class T>
int scan_abi(int op, const T* x, int64_t lo, int64_t mid, int64_t hi, T* y, nt2h_stream_t s) {
  ScanEnv env{s};
  return scan_entry<T>(op, PtrSrc<T>{x}, x != nullptr, lo, mid, hi, y, as_hip(s), env);
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_scan_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* y, nt2h_stream_t s) {
  return scan_abi<float>(op, x, lo, mid, hi, y, s);
}
int nt2h_scan_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* y, nt2h_stream_t s) {
  return scan_abi<double>(op, x, lo, mid, hi, y, s);
}

int nt2h_test_set_lookback_wait(unsigned spin_cap, int never_ready) {
  g_lb_spin_cap.store(spin_cap);
  g_lb_epoch_mask.store(never_ready ? 0x80000000u : 0u);
  return NT2H_OK;
}

}  // extern "C"
