// nt2h_fold_f32 / nt2h_fold_f64: the reduction_ functors of NT2 along one
// dimension of a contiguous table, (lo, mid, hi) reshaping
// (core/reduction/include/nt2/dsl/functions/reduction/run.hpp:48-93).  The
// kernels and their launch plans are in include/nt2hip/fold_kernels.hpp,
// shared with the C++ front end, which instantiates them over expressions.
#include <hip/hip_runtime.h>

#include "nt2h_common.hpp"
#include "../../include/nt2hip/fold_kernels.hpp"

namespace nt2h {

// The library's environment of a launch plan: the per-(device, stream)
// scratch and the thread's last HIP error.
struct LibEnv {
  nt2h_stream_t st;
  int scratch(size_t bytes, void** p) { return get_scratch(st, bytes, p); }
  int hip(hipError_t e) { return hip_check(e); }
};

template <class T>
int fold_abi(int op, const T* x, int64_t lo, int64_t mid, int64_t hi, T* out, nt2h_stream_t s) {
  LibEnv env{s};
  return fold_entry<T>(op, PtrSrc<T>{x}, x != nullptr, lo, mid, hi, out, as_hip(s), env);
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_fold_f32(int op, const float* x, int64_t lo, int64_t mid, int64_t hi, float* out, nt2h_stream_t s) {
  return fold_abi<float>(op, x, lo, mid, hi, out, s);
}
int nt2h_fold_f64(int op, const double* x, int64_t lo, int64_t mid, int64_t hi, double* out, nt2h_stream_t s) {
  return fold_abi<double>(op, x, lo, mid, hi, out, s);
}

}  // extern "C"
