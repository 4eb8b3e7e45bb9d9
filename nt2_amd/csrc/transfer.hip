// Host <-> device movement of tables, and a host-table Black-Scholes
// fwd+adj that overlaps the transfers with the compute.
//
// NT2's CUDA layer moves a host table through pinned staging buffers in
// blocks, one block per CUDA stream, so that the copy of one block overlaps
// the work on another (modules/arch/cuda/include/nt2/sdk/cuda/settings/
// specific_cuda.hpp: cu_buffers, transfer_htd / transfer_dth), and copies
// whole device tables with cudaMemcpy (memory/functions/cuda/copy.hpp:52-99).
// Here:
//  * nt2h_upload / nt2h_download copy between pageable host memory and the
//    device through three 8 MiB pinned staging buffers per device: the
//    host-side copy of chunk k (split over a few host threads) overlaps the
//    DMA of chunk k-1 (and, downwards, of chunk k+1);
//  * nt2h_bs_fwd_adj_f32_host runs the fused fwd+adj kernel over host
//    tables in blocks on two HIP streams: while one stream moves block j up,
//    computes it and moves its six outputs down, the host stages block j+1
//    for the other stream and unstages block j-1.  The two FP64 totals are
//    summed over blocks in block order, so the result does not depend on
//    the timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "nt2h_common.hpp"

namespace nt2h {
namespace {

// ------------------------------------------------ parallel host memcpy
// A small persistent pool: a bulk copy is cut into kWorkers slices.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool p;
    return p;
  }
  void copy(void* dst, const void* src, size_t bytes) {
    if (bytes < (size_t(2) << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    std::lock_guard<std::mutex> one(call_m_);  // one bulk copy at a time
    std::unique_lock<std::mutex> lk(m_);
    const size_t slice = ((bytes + kWorkers - 1) / kWorkers + 4095) / 4096 * 4096;  // kWorkers slices cover bytes
    pending_ = 0;
    for (int w = 0; w < kWorkers; ++w) {
      size_t off = slice * w;
      if (off >= bytes) break;
      size_t len = std::min(slice, bytes - off);
      tasks_.push_back({static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, len});
      ++pending_;
    }
    cv_.notify_all();
    done_.wait(lk, [&] { return pending_ == 0; });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }

 private:
  static constexpr int kWorkers = 4;
  struct Task {
    char* d;
    const char* s;
    size_t n;
  };
  CopyPool() {
    for (int w = 0; w < kWorkers; ++w) threads_.emplace_back([this] { run(); });
  }
  void run() {
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || !tasks_.empty(); });
      if (stop_ && tasks_.empty()) return;
      Task t = tasks_.back();
      tasks_.pop_back();
      lk.unlock();
      std::memcpy(t.d, t.s, t.n);
      lk.lock();
      if (--pending_ == 0) done_.notify_all();
    }
  }
  std::mutex call_m_, m_;
  std::condition_variable cv_, done_;
  std::vector<Task> tasks_;
  std::vector<std::thread> threads_;
  int pending_ = 0;
  bool stop_ = false;
};

// ------------------------------------------------------ staging buffers
constexpr size_t kChunk = size_t(8) << 20;
constexpr int kStage = 3;
struct Stager {
  void* pinned[kStage] = {};
  hipEvent_t ev[kStage] = {};
  bool ready = false;
};
std::mutex g_stage_mutex;  // staged copies on one device run one at a time
std::map<int, Stager> g_stagers;

int stager(Stager** out) {
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  Stager& st = g_stagers[dev];
  if (!st.ready) {
    for (int k = 0; k < kStage; ++k) {
      NT2H_HIP(hipHostMalloc(&st.pinned[k], kChunk, hipHostMallocDefault));
      NT2H_HIP(hipEventCreateWithFlags(&st.ev[k], hipEventDisableTiming));
    }
    st.ready = true;
  }
  *out = &st;
  return NT2H_OK;
}

// ------------------------------------------------ host-table pipeline
struct BSPipe {
  int64_t block = 0;
  hipStream_t st[2] = {};
  float* dev[2] = {};    // 12 tables of `block` floats per stream
  float* pin[2] = {};    // the same, pinned host
  double* dtot[2] = {};  // device totals per stream
  double* ptot[2] = {};  // pinned copy of them
  bool streams = false;
};
std::mutex g_pipe_mutex;
std::map<int, BSPipe> g_pipes;

int pipe_for(int64_t block, BSPipe** out) {
  int dev = 0;
  NT2H_HIP(hipGetDevice(&dev));
  BSPipe& p = g_pipes[dev];
  if (!p.streams) {
    for (int k = 0; k < 2; ++k) NT2H_HIP(hipStreamCreateWithFlags(&p.st[k], hipStreamNonBlocking));
    p.streams = true;
  }
  if (p.block < block) {
    for (int k = 0; k < 2; ++k) {
      if (p.dev[k]) {
        NT2H_HIP(hipStreamSynchronize(p.st[k]));
        NT2H_HIP(hipFree(p.dev[k]));
        NT2H_HIP(hipHostFree(p.pin[k]));
        NT2H_HIP(hipFree(p.dtot[k]));
        NT2H_HIP(hipHostFree(p.ptot[k]));
        p.dev[k] = p.pin[k] = nullptr;
        p.dtot[k] = p.ptot[k] = nullptr;
      }
      NT2H_HIP(hipMalloc((void**)&p.dev[k], sizeof(float) * 12 * block));
      NT2H_HIP(hipHostMalloc((void**)&p.pin[k], sizeof(float) * 12 * block, hipHostMallocDefault));
      NT2H_HIP(hipMalloc((void**)&p.dtot[k], 2 * sizeof(double)));
      NT2H_HIP(hipHostMalloc((void**)&p.ptot[k], 2 * sizeof(double), hipHostMallocDefault));
    }
    p.block = block;
  }
  *out = &p;
  return NT2H_OK;
}

}  // namespace
}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_upload(void* dst, const void* src, size_t bytes, nt2h_stream_t s) {
  if (bytes == 0) return NT2H_OK;
  if (!dst || !src) return NT2H_ERR_NULL_POINTER;
  std::lock_guard<std::mutex> lock(g_stage_mutex);
  Stager* sg = nullptr;
  NT2H_TRY(stager(&sg));
  hipStream_t st = as_hip(s);
  bool used[kStage] = {};
  size_t k = 0;
  for (size_t off = 0; off < bytes; off += kChunk, ++k) {
    const int b = (int)(k % kStage);
    const size_t len = std::min(kChunk, bytes - off);
    if (used[b]) NT2H_HIP(hipEventSynchronize(sg->ev[b]));  // its previous DMA is done
    CopyPool::get().copy(sg->pinned[b], static_cast<const char*>(src) + off, len);
    NT2H_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, sg->pinned[b], len, hipMemcpyHostToDevice, st));
    NT2H_HIP(hipEventRecord(sg->ev[b], st));
    used[b] = true;
  }
  NT2H_HIP(hipStreamSynchronize(st));
  return NT2H_OK;
}

int nt2h_download(void* dst, const void* src, size_t bytes, nt2h_stream_t s) {
  if (bytes == 0) return NT2H_OK;
  if (!dst || !src) return NT2H_ERR_NULL_POINTER;
  std::lock_guard<std::mutex> lock(g_stage_mutex);
  Stager* sg = nullptr;
  NT2H_TRY(stager(&sg));
  hipStream_t st = as_hip(s);
  const size_t nchunks = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t k) -> int {
    const int b = (int)(k % kStage);
    const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
    NT2H_HIP(hipMemcpyAsync(sg->pinned[b], static_cast<const char*>(src) + off, len, hipMemcpyDeviceToHost, st));
    NT2H_HIP(hipEventRecord(sg->ev[b], st));
    return NT2H_OK;
  };
  // keep kStage - 1 chunks in flight ahead of the host-side copy
  for (size_t k = 0; k < nchunks && k < (size_t)kStage - 1; ++k) NT2H_TRY(issue(k));
  for (size_t k = 0; k < nchunks; ++k) {
    if (k + kStage - 1 < nchunks) NT2H_TRY(issue(k + kStage - 1));  // reuses chunk k-1's buffer, copied out
    const int b = (int)(k % kStage);
    NT2H_HIP(hipEventSynchronize(sg->ev[b]));
    const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
    CopyPool::get().copy(static_cast<char*>(dst) + off, sg->pinned[b], len);
  }
  return NT2H_OK;
}

int nt2h_bs_fwd_adj_f32_host(const float* S, const float* K, const float* T, const float* r, const float* sig,
                             const float* w, float* price, float* dS, float* dK, float* dT, float* dr, float* dsig,
                             double* totals, int64_t n, int64_t block) {
  NT2H_TRY(check_n(n));
  NT2H_TRY(check_ptrs(n, S, K, T, r, sig, w, price, dS, dK, dT, dr, dsig));
  if (block <= 0) block = int64_t(1) << 20;
  block = (block + 3) / 4 * 4;  // keeps every staged table 16-byte aligned
  if (n == 0) {
    if (totals) totals[0] = totals[1] = 0.0;
    return NT2H_OK;
  }
  std::lock_guard<std::mutex> lock(g_pipe_mutex);
  BSPipe* p = nullptr;
  NT2H_TRY(pipe_for(std::min(block, (n + 3) / 4 * 4), &p));
  const int64_t B = p->block < block ? p->block : block;
  const float* in[6] = {S, K, T, r, sig, w};
  float* out[6] = {price, dS, dK, dT, dr, dsig};
  const int64_t nb = (n + B - 1) / B;
  double acc[2] = {0.0, 0.0};
  int64_t len_of[2] = {0, 0};
  int64_t start_of[2] = {0, 0};
  auto finish = [&](int k) -> int {  // block on stream k is done: unstage it
    NT2H_HIP(hipStreamSynchronize(p->st[k]));
    for (int a = 0; a < 6; ++a)
      CopyPool::get().copy(out[a] + start_of[k], p->pin[k] + (6 + a) * B, sizeof(float) * len_of[k]);
    acc[0] += p->ptot[k][0];
    acc[1] += p->ptot[k][1];
    return NT2H_OK;
  };
  for (int64_t j = 0; j < nb; ++j) {
    const int k = (int)(j & 1);
    if (j >= 2) NT2H_TRY(finish(k));  // block j-2, in block order
    const int64_t j0 = j * B, len = std::min(B, n - j0);
    float* pin = p->pin[k];
    float* dev = p->dev[k];
    for (int a = 0; a < 6; ++a) CopyPool::get().copy(pin + a * B, in[a] + j0, sizeof(float) * len);
    for (int a = 0; a < 6; ++a)
      NT2H_HIP(hipMemcpyAsync(dev + a * B, pin + a * B, sizeof(float) * len, hipMemcpyHostToDevice, p->st[k]));
    NT2H_TRY(nt2h_bs_fwd_adj_f32(dev, dev + B, dev + 2 * B, dev + 3 * B, dev + 4 * B, dev + 5 * B, dev + 6 * B,
                                 dev + 7 * B, dev + 8 * B, dev + 9 * B, dev + 10 * B, dev + 11 * B, p->dtot[k], len,
                                 reinterpret_cast<nt2h_stream_t>(p->st[k])));
    for (int a = 0; a < 6; ++a)
      NT2H_HIP(hipMemcpyAsync(pin + (6 + a) * B, dev + (6 + a) * B, sizeof(float) * len, hipMemcpyDeviceToHost,
                              p->st[k]));
    NT2H_HIP(hipMemcpyAsync(p->ptot[k], p->dtot[k], 2 * sizeof(double), hipMemcpyDeviceToHost, p->st[k]));
    start_of[k] = j0;
    len_of[k] = len;
  }
  // drain the last two blocks in block order
  for (int64_t j = std::max<int64_t>(nb - 2, 0); j < nb; ++j) NT2H_TRY(finish((int)(j & 1)));
  if (totals) {
    totals[0] = acc[0];
    totals[1] = acc[1];
  }
  return NT2H_OK;
}

}  // extern "C"
