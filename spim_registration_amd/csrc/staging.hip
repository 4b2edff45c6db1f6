// Copies between pageable host arrays (Java arrays handed over by JNA, numpy
// arrays) and device memory for the entry points whose contract is host
// arrays.  Large copies go through two pinned staging buffers per device: a
// chunk is copied between the pageable array and one staging buffer by
// several host threads while the DMA engine moves the previous chunk, so the
// single-threaded pageable copy of the runtime is not the bound.  Small
// copies use the runtime directly.  Each call returns when its copy is done.
// The two 16 MiB pinned buffers of a device are allocated at its first large
// copy and held until the process ends; large copies on one device are
// serialised by a mutex (the JNA hooks are called from one thread per device).
// The all-or-nothing variant keeps one more pinned buffer per device, as large
// as the largest such copy so far, also held until the process ends.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "mvd_internal.h"

namespace mvd {
namespace {

constexpr size_t kStageChunk = 16u << 20;   // bytes per staging buffer
constexpr size_t kStageMin = 8u << 20;      // smaller copies use the runtime
constexpr int kMaxDevices = 64;

struct Stage {
  std::mutex mu;
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ready = false;
  void* whole = nullptr;  // pinned buffer of the all-or-nothing copies (grows, held)
  size_t whole_cap = 0;
};
Stage g_stage[kMaxDevices];

int copy_threads() {
  static const int n = [] {
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(8u, hw ? hw / 2 : 1u));
  }();
  return n;
}

// memcpy split over a few threads (the calling thread takes the first part)
void parallel_copy(void* dst, const void* src, size_t bytes) {
  const int t = bytes >= (2u << 20) ? copy_threads() : 1;
  if (t <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = (bytes / t + 63) & ~size_t(63);
  std::vector<std::thread> th;
  th.reserve(t - 1);
  for (int i = 1; i < t; ++i) {
    const size_t off = part * i;
    if (off >= bytes) break;
    const size_t n = std::min(part, bytes - off);
    th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, n); });
  }
  std::memcpy(dst, src, std::min(part, bytes));
  for (auto& x : th) x.join();
}

hipError_t stage_for(int dev, Stage** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  Stage& s = g_stage[dev];
  if (!s.ready) {
    for (int b = 0; b < 2; ++b) {
      hipError_t e = hipHostMalloc(&s.buf[b], kStageChunk, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev[b], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    s.ready = true;
  }
  *out = &s;
  return hipSuccess;
}

}  // namespace

hipError_t staged_h2d(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  if (bytes < kStageMin) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
    return e == hipSuccess ? hipStreamSynchronize(stream) : e;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_stage[std::clamp(dev, 0, kMaxDevices - 1)].mu);
  Stage* s = nullptr;
  if ((e = stage_for(dev, &s)) != hipSuccess) return e;
  size_t done = 0;
  for (int k = 0; done < bytes && e == hipSuccess; ++k) {
    const int b = k & 1;
    const size_t n = std::min(kStageChunk, bytes - done);
    if (k >= 2) e = hipEventSynchronize(s->ev[b]);  // its previous DMA has read the buffer
    if (e != hipSuccess) break;
    parallel_copy(s->buf[b], static_cast<const char*>(src) + done, n);
    e = hipMemcpyAsync(static_cast<char*>(dst) + done, s->buf[b], n, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(s->ev[b], stream);
    done += n;
  }
  const hipError_t es = hipStreamSynchronize(stream);
  return e != hipSuccess ? e : es;
}

hipError_t staged_d2h(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  if (bytes < kStageMin) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream);
    return e == hipSuccess ? hipStreamSynchronize(stream) : e;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_stage[std::clamp(dev, 0, kMaxDevices - 1)].mu);
  Stage* s = nullptr;
  if ((e = stage_for(dev, &s)) != hipSuccess) return e;
  // DMA of chunk k into buffer k & 1, then the host copy of chunk k - 1
  size_t done = 0, prev_off = 0, prev_n = 0;
  int k = 0;
  for (; done < bytes && e == hipSuccess; ++k) {
    const int b = k & 1;
    const size_t n = std::min(kStageChunk, bytes - done);
    e = hipMemcpyAsync(s->buf[b], static_cast<const char*>(src) + done, n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipEventRecord(s->ev[b], stream);
    if (e == hipSuccess && k >= 1) {
      const int pb = (k - 1) & 1;
      e = hipEventSynchronize(s->ev[pb]);
      if (e == hipSuccess) parallel_copy(static_cast<char*>(dst) + prev_off, s->buf[pb], prev_n);
    }
    prev_off = done;
    prev_n = n;
    done += n;
  }
  if (e == hipSuccess && k >= 1) {
    const int pb = (k - 1) & 1;
    e = hipEventSynchronize(s->ev[pb]);
    if (e == hipSuccess) parallel_copy(static_cast<char*>(dst) + prev_off, s->buf[pb], prev_n);
  }
  const hipError_t es = hipStreamSynchronize(stream);
  return e != hipSuccess ? e : es;
}

// Device to host where the host array must stay as it was unless the whole
// copy succeeds: the DMA fills one pinned buffer of the full size, and only
// after it has completed without error is the host array overwritten.
hipError_t staged_d2h_all_or_nothing(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  Stage& s = g_stage[dev];
  std::lock_guard<std::mutex> lk(s.mu);
  if (s.whole_cap < bytes) {
    if (s.whole) (void)hipHostFree(s.whole);
    s.whole = nullptr;
    s.whole_cap = 0;
    if ((e = hipHostMalloc(&s.whole, bytes, hipHostMallocDefault)) != hipSuccess) return e;
    s.whole_cap = bytes;
  }
  if ((e = hipMemcpyAsync(s.whole, src, bytes, hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  parallel_copy(dst, s.whole, bytes);
  return hipSuccess;
}

}  // namespace mvd
