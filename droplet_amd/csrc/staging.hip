// Copies between pageable host memory and the device through pinned staging.
//
// A hipMemcpy from or to pageable memory goes through the runtime's own
// bounce buffers one piece at a time (the 33 MB RGBA8 image of a 4K frame
// read back at ~3 GB/s). Here a context keeps a few pinned chunks
// (hipHostMalloc): the DMA engine moves one chunk while host threads copy
// another between a pinned chunk and the caller's memory, so the DMA and
// the host copy overlap and the host copy runs on several cores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "dl_internal.h"

namespace dl {

namespace {

// A small persistent pool: parallel_copy hands each worker one slice of a
// memcpy and waits for all of them.
class CopyPool {
 public:
  static CopyPool &get() {
    static CopyPool pool;
    return pool;
  }
  size_t workers() const { return threads_.size() + 1; }

  void parallel_copy(void *dst, const void *src, size_t bytes) {
    const size_t parts = std::min<size_t>(workers(), std::max<size_t>(1, bytes >> 20));  // >= 1 MiB per part
    if (parts <= 1) {
      std::memcpy(dst, src, bytes);
      return;
    }
    // one job at a time (contexts on several host threads share the pool)
    std::lock_guard<std::mutex> job(job_mu_);
    const size_t step = ((bytes + parts - 1) / parts + 63) & ~(size_t)63;
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = static_cast<uint8_t *>(dst);
      src_ = static_cast<const uint8_t *>(src);
      bytes_ = bytes;
      step_ = step;
      next_ = 1;  // slice 0 is the caller's
      parts_ = parts;
      pending_ = parts - 1;
      gen_++;
    }
    cv_.notify_all();
    copy_slice(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  CopyPool() {
    const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
    const unsigned n = std::min(8u, hw) - 1;
    for (unsigned i = 0; i < n; i++) threads_.emplace_back([this] { run(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      gen_++;
    }
    cv_.notify_all();
    for (auto &t : threads_) t.join();
  }
  void copy_slice(size_t k) {
    const size_t off = k * step_;
    if (off < bytes_) std::memcpy(dst_ + off, src_ + off, std::min(step_, bytes_ - off));
  }
  void run() {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      while (next_ < parts_) {
        const size_t k = next_++;
        lk.unlock();
        copy_slice(k);
        lk.lock();
        if (--pending_ == 0) done_.notify_all();
      }
    }
  }

  std::vector<std::thread> threads_;
  std::mutex job_mu_, mu_;
  std::condition_variable cv_, done_;
  uint8_t *dst_ = nullptr;
  const uint8_t *src_ = nullptr;
  size_t bytes_ = 0, step_ = 0, next_ = 0, parts_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

constexpr size_t kStageChunk = size_t(DL_STAGE_CHUNK_MB) << 20;
constexpr int kSlots = HostStaging::kSlots;

}  // namespace

HostStaging::~HostStaging() {
  for (int k = 0; k < kSlots; k++) {
    if (ev_[k]) (void)hipEventDestroy(ev_[k]);
    if (buf_[k]) (void)hipHostFree(buf_[k]);
  }
}

void HostStaging::ensure() {
  if (buf_[0]) return;
  for (int k = 0; k < kSlots; k++) {
    DL_HIP(hipHostMalloc(&buf_[k], kStageChunk, hipHostMallocDefault));
    DL_HIP(hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming));
  }
}

void HostStaging::to_device(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (bytes < (size_t(1) << 20)) {
    // small copies: the runtime's own path costs less than the hand-off
    DL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  ensure();
  const uint8_t *sp = static_cast<const uint8_t *>(src);
  uint8_t *dp = static_cast<uint8_t *>(dst);
  bool used[kSlots] = {};
  for (size_t off = 0, k = 0; off < bytes; off += kStageChunk, k++) {
    const int slot = (int)(k % kSlots);
    const size_t len = std::min(kStageChunk, bytes - off);
    // the DMA that read this chunk kSlots steps ago must be done before it is refilled
    if (used[slot]) DL_HIP(hipEventSynchronize(ev_[slot]));
    CopyPool::get().parallel_copy(buf_[slot], sp + off, len);
    DL_HIP(hipMemcpyAsync(dp + off, buf_[slot], len, hipMemcpyHostToDevice, s));
    DL_HIP(hipEventRecord(ev_[slot], s));
    used[slot] = true;
  }
  // the caller's memory may change after return; the pinned chunks are the
  // DMA's source, so the last ones must be read before they are reused
  for (int k = 0; k < kSlots; k++)
    if (used[k]) DL_HIP(hipEventSynchronize(ev_[k]));
}

void HostStaging::to_host(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if (bytes < (size_t(1) << 20)) {
    DL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    return;
  }
  ensure();
  const uint8_t *sp = static_cast<const uint8_t *>(src);
  uint8_t *dp = static_cast<uint8_t *>(dst);
  const size_t chunks = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t k) {
    const int slot = (int)(k % kSlots);
    const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
    DL_HIP(hipMemcpyAsync(buf_[slot], sp + off, len, hipMemcpyDeviceToHost, s));
    DL_HIP(hipEventRecord(ev_[slot], s));
  };
  for (size_t k = 0; k + 1 < (size_t)kSlots && k < chunks; k++) issue(k);
  for (size_t k = 0; k < chunks; k++) {
    // chunks k + 1 .. k + kSlots - 1 land in the other slots while the host
    // copies chunk k out; the slot of chunk k + kSlots - 1 held chunk k - 1,
    // which the last step copied out
    if (k + kSlots - 1 < chunks) issue(k + kSlots - 1);
    const int slot = (int)(k % kSlots);
    const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
    DL_HIP(hipEventSynchronize(ev_[slot]));
    CopyPool::get().parallel_copy(dp + off, buf_[slot], len);
  }
}

}  // namespace dl
