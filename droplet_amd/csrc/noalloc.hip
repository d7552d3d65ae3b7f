// The stable u32 key/value sort without allocation: dl_sort_pairs_u32_noalloc.
//
// dl_sort_pairs_u32 stages host arrays through library-owned device buffers,
// which it grows with hipMalloc. This variant sorts keys and values that
// already live in device memory, in place, for callers that cannot allocate
// (code under stream capture, serving loops, harnesses that forbid
// hipMalloc). The LSD passes ping-pong between the caller's arrays and two
// arrays of the same length, and the radix sort needs its digit counts and
// their scan; all of that lives in `scratch`, a file-scope device array of a
// fixed size, so a call allocates nothing. Every GPU holds its own copy of
// the array until the process ends; its device address is looked up once
// per device when the first context on that device is created.
//
// Every pass writes all of its digit counts and scan levels before reading
// them, so no part of the scratch has to start at zero. Calls on one GPU
// share the array: calls that overlap on one GPU (from several threads or
// contexts on different streams) race and give wrong results.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <string>

#include "dl_internal.h"

namespace dl {

namespace {

// the largest sort: 2^20 pairs (8 MiB of ping-pong arrays) and the radix
// workspace of that size (digit counts of 256 tiles x 512 digits, their
// offsets and the digit totals: about 1.1 MB)
constexpr size_t kNoAllocMaxN = size_t(1) << 20;
constexpr size_t kNoAllocWsBytes = size_t(2) << 20;
constexpr size_t kScratchBytes = 2 * kNoAllocMaxN * sizeof(uint32_t) + kNoAllocWsBytes;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<void *, kMaxDevices> g_scratch{};  // device address of `scratch` per device (null: not looked up)

}  // namespace

// Looks up this device's copy of `scratch` once (the device must be
// current). Called by dl_context_create.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) fail(DL_ERR_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  if (radix_workspace_bytes(kNoAllocMaxN, 32) > kNoAllocWsBytes)
    fail(DL_ERR_INTERNAL, "no-alloc sort: the radix workspace outgrew its scratch");
  void *ptr = nullptr;
  DL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(DL_ERR_DEVICE, "hipGetSymbolAddress returned no address for the sort scratch");
  g_scratch[device] = ptr;
}

uint64_t noalloc_max_pairs() { return kNoAllocMaxN; }

// Stable sort of n (key, value) pairs in device memory by the low key_bits
// bits of the keys, in place, on the context's stream; returns when done.
void noalloc_sort_pairs(Context &c, uint32_t *d_keys, uint32_t *d_vals, size_t n, int key_bits) {
  if (n > kNoAllocMaxN)
    fail(DL_ERR_CAPACITY, "dl_sort_pairs_u32_noalloc sorts at most " + std::to_string(kNoAllocMaxN) + " pairs, got " +
                              std::to_string(n));
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c.device >= 0 && c.device < kMaxDevices) base = g_scratch[c.device];
  }
  if (!base) fail(DL_ERR_INTERNAL, "the sort scratch of this device was not set up");
  c.make_current();
  hipStream_t s = c.stream;
  uint8_t *p = static_cast<uint8_t *>(base);
  uint32_t *tmp_keys = reinterpret_cast<uint32_t *>(p);
  uint32_t *tmp_vals = reinterpret_cast<uint32_t *>(p + kNoAllocMaxN * sizeof(uint32_t));
  void *ws = p + 2 * kNoAllocMaxN * sizeof(uint32_t);
  const bool in_tmp = radix_sort_pairs(s, d_keys, d_vals, tmp_keys, tmp_vals, n, key_bits, ws);
  if (in_tmp) {
    DL_HIP(hipMemcpyAsync(d_keys, tmp_keys, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    DL_HIP(hipMemcpyAsync(d_vals, tmp_vals, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
  DL_HIP(hipStreamSynchronize(s));
}

}  // namespace dl
