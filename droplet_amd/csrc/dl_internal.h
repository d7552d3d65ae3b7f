// Internal declarations shared by the translation units of libdroplet_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/droplet_hip.h"

namespace dl {

// ---------------------------------------------------------------------------
// errors: every entry point converts failures into a status code plus a
// thread-local message (dl_last_error).
// ---------------------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

void set_last_error(const std::string &msg);
[[noreturn]] void fail(int code, const std::string &msg);
void check_hip(hipError_t e, const char *what);

#define DL_HIP(call) ::dl::check_hip((call), #call)

// Runs f() and maps exceptions to status codes.
template <class F>
int guard(F &&f) {
  try {
    f();
    return DL_OK;
  } catch (const Error &e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return DL_ERR_INTERNAL;
  } catch (...) {
    set_last_error("unknown error");
    return DL_ERR_INTERNAL;
  }
}

// ---------------------------------------------------------------------------
// device memory: a grow-only buffer owned by the library
// ---------------------------------------------------------------------------
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  // ensure at least `bytes` of capacity; contents are not preserved on growth
  void reserve(size_t bytes);
  void release();
  void *get() const { return ptr_; }
  template <class T>
  T *as() const { return static_cast<T *>(ptr_); }
  size_t capacity() const { return cap_; }

 private:
  void *ptr_ = nullptr;
  size_t cap_ = 0;
};

// Copies between pageable host memory and the device through kSlots pinned
// chunks, with the host side spread over a few threads (staging.hip). Both
// return when the caller's memory may be reused: to_device once the last
// chunk was read by the DMA, to_host once the data is in dst (the stream is
// synchronised). Measured on the 6M-splat scene (1.42 GB): 3 x 16 MB
// uploads at 47.5 GB/s and renders a host-array 4K frame in 31.2 ms; 2 x 8 MB
// uploaded a scene at 46.5 GB/s but took 54.3 ms for that frame; 2 x 32 MB
// and 4 x 32 MB measured within 1 ms of 3 x 16 MB. Pinned DMA alone runs at
// 57.6 GB/s and an 8-thread host copy into pinned memory at 110 GB/s.
#ifndef DL_STAGE_SLOTS
#define DL_STAGE_SLOTS 3
#endif
#ifndef DL_STAGE_CHUNK_MB
#define DL_STAGE_CHUNK_MB 16
#endif

class HostStaging {
 public:
  static constexpr int kSlots = DL_STAGE_SLOTS;
  HostStaging() = default;
  ~HostStaging();
  HostStaging(const HostStaging &) = delete;
  HostStaging &operator=(const HostStaging &) = delete;
  void to_device(void *dst, const void *src, size_t bytes, hipStream_t s);
  void to_host(void *dst, const void *src, size_t bytes, hipStream_t s);

 private:
  void ensure();
  void *buf_[kSlots] = {};
  hipEvent_t ev_[kSlots] = {};
};

// Device scratch of the voxelizer and of the voxel grid's compaction, kept
// by the context between calls (grow-only): allocating ~0.9 GB of pair
// arrays and workspaces per call of the 1M-Gaussian workload left the GPU
// idle for ~2 ms of each call. Calls on one context run one at a time.
struct VoxelScratch {
  VoxelScratch() = default;
  ~VoxelScratch();
  VoxelScratch(const VoxelScratch &) = delete;
  VoxelScratch &operator=(const VoxelScratch &) = delete;
  // voxelize: Gaussian arrays and records, per-Gaussian pair counts and
  // offsets, the (block, Gaussian) pairs and their sort, block ranges
  DeviceBuffer pos, rot, ls, op, rec, partials, counts, offs, scan_ws, small;
  DeviceBuffer keys, vals, tkeys, tvals, sort_ws, ranges;
  // compaction of non-empty blocks
  DeviceBuffer flags, fpos, omask, oidx, cscan_ws, csmall;
  // timing events of a voxelize call and a pinned host block for its small
  // readbacks (the bounds, the pair totals), created on first use
  hipEvent_t ev[4] = {};
  static constexpr size_t kHostSmallBytes = 1024;
  void *host_small = nullptr;
  void ensure_events();
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  std::string arch;
  // scratch for the standalone sort entry point
  DeviceBuffer sort_keys, sort_vals, sort_tmp_keys, sort_tmp_vals, sort_ws;
  VoxelScratch vox;
  // pinned staging of large host <-> device copies
  HostStaging staging;
  void make_current() const;
};

// ---------------------------------------------------------------------------
// device-wide primitives (scan_sort.hip)
// ---------------------------------------------------------------------------
// Exclusive scan of n u32 values; writes the total to *d_total (device) when
// d_total is not null. `ws` must hold scan_workspace_bytes(n).
size_t scan_workspace_bytes(size_t n);
void exclusive_scan_u32(hipStream_t s, const uint32_t *d_in, uint32_t *d_out, size_t n,
                        uint32_t *d_total, void *ws);
// The same with the element count on the device: min(n_cap, d_n[0] * mult)
// elements; launches are sized by n_cap and ws holds scan_workspace_bytes(n_cap).
void exclusive_scan_u32_dev(hipStream_t s, const uint32_t *d_in, uint32_t *d_out, size_t n_cap,
                            const uint32_t *d_n, uint32_t mult, uint32_t *d_total, void *ws);

// Stable LSD radix sort of (key, value) pairs on bits [0, key_bits) of the
// key. Results end in either (keys, vals) or (tmp_keys, tmp_vals); the
// return value is true when they are in the tmp buffers. With index_vals
// the input values are not read: value i is i (vals is still scratch).
// Digits are at most max_digit_bits (1..9) wide; passes = ceil(key_bits /
// max_digit_bits).
// With map.kmin set, the digits are those of (key == map.invisible ?
// map.sentinel : key - *map.kmin), kmin read on the device; the keys are
// moved unchanged. When every other key lies in [kmin, kmin + sentinel) the
// result equals the sort of the raw keys (the map is order-preserving), in
// fewer passes when sentinel < 2^key_bits.
// With low_bit > 0 the passes cover bits [low_bit, key_bits) only: the
// result is sorted by the whole key when the input is already sorted by the
// low bits among keys whose high bits are equal (the sort is stable).
struct RadixKeyMap {
  const uint32_t *kmin = nullptr;
  uint32_t invisible = 0, sentinel = 0;
};
size_t radix_workspace_bytes(size_t n, int key_bits);
bool radix_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys,
                      uint32_t *tmp_vals, size_t n, int key_bits, void *ws, int max_digit_bits = 8,
                      const RadixKeyMap &map = RadixKeyMap{}, bool index_vals = false, int low_bit = 0);
int radix_passes(int key_bits, int max_digit_bits);

// Segment table of the segmented sort (device pointers): segments are the
// contiguous element ranges [seg_start[s], seg_start[s+1]) with seg_start[0]
// = 0; blk0 / blk_seg are the sort's own tile tables.
struct SegGeom {
  const uint32_t *seg_start = nullptr;
  const uint32_t *blk0 = nullptr;
  const uint32_t *blk_seg = nullptr;
  uint32_t nseg = 0;
};

// Stable LSD radix sort of every segment on its own, with the segment bounds
// read on the device (d_seg_start, nseg + 1 entries, at most max_elems
// elements in all): no host synchronisation, so the launches can be captured
// in a graph. Result placement as radix_sort_pairs (it depends only on the
// pass count). With index_vals the input values are not read: value i is i
// (vals is still scratch).
size_t seg_radix_workspace_bytes(size_t max_elems, uint32_t nseg);
bool seg_radix_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys, uint32_t *tmp_vals,
                          const uint32_t *d_seg_start, uint32_t nseg, size_t max_elems, int key_bits,
                          int max_digit_bits, void *ws, bool index_vals = false);

// The no-alloc sort (noalloc.hip): noalloc_setup looks up the device
// address of its scratch array for `device` (current) once; the sort runs
// in that scratch, in place on device arrays, and fails with
// DL_ERR_CAPACITY above noalloc_max_pairs() pairs.
void noalloc_setup(int device);
uint64_t noalloc_max_pairs();
void noalloc_sort_pairs(Context &c, uint32_t *d_keys, uint32_t *d_vals, size_t n, int key_bits);

// off[b] = lower_bound(sorted keys, b) for b < num_bins and off[num_bins] = n
// (all zeros when n == 0): the result of find-boundaries.ts:17-99.
void bin_ranges(hipStream_t s, const uint32_t *d_keys, uint32_t n, uint32_t num_bins, uint32_t *d_off);

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

inline uint32_t bit_width_u32(uint32_t v) {
  uint32_t b = 0;
  while (v) {
    b++;
    v >>= 1;
  }
  return b;
}

}  // namespace dl

// opaque handle of include/droplet_hip.h
struct dl_context {
  dl::Context c;
};
