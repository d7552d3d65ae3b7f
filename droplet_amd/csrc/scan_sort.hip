// Device-wide exclusive scan and stable LSD radix sort (u32 key / u32 value)
// for gfx950.
//
// Both run over 4096-element tiles: the scan in 256-thread workgroups (4
// wave64 of 16 elements per lane), the sort in 512-thread workgroups (8
// wave64 of 8 elements per lane).
// The sort is the replacement for src/gpu/radix-sort.ts:91-216, whose scatter
// takes slots with atomicAdd and is therefore unstable; this one ranks every
// element inside its wave with ballots in input order, so equal keys keep
// their input order across passes (stable).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dl_common.h"
#include "dl_internal.h"

namespace dl {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTileN = kBlock * kItems;  // 4096
constexpr int kWaves = kBlock / 64;
// digits of one radix pass: at most 9 bits
constexpr int kMaxDigitBits = 9;
constexpr uint32_t kMaxDigits = 1u << kMaxDigitBits;
// Up to this many tiles, a scan's down-sweep sums the earlier tiles' totals
// itself (at most 16 loads per thread from L2) instead of a separate scan of
// the totals.
constexpr int kSumPartialsMax = 4096;

// Radix sort tiles: kRBlock threads of kRItems keys each per workgroup,
// 4096 keys per tile. Measured on the 6M-key depth sort (three passes):
// 512 x 8 takes 176 us; 256 x 16 199 us (113-120 VGPRs leave 3-4 waves
// per SIMD); 1024 x 4 234 us and 512 x 16 or 256 x 32 (8192-key tiles)
// 250-280 us, the LDS of the larger tiles or waves limiting occupancy.
// Ranking by ballots in the upsweep as in the downsweep (one LDS update per
// distinct digit) cost 10 us per pass more than LDS atomics. Tiles below
// 4096 keys would outgrow the no-alloc sort's workspace.
#ifndef DL_RADIX_THREADS
#define DL_RADIX_THREADS 512
#endif
#ifndef DL_RADIX_ITEMS
#define DL_RADIX_ITEMS 8
#endif
constexpr int kRBlock = DL_RADIX_THREADS;
constexpr int kRItems = DL_RADIX_ITEMS;
constexpr int kRTile = kRBlock * kRItems;
constexpr int kRWaves = kRBlock / 64;
constexpr int kRWaveItems = 64 * kRItems;  // elements of one wave's stripe
// Waves per SIMD the downsweep is compiled for: six (78 VGPRs, three
// 512-thread workgroups per CU). Measured at 6M keys / 25 bits: 145 us for the
// sort at six, 152 at eight (64 VGPRs with spills), 152 at four. A
// persistent downsweep (each workgroup sorting tiles in turn, the next
// tile's loads issued before the current tile's writes) took 157 us at two
// workgroups per CU: its loop needed 108 VGPRs.
#ifndef DL_RADIX_WAVES_PER_EU
#define DL_RADIX_WAVES_PER_EU 6
#endif
constexpr int kRWaveSlots = DL_RADIX_WAVES_PER_EU;

__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Block-wide exclusive scan of one value per thread (WAVES waves); returns
// the exclusive prefix and writes the block total to *total (all threads).
template <int WAVES = kWaves>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wave, uint32_t *total) {
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < WAVES; w++) {
    uint32_t c = s_wave[w];
    if (w < wave) wave_off += c;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return wave_off + inc - v;
}

// -------------------------------- scan ------------------------------------
// Element count of a scan level: the static n, or with a device count
// min(n, d_n[0] * mult) (a level sized on the device, capacity n).
__device__ inline size_t level_n(size_t n, const uint32_t *d_n, uint32_t mult) {
  return d_n ? min(n, (size_t)d_n[0] * mult) : n;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t *__restrict__ in, size_t n_cap,
                                                        uint32_t *__restrict__ partials,
                                                        const uint32_t *__restrict__ d_n, uint32_t mult,
                                                        uint32_t *__restrict__ d_next_n) {
  __shared__ uint32_t s_wave[kWaves];
  const size_t n = level_n(n_cap, d_n, mult);
  if (d_next_n && blockIdx.x == 0 && threadIdx.x == 0) *d_next_n = (uint32_t)((n + kTileN - 1) / kTileN);
  const size_t base = (size_t)blockIdx.x * kTileN;
  if (base >= n) return;
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    size_t i = base + (size_t)k * kBlock + threadIdx.x;
    if (i < n) acc += in[i];
  }
  acc = wave_sum(acc);
  if (__lane_id() == 0) s_wave[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kWaves; w++) t += s_wave[w];
    partials[blockIdx.x] = t;
  }
}

// Exclusive scan of one 4096-element tile plus the tile's prefix. With
// SUM_PARTIALS the prefix is the sum of the per-tile totals of all earlier
// tiles (the unscanned partials, read from L2 by every tile); otherwise it is
// prefixes[tile] (already scanned) or 0. The last tile also writes the grand
// total.
template <bool SUM_PARTIALS>
__global__ __launch_bounds__(kBlock) void k_scan_down(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                      size_t n_cap, const uint32_t *__restrict__ prefixes,
                                                      uint32_t *__restrict__ d_total,
                                                      const uint32_t *__restrict__ d_n, uint32_t mult) {
  __shared__ uint32_t s_data[kTileN];
  __shared__ uint32_t s_wave[kWaves];
  const size_t n = level_n(n_cap, d_n, mult);
  const size_t base = (size_t)blockIdx.x * kTileN;
  if (n == 0) {
    if (d_total && blockIdx.x == 0 && threadIdx.x == 0) *d_total = 0;
    return;
  }
  if (base >= n) return;
  uint32_t prefix = 0;
  if (SUM_PARTIALS) {
    uint32_t acc = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kBlock) acc += prefixes[j];
    acc = wave_sum(acc);
    if (__lane_id() == 0) s_wave[threadIdx.x >> 6] = acc;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kWaves; w++) prefix += s_wave[w];
    __syncthreads();
  } else if (prefixes) {
    prefix = prefixes[blockIdx.x];
  }
  const bool last = base + kTileN >= n;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    int li = k * kBlock + threadIdx.x;
    size_t i = base + li;
    s_data[li] = i < n ? in[i] : 0u;
  }
  __syncthreads();
  uint32_t v[kItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    v[k] = s_data[threadIdx.x * kItems + k];
    sum += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, s_wave, &total);
  run += prefix;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    s_data[threadIdx.x * kItems + k] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    int li = k * kBlock + threadIdx.x;
    size_t i = base + li;
    if (i < n) out[i] = s_data[li];
  }
  if (d_total && last && threadIdx.x == 0) *d_total = prefix + total;
}

// ------------------------------- radix ------------------------------------
// A sort block owns one kRTile-element tile. Unsegmented: tile b is elements
// [kRTile b, kRTile (b + 1)) and the count of digit d of block b sits at
// d * nblocks + b. Segmented (SegGeom::seg_start set): segments are
// contiguous element ranges from 0, tiles never straddle a segment, and the
// counts of segment s are laid out [digit][tile of s] from R * blk0[s]
// (R = digits of the pass), so one flat exclusive scan orders (segment,
// digit, tile) and every tile's digit base is its absolute output position.
struct TileSpan {
  size_t i0;        // first element
  uint32_t n;       // elements (0: block unused)
  size_t cbase;     // count index of digit 0
  uint32_t cstride; // count index step per digit
};

__device__ inline TileSpan tile_span(const SegGeom &g, size_t n, uint32_t nblocks, uint32_t b, uint32_t ndig) {
  TileSpan t;
  if (!g.seg_start) {
    t.i0 = (size_t)b * kRTile;
    t.n = t.i0 < n ? (uint32_t)min((size_t)kRTile, n - t.i0) : 0u;
    t.cbase = b;
    t.cstride = nblocks;
    return t;
  }
  t.n = 0;
  t.i0 = 0;
  t.cbase = 0;
  t.cstride = 0;
  if (b >= g.blk0[g.nseg]) return t;
  const uint32_t sg = g.blk_seg[b];
  const uint32_t b0 = g.blk0[sg], nb = g.blk0[sg + 1] - b0, k = b - b0;
  const uint32_t s0 = g.seg_start[sg] + k * (uint32_t)kRTile;
  t.i0 = s0;
  t.n = min((uint32_t)kRTile, g.seg_start[sg + 1] - s0);
  t.cbase = (size_t)ndig * b0 + k;
  t.cstride = nb;
  return t;
}

// The key whose digits a pass sorts on (RadixKeyMap)
__device__ inline uint32_t mapped_key(uint32_t k, const RadixKeyMap &m, uint32_t kmin) {
  return m.kmin ? (k == m.invisible ? m.sentinel : k - kmin) : k;
}

// Tile layout: wave w owns elements [w*kRWaveItems, (w+1)*kRWaveItems) of
// the tile; item k of lane l is element w*kRWaveItems + k*64 + l. Loads are
// coalesced and the (k, lane) processing order is the input order, which
// makes ranks stable.

__global__ __launch_bounds__(kRBlock) void k_radix_upsweep(const uint32_t *__restrict__ keys, size_t n, int shift,
                                                           int bits, uint32_t *__restrict__ counts,
                                                           uint32_t nblocks, SegGeom g, RadixKeyMap km) {
  const uint32_t mask = (1u << bits) - 1u;
  const TileSpan ts = tile_span(g, n, nblocks, blockIdx.x, mask + 1u);
  if (ts.n == 0 && g.seg_start) return;  // unused capacity block
  const uint32_t kmin = km.kmin ? *km.kmin : 0u;
  // all of a thread's loads are issued before its first LDS atomic (loads
  // and atomics interleaved left one load in flight per thread: 14 us per
  // pass against 10 at 6M keys)
  __shared__ uint32_t s_hist[kMaxDigits];
  for (uint32_t d = threadIdx.x; d < kMaxDigits; d += kRBlock) s_hist[d] = 0;
  uint32_t key[kRItems];
#pragma unroll
  for (int k = 0; k < kRItems; k++) {
    const uint32_t li = (uint32_t)k * kRBlock + threadIdx.x;
    key[k] = li < ts.n ? keys[ts.i0 + li] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRItems; k++) {
    const uint32_t li = (uint32_t)k * kRBlock + threadIdx.x;
    if (li < ts.n) atomicAdd(&s_hist[(mapped_key(key[k], km, kmin) >> shift) & mask], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d <= mask; d += kRBlock) counts[ts.cbase + (size_t)d * ts.cstride] = s_hist[d];
}

// Row scan of the digit-major counts (unsegmented layout): workgroup d
// writes the exclusive scan of counts[d nb, (d + 1) nb), each tile's place in
// digit d's run, and the digit's total rowtot[d]. One launch in place of a
// flat scan's two; the downsweep adds the digit bases.
__global__ __launch_bounds__(kBlock) void k_radix_rowscan(const uint32_t *__restrict__ counts,
                                                          uint32_t *__restrict__ offsets, uint32_t nb,
                                                          uint32_t *__restrict__ rowtot) {
  __shared__ uint32_t s_data[kTileN];
  __shared__ uint32_t s_wave[kWaves];
  const size_t row = (size_t)blockIdx.x * nb;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nb; c0 += kTileN) {
    const uint32_t m = min((uint32_t)kTileN, nb - c0);
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      const uint32_t li = (uint32_t)k * kBlock + threadIdx.x;
      s_data[li] = li < m ? counts[row + c0 + li] : 0u;
    }
    __syncthreads();
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      v[k] = s_data[threadIdx.x * kItems + k];
      sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, s_wave, &total) + carry;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      s_data[threadIdx.x * kItems + k] = run;
      run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      const uint32_t li = (uint32_t)k * kBlock + threadIdx.x;
      if (li < m) offsets[row + c0 + li] = s_data[li];
    }
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) rowtot[blockIdx.x] = carry;
}

// Workgroups are dealt to the 8 XCDs in turn (workgroup i to XCD i mod 8).
// A launch of 8 xq workgroups has workgroup i sort tile (i mod 8) xq + i / 8,
// so each XCD sorts a contiguous run of tiles: the runs of one digit that
// neighbouring tiles write are adjacent in the output, and written through
// one L2 they fill whole lines.
__device__ inline uint32_t xcd_tile(uint32_t xq) { return (blockIdx.x & 7u) * xq + (blockIdx.x >> 3); }

// Stable scatter of one radix pass, one tile per workgroup.
// With rowtot set, offsets[] holds each tile's position inside its digit's
// run (the row scan above) and the downsweep adds the digit's base, the sum
// of rowtot over the smaller digits; without it offsets[] are absolute.
template <int BITS>
__global__ __launch_bounds__(kRBlock, kRWaveSlots) void k_radix_downsweep(const uint32_t *__restrict__ keys_in,
                                                            const uint32_t *__restrict__ vals_in,
                                                            uint32_t *__restrict__ keys_out,
                                                            uint32_t *__restrict__ vals_out, size_t n, int shift,
                                                            const uint32_t *__restrict__ offsets,
                                                            const uint32_t *__restrict__ rowtot,
                                                            uint32_t nblocks, SegGeom g, RadixKeyMap km,
                                                            uint32_t xq) {
  constexpr uint32_t kDigits = 1u << BITS;
  constexpr uint32_t kMask = kDigits - 1u;
  // digit slots: one per thread, two per thread for 9-bit digits
  constexpr uint32_t kSlots = kDigits > (uint32_t)kRBlock ? kDigits : (uint32_t)kRBlock;
  constexpr uint32_t kPer = kSlots / kRBlock;
  static_assert(kRWaves * kSlots <= 2 * kRTile, "wave histograms must fit the staging tile");
  // the per-wave digit histograms are dead once every element has its tile
  // position, so the staged tile (keys, then values) overlays them
  __shared__ uint32_t s_stage[2 * kRTile];
  uint32_t (*s_whist)[kSlots] = reinterpret_cast<uint32_t (*)[kSlots]>(s_stage);
  uint32_t *s_keys = s_stage;
  uint32_t *s_vals = s_stage + kRTile;
  __shared__ uint32_t s_start[kSlots];
  __shared__ uint32_t s_gbase[kSlots];
  __shared__ uint32_t s_wave[kRWaves];

  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const uint32_t b = xcd_tile(xq);
  if (b >= nblocks) return;
  const TileSpan ts = tile_span(g, n, nblocks, b, kDigits);
  if (ts.n == 0) return;
  const size_t tile0 = ts.i0;
  const uint32_t tile_n = ts.n;
  const uint32_t kmin = km.kmin ? *km.kmin : 0u;

  for (uint32_t d = threadIdx.x; d < kRWaves * kSlots; d += kRBlock) (&s_whist[0][0])[d] = 0;
  // this tile's digit offsets and the digit totals do not depend on the
  // keys: their loads go out with the keys' (thread t owns digits
  // t*kPer .. t*kPer + kPer - 1, in digit order)
  uint32_t goff[kPer], dtot[kPer];
#pragma unroll
  for (uint32_t j = 0; j < kPer; j++) {
    const uint32_t d = threadIdx.x * kPer + j;
    goff[j] = d < kDigits ? offsets[ts.cbase + (size_t)d * ts.cstride] : 0u;
    dtot[j] = (rowtot && d < kDigits) ? rowtot[d] : 0u;
  }
  __syncthreads();

  uint32_t key[kRItems], val[kRItems], rank[kRItems];
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < kRItems; k++) {
    const uint32_t li = (uint32_t)wave * kRWaveItems + (uint32_t)k * 64u + (uint32_t)lane;
    const bool valid = li < tile_n;
    key[k] = valid ? keys_in[tile0 + li] : 0u;
    // without vals_in the values are the elements' indices
    val[k] = valid ? (vals_in ? vals_in[tile0 + li] : (uint32_t)(tile0 + li)) : 0u;
    const uint32_t d = (mapped_key(key[k], km, kmin) >> shift) & kMask;
    uint64_t peers = __ballot(valid);
    if (!valid) peers = 0;
#pragma unroll
    for (int b = 0; b < BITS; b++) {
      const int set = (d >> b) & 1u;
      const uint64_t bal = __ballot(set);
      peers &= set ? bal : ~bal;
    }
    const int leader = peers ? (__ffsll((long long)peers) - 1) : 0;
    // one leader per distinct digit: its add does not race within the wave
    uint32_t old = 0;
    if (valid && lane == leader) old = atomicAdd(&s_whist[wave][d], (uint32_t)__popcll(peers));
    old = (uint32_t)__shfl((int)old, leader, 64);
    rank[k] = old + (uint32_t)__popcll(peers & lt);
  }
  __syncthreads();
  // per digit: wave offsets and block count, then block-wide digit starts
  {
    uint32_t cnt[kPer];
    uint32_t tsum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      const uint32_t d = threadIdx.x * kPer + j;
      uint32_t c = 0;
#pragma unroll
      for (int w = 0; w < kRWaves; w++) {
        const uint32_t x = s_whist[w][d];
        s_whist[w][d] = c;
        c += x;
      }
      cnt[j] = d < kDigits ? c : 0u;
      tsum += cnt[j];
    }
    uint32_t total;
    uint32_t start = block_exclusive_scan<kRWaves>(tsum, s_wave, &total);
    // digit bases: exclusive scan of the digit totals (row-local offsets)
    uint32_t dsum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) dsum += dtot[j];
    uint32_t dbase = rowtot ? block_exclusive_scan<kRWaves>(dsum, s_wave, &total) : 0u;
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) {
      const uint32_t d = threadIdx.x * kPer + j;
      s_start[d] = start;
      start += cnt[j];
      s_gbase[d] = dbase + goff[j];
      dbase += dtot[j];
    }
  }
  __syncthreads();
  // tile positions first (the last reads of the wave histograms), then the
  // staged tile overwrites them
#pragma unroll
  for (int k = 0; k < kRItems; k++) {
    const uint32_t li = (uint32_t)wave * kRWaveItems + (uint32_t)k * 64u + (uint32_t)lane;
    if (li < tile_n) {
      const uint32_t d = (mapped_key(key[k], km, kmin) >> shift) & kMask;
      rank[k] += s_start[d] + s_whist[wave][d];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRItems; k++) {
    const uint32_t li = (uint32_t)wave * kRWaveItems + (uint32_t)k * 64u + (uint32_t)lane;
    if (li < tile_n) {
      s_keys[rank[k]] = key[k];
      s_vals[rank[k]] = val[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRItems; k++) {
    const uint32_t p = (uint32_t)k * kRBlock + threadIdx.x;
    if (p < tile_n) {
      const uint32_t kk = s_keys[p];
      const uint32_t d = (mapped_key(kk, km, kmin) >> shift) & kMask;
      const size_t dst = (size_t)s_gbase[d] + (p - s_start[d]);
      keys_out[dst] = kk;
      vals_out[dst] = s_vals[p];
    }
  }
}

// off[b] = lower_bound(keys, b) for b < num_bins, off[num_bins] = n; every
// entry is written exactly once (keys sorted, all < num_bins, n >= 1).
__global__ __launch_bounds__(256) void k_bin_ranges(const uint32_t *__restrict__ keys, uint32_t n,
                                                    uint32_t num_bins, uint32_t *__restrict__ off) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t cur = keys[i];
  if (i == 0) {
    for (uint32_t t = 0; t <= cur; t++) off[t] = 0;
  } else {
    const uint32_t prev = keys[i - 1];
    for (uint32_t t = prev + 1; t <= cur; t++) off[t] = i;
  }
  if (i == n - 1)
    for (uint32_t t = cur + 1; t <= num_bins; t++) off[t] = n;
}

// The same offsets, four keys per thread from one 16-byte load (keys
// 16-byte aligned): each boundary between consecutive keys writes the bins
// it passes, the first and last keys the bins before and after them.
__global__ __launch_bounds__(256) void k_bin_ranges4(const uint32_t *__restrict__ keys, uint32_t n,
                                                     uint32_t num_bins, uint32_t *__restrict__ off) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  const uint32_t i0 = 4u * j;
  if (i0 >= n) return;
  uint32_t k[4];
  if (i0 + 4u <= n) {
    const uint4 v = reinterpret_cast<const uint4 *>(keys)[j];
    k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w;
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) k[q] = i0 + q < n ? keys[i0 + q] : 0u;
  }
  uint32_t prev = i0 > 0 ? keys[i0 - 1] : 0u;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint32_t i = i0 + (uint32_t)q;
    if (i >= n) break;
    const uint32_t cur = k[q];
    const uint32_t t0 = i == 0 ? 0u : prev + 1u;
    for (uint32_t t = t0; t <= cur; t++) off[t] = i;
    if (i == n - 1)
      for (uint32_t t = cur + 1; t <= num_bins; t++) off[t] = n;
    prev = cur;
  }
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Segment tables: units[s] = ceil(len(s) / unit), then (after a scan into
// first[]) owner[u] = s for every unit u of segment s.
__global__ __launch_bounds__(256) void k_iota_u32(uint32_t *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_seg_units(const uint32_t *__restrict__ seg_start, uint32_t nseg, uint32_t unit,
                                                   uint32_t *__restrict__ units) {
  const uint32_t sg = blockIdx.x * 256 + threadIdx.x;
  if (sg < nseg) units[sg] = (seg_start[sg + 1] - seg_start[sg] + unit - 1) / unit;
}

__global__ __launch_bounds__(256) void k_unit_owner(const uint32_t *__restrict__ first, uint32_t nseg,
                                                    uint32_t *__restrict__ owner) {
  const uint32_t sg = blockIdx.x * 256 + threadIdx.x;
  if (sg >= nseg) return;
  for (uint32_t u = first[sg]; u < first[sg + 1]; u++) owner[u] = sg;
}

using DownFn = void (*)(const uint32_t *, const uint32_t *, uint32_t *, uint32_t *, size_t, int,
                        const uint32_t *, const uint32_t *, uint32_t, SegGeom, RadixKeyMap, uint32_t);

DownFn downsweep_for(int bits) {
  switch (bits) {
    case 1: return k_radix_downsweep<1>;
    case 2: return k_radix_downsweep<2>;
    case 3: return k_radix_downsweep<3>;
    case 4: return k_radix_downsweep<4>;
    case 5: return k_radix_downsweep<5>;
    case 6: return k_radix_downsweep<6>;
    case 7: return k_radix_downsweep<7>;
    case 9: return k_radix_downsweep<9>;
    default: return k_radix_downsweep<8>;
  }
}

// per level: the partials (256-byte aligned) and a 256-byte slot for the
// level's device count
size_t scan_levels_bytes(size_t n) {
  size_t bytes = 0;
  while (n > (size_t)kTileN) {
    n = div_up(n, kTileN);
    bytes += ((n * sizeof(uint32_t) + 255) / 256) * 256 + 256;
  }
  return bytes + 256;
}

// Reduce-then-scan, recursive over the per-tile partials. With d_n the
// element count is read on the device (capacity n, see level_n) and every
// launch is sized by the capacity, so no host synchronisation is needed.
void scan_rec(hipStream_t s, const uint32_t *in, uint32_t *out, size_t n, uint32_t *d_total, uint8_t *ws,
              const uint32_t *d_n = nullptr, uint32_t mult = 1) {
  if (n == 0) {
    // a fill kernel rather than a memset node, so that the scan can be captured in a graph
    if (d_total) k_fill_u32<<<1, 256, 0, s>>>(d_total, 1, 0u);
    DL_HIP(hipGetLastError());
    return;
  }
  if (n <= (size_t)kTileN) {
    k_scan_down<false><<<1, kBlock, 0, s>>>(in, out, n, nullptr, d_total, d_n, mult);
    DL_HIP(hipGetLastError());
    return;
  }
  const uint32_t nb = div_up(n, kTileN);
  uint32_t *partials = reinterpret_cast<uint32_t *>(ws);
  uint8_t *p = ws + ((nb * sizeof(uint32_t) + 255) / 256) * 256;
  uint32_t *d_nb = d_n ? reinterpret_cast<uint32_t *>(p) : nullptr;
  uint8_t *next = p + 256;
  k_scan_reduce<<<nb, kBlock, 0, s>>>(in, n, partials, d_n, mult, d_nb);
  DL_HIP(hipGetLastError());
  if (nb <= (uint32_t)kSumPartialsMax) {
    // two launches: every tile sums the totals of the tiles before it
    k_scan_down<true><<<nb, kBlock, 0, s>>>(in, out, n, partials, d_total, d_n, mult);
  } else {
    scan_rec(s, partials, partials, nb, nullptr, next, d_nb, 1);
    k_scan_down<false><<<nb, kBlock, 0, s>>>(in, out, n, partials, d_total, d_n, mult);
  }
  DL_HIP(hipGetLastError());
}

// Fewest passes of at most max_bits, with widths as even as possible: 25
// bits at 9 are 8 + 8 + 9, not 9 + 9 + 7, since a 9-bit pass costs more than
// an 8-bit one. The wider digits are the high ones: high digits of depth keys
// take few values, so a wide high pass scatters into long runs, while a wide
// low pass splits every tile into 512 short runs of partly written lines.
void pass_layout(int key_bits, int max_bits, int *passes, int bits[32]) {
  max_bits = std::min(kMaxDigitBits, std::max(1, max_bits));
  const int p = std::max(1, (key_bits + max_bits - 1) / max_bits);
  const int base = key_bits / p, wide = key_bits % p;
  for (int i = 0; i < p; i++) bits[i] = base + (i >= p - wide ? 1 : 0);
  *passes = p;
}

}  // namespace

size_t scan_workspace_bytes(size_t n) { return scan_levels_bytes(n); }

void exclusive_scan_u32(hipStream_t s, const uint32_t *d_in, uint32_t *d_out, size_t n, uint32_t *d_total,
                        void *ws) {
  scan_rec(s, d_in, d_out, n, d_total, static_cast<uint8_t *>(ws));
}

void exclusive_scan_u32_dev(hipStream_t s, const uint32_t *d_in, uint32_t *d_out, size_t n_cap, const uint32_t *d_n,
                            uint32_t mult, uint32_t *d_total, void *ws) {
  scan_rec(s, d_in, d_out, n_cap, d_total, static_cast<uint8_t *>(ws), d_n, mult);
}

void bin_ranges(hipStream_t s, const uint32_t *d_keys, uint32_t n, uint32_t num_bins, uint32_t *d_off) {
  if (n == 0) {
    k_fill_u32<<<div_up((size_t)num_bins + 1, 256), 256, 0, s>>>(d_off, (size_t)num_bins + 1, 0u);
  } else if ((reinterpret_cast<uintptr_t>(d_keys) & 15u) == 0) {
    k_bin_ranges4<<<div_up(div_up(n, 4), 256), 256, 0, s>>>(d_keys, n, num_bins, d_off);
  } else {
    k_bin_ranges<<<div_up(n, 256), 256, 0, s>>>(d_keys, n, num_bins, d_off);
  }
  DL_HIP(hipGetLastError());
}

// Downsweep launch over nb tiles: 8 xq workgroups, xq tiles per XCD
static inline uint32_t downsweep_xq(uint32_t nb) { return div_up(nb, 8); }
static inline uint32_t downsweep_grid(uint32_t nb) { return 8u * div_up(nb, 8); }

// counts and offsets ([digit][tile], 512 digits) and the digit totals
size_t radix_workspace_bytes(size_t n, int key_bits) {
  (void)key_bits;
  const size_t nb = std::max<size_t>(1, div_up(n, kRTile));
  const size_t counts = ((kMaxDigits * nb * sizeof(uint32_t) + 255) / 256) * 256;
  return counts * 2 + kMaxDigits * sizeof(uint32_t);
}

bool radix_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys, uint32_t *tmp_vals,
                      size_t n, int key_bits, void *ws, int max_digit_bits, const RadixKeyMap &map,
                      bool index_vals, int low_bit) {
  if (key_bits > 32) key_bits = 32;
  if (low_bit < 0) low_bit = 0;
  if (n <= 1 || key_bits <= low_bit) {
    // nothing to reorder; index values are still owed: value i is i
    if (index_vals && n > 0) {
      k_iota_u32<<<div_up(n, 256), 256, 0, s>>>(vals, n);
      DL_HIP(hipGetLastError());
    }
    return false;
  }
  const uint32_t nb = div_up(n, kRTile);
  uint8_t *w = static_cast<uint8_t *>(ws);
  const size_t counts_bytes = ((kMaxDigits * (size_t)nb * sizeof(uint32_t) + 255) / 256) * 256;
  uint32_t *counts = reinterpret_cast<uint32_t *>(w);
  uint32_t *offsets = reinterpret_cast<uint32_t *>(w + counts_bytes);
  uint32_t *rowtot = reinterpret_cast<uint32_t *>(w + 2 * counts_bytes);
  int passes, bits[32];
  pass_layout(key_bits - low_bit, max_digit_bits, &passes, bits);
  uint32_t *ki = keys, *vi = vals, *ko = tmp_keys, *vo = tmp_vals;
  int shift = low_bit;
  const SegGeom none{};
  for (int p = 0; p < passes; p++) {
    const uint32_t ndig = 1u << bits[p];
    k_radix_upsweep<<<nb, kRBlock, 0, s>>>(ki, n, shift, bits[p], counts, nb, none, map);
    DL_HIP(hipGetLastError());
    k_radix_rowscan<<<ndig, kBlock, 0, s>>>(counts, offsets, nb, rowtot);
    DL_HIP(hipGetLastError());
    downsweep_for(bits[p])<<<downsweep_grid(nb), kRBlock, 0, s>>>(ki, (p == 0 && index_vals) ? nullptr : vi, ko,
                                                                  vo, n, shift, offsets, rowtot, nb, none, map,
                                                                  downsweep_xq(nb));
    DL_HIP(hipGetLastError());
    std::swap(ki, ko);
    std::swap(vi, vo);
    shift += bits[p];
  }
  return ki == tmp_keys;
}

int radix_passes(int key_bits, int max_digit_bits) {
  if (key_bits <= 0) return 0;
  int passes, bits[32];
  pass_layout(std::min(32, key_bits), max_digit_bits, &passes, bits);
  return passes;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t seg_radix_workspace_bytes(size_t max_elems, uint32_t nseg) {
  const size_t mb = (size_t)div_up(max_elems, kRTile) + nseg + 1;
  return align256((size_t)nseg * 4 + 4) + align256(((size_t)nseg + 1) * 4) + align256(mb * 4) +
         2 * align256(256 * mb * 4) + std::max(scan_levels_bytes(nseg), scan_levels_bytes(256 * mb));
}

bool seg_radix_sort_pairs(hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *tmp_keys, uint32_t *tmp_vals,
                          const uint32_t *d_seg_start, uint32_t nseg, size_t max_elems, int key_bits,
                          int max_digit_bits, void *ws, bool index_vals) {
  if (nseg == 0 || key_bits <= 0) {
    if (index_vals && max_elems > 0) {
      k_iota_u32<<<div_up(max_elems, 256), 256, 0, s>>>(vals, max_elems);
      DL_HIP(hipGetLastError());
    }
    return false;
  }
  if (key_bits > 32) key_bits = 32;
  const uint32_t mb = div_up(max_elems, kRTile) + nseg + 1;
  uint8_t *w = static_cast<uint8_t *>(ws);
  uint32_t *units = reinterpret_cast<uint32_t *>(w);
  w += align256((size_t)nseg * 4 + 4);
  uint32_t *blk0 = reinterpret_cast<uint32_t *>(w);
  w += align256(((size_t)nseg + 1) * 4);
  uint32_t *blk_seg = reinterpret_cast<uint32_t *>(w);
  w += align256((size_t)mb * 4);
  uint32_t *counts = reinterpret_cast<uint32_t *>(w);
  w += align256(256 * (size_t)mb * 4);
  uint32_t *offsets = reinterpret_cast<uint32_t *>(w);
  w += align256(256 * (size_t)mb * 4);
  uint8_t *scan_ws = w;
  // tile table: blk0[s] = first tile of segment s, blk0[nseg] = tiles in use
  k_seg_units<<<div_up(nseg, 256), 256, 0, s>>>(d_seg_start, nseg, (uint32_t)kRTile, units);
  DL_HIP(hipGetLastError());
  scan_rec(s, units, blk0, nseg, blk0 + nseg, scan_ws);
  k_unit_owner<<<div_up(nseg, 256), 256, 0, s>>>(blk0, nseg, blk_seg);
  DL_HIP(hipGetLastError());
  const SegGeom g{d_seg_start, blk0, blk_seg, nseg};
  int passes, bits[32];
  // the segmented counts are sized for 8-bit digits
  pass_layout(key_bits, std::min(max_digit_bits, 8), &passes, bits);
  uint32_t *ki = keys, *vi = vals, *ko = tmp_keys, *vo = tmp_vals;
  int shift = 0;
  for (int p = 0; p < passes; p++) {
    const size_t ndig = (size_t)1 << bits[p];
    k_radix_upsweep<<<mb, kRBlock, 0, s>>>(ki, 0, shift, bits[p], counts, mb, g, RadixKeyMap{});
    DL_HIP(hipGetLastError());
    scan_rec(s, counts, offsets, ndig * mb, nullptr, scan_ws);
    downsweep_for(bits[p])<<<downsweep_grid(mb), kRBlock, 0, s>>>(ki, (p == 0 && index_vals) ? nullptr : vi, ko,
                                                                  vo, 0, shift, offsets, nullptr, mb, g,
                                                                  RadixKeyMap{}, downsweep_xq(mb));
    DL_HIP(hipGetLastError());
    std::swap(ki, ko);
    std::swap(vi, vo);
    shift += bits[p];
  }
  return ki == tmp_keys;
}

}  // namespace dl
