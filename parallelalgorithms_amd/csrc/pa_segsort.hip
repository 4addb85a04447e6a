// pa_segsort.hip -- segmented (batched) LSD radix sort of many u32 arrays in
// one batch on gfx950.
//
// No reference counterpart: the reference sorts one array per call
// (SortRadixPar, RadixSortLsdParallel.h:367-387) and only benchmarks bundling
// many small tasks (BundleSmallWorkItems.cpp:36-144).  Each segment's result
// equals SortRadixPar on that segment (a stable sort of scalar keys is unique).
//
// Size classes:
//   len <= 1              nothing to do.
//   len <= 16384          one workgroup sorts the segment entirely in LDS
//                         (4 stable 8-bit passes over LDS, one read and one
//                         write of HBM per key), in three workgroup sizes
//                         (256, 512, 1024 threads x 16 keys) so a segment
//                         pays for at most twice its length in LDS slots.
//                         One launch per size for all of them.
//   len <= 65536          a persistent kernel: workgroups pull whole segments
//                         (longest first) from a device counter and sort each
//                         alone, streaming it through LDS in 8192-key chunks
//                         with running digit offsets (no look-back), when the
//                         batch holds enough of them to balance the
//                         workgroups; the longest ones otherwise go to the
//                         onesweep class.
//   larger                segmented onesweep: one fused read builds every
//                         segment's 4 digit histograms, then one onesweep
//                         launch per digit covers every tile of every large
//                         segment; each segment has its own look-back chain,
//                         and tiles are claimed in groups of a few segments.
// So the whole batch costs at most 10 launches however many segments it
// holds, where per-array calls cost 6 launches per array.
#include <algorithm>
#include <cstdio>
#include <string>
#include <atomic>
#include <cstdlib>
#include <vector>

#include "pa_sort_kernels.h"

namespace pa {

// LDS classes: segments of up to kLdsLimit[c] keys are sorted by one
// workgroup of kLdsLimit[c] / kSmallItems threads.
constexpr int kSmallItems = 16;
constexpr int kLdsClasses = 3;
constexpr uint32_t kLdsLimit[kLdsClasses] = {4096, 8192, 16384};
constexpr uint32_t kLdsMax = kLdsLimit[kLdsClasses - 1];
constexpr uint64_t kQueueDefaultMax = 65536;

// Block of the segmented histogram read (one tile per block).
constexpr int kSegBlock = 512;

// Shapes of the segmented sort (the PA_TUNE_SEG_U32 tuning target): the
// onesweep class's tile (block x items, staged in `split` rounds, ranked in
// 16-lane units) and digit width, and the longest segment of the queue
// class.  Widths of 8, 6 and 4 bits give 4, 6 and 8 passes: an even count
// leaves every large segment's result in place without a copy back.
struct SegPassArgs {
  const uint32_t* counts;
  uint32_t* lb;
  uint32_t* ctr;
  uint32_t ntiles, shift, gstride;
  const uint64_t *tile_info, *tin, *tout;
  const uint32_t *len, *row;
};
struct SegVariant {
  int block, items, lbw, split, rb;
  uint64_t queue_max;
  int (*pass)(hipStream_t s, const SegPassArgs& a);
  int (*hist)(hipStream_t s, const uint64_t* tile_info, const uint64_t* seg_in, const uint32_t* seg_len,
              uint32_t ntiles, uint32_t tile, uint32_t* counts, uint4* zero_ptr, size_t zero_vec);
  uint32_t tile() const { return (uint32_t)(block * items); }
  int passes() const { return (32 + rb - 1) / rb; }
  uint32_t bins() const { return 1u << rb; }
};
// Claim order: the large segments are cut into consecutive groups of at least
// kSegGroupTiles tiles.  Groups are claimed in order; inside a group, tile k
// of a segment with T tiles goes to bucket floor((2k + 1) * kSegBuckets / (2T)),
// buckets in order, so a short segment's tiles are spread over its group and
// a tile's predecessor in its segment was claimed well before it, while the
// tiles in flight at any time come from a few groups only (spreading every
// segment over the whole launch keeps tiles of hundreds of segments in flight
// and was measured 25% slower on 10,000 arrays).  A segment of
// kSegGroupTiles tiles or more is a group of its own and is claimed in order,
// as the plain sort claims its tiles.
constexpr uint32_t kSegBuckets = 1024;
constexpr uint64_t kSegGroupTiles = 64;

// Stable ranks within a wave: rank[i] = the wave's earlier keys with the same
// digit (earlier items, then lower lanes of the same item) plus wh[digit] on
// entry; wh[d] is advanced by the wave's count of d.  Peers of one digit are
// found by every lane ORing its bit into the wave's LDS slot for the digit;
// the lowest peer clears the slot, so wm is all zero on entry and on exit.
template <int ITEMS>
__device__ __forceinline__ void wave_rank(const uint32_t (&key)[ITEMS], uint32_t shift, uint32_t* __restrict__ wh,
                                          uint64_t* __restrict__ wm, uint32_t lane, uint32_t (&rank)[ITEMS]) {
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t d = (key[i] >> shift) & 0xFFu;
    atomicOr(reinterpret_cast<unsigned long long*>(&wm[d]), 1ull << lane);
    const uint64_t peers = wm[d];
    const uint32_t below = dev::count_below(peers);
    const uint32_t base = wh[d];
    rank[i] = base + below;
    if (below == 0) {
      wh[d] = base + (uint32_t)__popcll(peers);
      wm[d] = 0ull;
    }
  }
}

// Turns the per-wave digit counts s_whist[w * 256 + d] into the start of
// (wave w, digit d) in the block's digit-sorted order; when s_dstart is given
// it receives each digit's start.  Threads 0..255 take one digit each.  Call
// after a barrier that follows the counting; ends with a barrier.
template <int WAVES>
__device__ __forceinline__ void block_digit_starts(uint32_t* __restrict__ s_whist, uint32_t* __restrict__ s_wsum,
                                                   uint32_t* __restrict__ s_dstart) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t cnt = 0, inc = 0;
  if (tid < 256) {
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = s_whist[w * 256 + tid];
      s_whist[w * 256 + tid] = cnt;
      cnt += c;
    }
    inc = dev::wave_incl_scan(cnt, lane);
    if (lane == 63) s_wsum[wave] = inc;
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t pre = 0;
    for (uint32_t w = 0; w < wave; ++w) pre += s_wsum[w];
    const uint32_t start = pre + inc - cnt;
    if (s_dstart) s_dstart[tid] = start;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) s_whist[w * 256 + tid] += start;
  }
  __syncthreads();
}

// Whole-segment LDS sort.  Padding slots hold 0xFFFFFFFF: a stable sort
// keeps them after every real key, so the first len slots are the answer.
template <int BLOCK>
__device__ __forceinline__ void lds_sort_segment(uint32_t* __restrict__ seg, uint32_t n) {
  constexpr int WAVES = BLOCK / 64;
  constexpr int WAVE_SPAN = 64 * kSmallItems;
  constexpr int TILE = BLOCK * kSmallItems;
  static_assert(BLOCK >= 256, "one thread per digit in the scan");
  __shared__ uint32_t s_keys[TILE];
  __shared__ uint32_t s_whist[WAVES * 256];
  __shared__ uint64_t s_match[WAVES * 256];
  __shared__ uint32_t s_wsum[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t key[kSmallItems];
  const uint32_t wbase = wave * WAVE_SPAN + lane;
#pragma unroll
  for (int i = 0; i < kSmallItems; ++i) {
    const uint32_t idx = wbase + i * 64;
    key[i] = idx < n ? seg[idx] : 0xFFFFFFFFu;
  }
  uint32_t* wh = s_whist + wave * 256;
  uint64_t* wm = s_match + wave * 256;
  for (uint32_t i = tid; i < WAVES * 256; i += BLOCK) s_match[i] = 0ull;
  for (uint32_t shift = 0; shift < 32; shift += 8) {
    for (uint32_t i = tid; i < WAVES * 256; i += BLOCK) s_whist[i] = 0;
    __syncthreads();
    uint32_t rank[kSmallItems];
    wave_rank<kSmallItems>(key, shift, wh, wm, lane, rank);
    __syncthreads();
    block_digit_starts<WAVES>(s_whist, s_wsum, nullptr);
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) s_keys[wh[(key[i] >> shift) & 0xFFu] + rank[i]] = key[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSmallItems; ++i) key[i] = s_keys[wbase + i * 64];
    __syncthreads();
  }
  for (uint32_t j = tid; j < n; j += BLOCK) seg[j] = s_keys[j];
}

// Mid-size class: a persistent kernel.  Each workgroup pulls the next segment
// from a device counter over a list ordered by length (longest first) and
// sorts it alone: one read builds the segment's 4 digit histograms, then each
// of the 4 passes streams the segment in chunks of kQueueChunk keys in input
// order.  A chunk is ranked by the pass's digit in LDS, staged in digit order
// and written to the running offset of each digit, which the workgroup keeps
// in LDS, so the passes need no cross-workgroup prefix at all.  Ping-pong
// between the segment and its tmp; after 4 passes the result is back in the
// segment.  Every workgroup leaves once the counter passes the list.
constexpr int kQueueBlock = 512;
constexpr uint32_t kQueueChunk = kQueueBlock * kSmallItems;
constexpr size_t kQueueGrid = 512;  // two workgroups per CU on the 256 CUs of an MI355X

// Longest segment served by the queue: the shape's, unless PA_SEG_QUEUE_MAX
// overrides it (read per plan, so sweeps can change it between calls); at
// most kLdsMax disables the class.
struct SegVariant;
static uint64_t seg_queue_max(const SegVariant& v);

__global__ __launch_bounds__(kQueueBlock) void k_seg_queue(const uint64_t* __restrict__ seg_ptr,
                                                           const uint64_t* __restrict__ seg_tmp,
                                                           const uint32_t* __restrict__ seg_len, uint32_t nseg,
                                                           uint32_t* __restrict__ next) {
  constexpr int WAVES = kQueueBlock / 64;
  constexpr int WAVE_SPAN = 64 * kSmallItems;
  __shared__ uint32_t s_keys[kQueueChunk];
  __shared__ uint32_t s_whist[WAVES * 256];
  __shared__ uint64_t s_match[WAVES * 256];
  __shared__ uint32_t s_cnt[4 * 256];
  __shared__ uint32_t s_off[256];
  __shared__ uint32_t s_dstart[256];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_seg;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t* wh = s_whist + wave * 256;
  uint64_t* wm = s_match + wave * 256;
  for (uint32_t i = tid; i < WAVES * 256; i += kQueueBlock) s_match[i] = 0ull;
  for (;;) {
    if (tid == 0) s_seg = atomicAdd(next, 1u);
    for (uint32_t i = tid; i < 4 * 256; i += kQueueBlock) s_cnt[i] = 0;
    __syncthreads();
    const uint32_t g = s_seg;
    if (g >= nseg) break;  // uniform: every thread read the same s_seg
    uint32_t* src = reinterpret_cast<uint32_t*>(seg_ptr[g]);
    uint32_t* dst = reinterpret_cast<uint32_t*>(seg_tmp[g]);
    const uint32_t m = seg_len[g];
    // The segment's 4 digit histograms in one read.
    for (uint32_t i0 = 0; i0 < m; i0 += 4 * kQueueBlock) {
      uint32_t x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t idx = i0 + u * kQueueBlock + tid;
        x[u] = idx < m ? src[idx] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (i0 + u * kQueueBlock + tid >= m) continue;
#pragma unroll
        for (int d = 0; d < 4; ++d) atomicAdd(&s_cnt[d * 256 + ((x[u] >> (8 * d)) & 0xFFu)], 1u);
      }
    }
    __syncthreads();
    for (uint32_t shift = 0; shift < 32; shift += 8) {
      // Running output offset of every digit: exclusive scan of its counts.
      if (tid < 256) {
        const uint32_t c = s_cnt[(shift / 8) * 256 + tid];
        const uint32_t inc = dev::wave_incl_scan(c, lane);
        if (lane == 63) s_wsum[wave] = inc;
        s_off[tid] = inc - c;
      }
      __syncthreads();
      if (tid < 256)
        for (uint32_t w = 0; w < wave; ++w) s_off[tid] += s_wsum[w];
      for (uint32_t c0 = 0; c0 < m; c0 += kQueueChunk) {
        const uint32_t cl = min(kQueueChunk, m - c0);
        uint32_t key[kSmallItems];
        const uint32_t wbase = wave * WAVE_SPAN + lane;
#pragma unroll
        for (int i = 0; i < kSmallItems; ++i) {
          const uint32_t idx = wbase + i * 64;
          key[i] = idx < cl ? src[c0 + idx] : 0xFFFFFFFFu;
        }
        for (uint32_t i = tid; i < WAVES * 256; i += kQueueBlock) s_whist[i] = 0;
        __syncthreads();
        uint32_t rank[kSmallItems];
        wave_rank<kSmallItems>(key, shift, wh, wm, lane, rank);
        __syncthreads();
        block_digit_starts<WAVES>(s_whist, s_wsum, s_dstart);
#pragma unroll
        for (int i = 0; i < kSmallItems; ++i) s_keys[wh[(key[i] >> shift) & 0xFFu] + rank[i]] = key[i];
        __syncthreads();
        // Padding keys (digit 255, after every real key) fill chunk slots
        // [cl, kQueueChunk), so j < cl writes exactly the real keys.
        for (uint32_t j = tid; j < cl; j += kQueueBlock) {
          const uint32_t k = s_keys[j];
          const uint32_t d = (k >> shift) & 0xFFu;
          dst[s_off[d] + (j - s_dstart[d])] = k;
        }
        __syncthreads();
        if (tid < 256) s_off[tid] += (tid < 255 ? s_dstart[tid + 1] : kQueueChunk) - s_dstart[tid];
      }
      __syncthreads();
      uint32_t* t = src;
      src = dst;
      dst = t;
    }
  }
}

// One segment per workgroup, from the plan's (address, length) lists.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_seg_small(const uint64_t* __restrict__ ptr,
                                                     const uint32_t* __restrict__ len) {
  lds_sort_segment<BLOCK>(reinterpret_cast<uint32_t*>(ptr[blockIdx.x]), len[blockIdx.x]);
}

// A single small array (plain sort of up to kLdsMax keys): one launch.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sort_small(uint32_t* __restrict__ keys, uint32_t n) {
  lds_sort_segment<BLOCK>(keys, n);
}

// Clears the segments' digit counts and the 16 tile / queue counters (one
// per digit pass, up to 8, then the queue's).
__global__ __launch_bounds__(256) void k_seg_prepare(uint32_t* __restrict__ counts, size_t ncounts,
                                                     uint32_t* __restrict__ counters) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncounts; i += stride) counts[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 16) counters[threadIdx.x] = 0;
}

// One tile of `tile` keys per block: the D digit histograms (RB-bit digits)
// of the tile's keys into its segment's counts [D][R]; also clears the
// look-back table.
template <int RB>
__global__ __launch_bounds__(kSegBlock) void k_seg_hist(const uint64_t* __restrict__ tile_info,
                                                        const uint64_t* __restrict__ seg_in,
                                                        const uint32_t* __restrict__ seg_len, uint32_t tile,
                                                        uint32_t* __restrict__ counts, uint4* __restrict__ zero_ptr,
                                                        size_t zero_vec) {
  constexpr uint32_t R = 1u << RB, D = (32 + RB - 1) / RB;
  __shared__ uint32_t s_hist[D * R];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < D * R; i += kSegBlock) s_hist[i] = 0;
  const size_t stride = (size_t)gridDim.x * kSegBlock;
  for (size_t i = (size_t)blockIdx.x * kSegBlock + tid; i < zero_vec; i += stride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  const uint64_t info = tile_info[blockIdx.x];
  const uint32_t seg = (uint32_t)(info >> 32), local = (uint32_t)info;
  const uint32_t* k = reinterpret_cast<const uint32_t*>(seg_in[seg]);
  const uint32_t n = seg_len[seg];
  const uint32_t lo = local * tile, hi = min(lo + tile, n);
  // Eight loads in flight per thread before any counting (segments start at
  // any offset, so the loads are 4-byte, coalesced across the wave).
  constexpr int U = 8;
  for (uint32_t i0 = lo + tid; i0 < hi; i0 += U * kSegBlock) {
    uint32_t x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t idx = i0 + u * kSegBlock;
      x[u] = idx < hi ? k[idx] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i0 + u * kSegBlock >= hi) break;  // the indices grow with u
#pragma unroll
      for (uint32_t d = 0; d < D; ++d) {
        const uint32_t dg[1] = {dev::digit_bits<uint32_t, RB>(x[u], RB * d)};
        dev::lds_count<1>(s_hist + d * R, dg);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < D * R; i += kSegBlock)
    if (s_hist[i]) atomicAdd(&counts[(size_t)seg * D * R + i], s_hist[i]);
}

template <int RB>
int launch_seg_hist(hipStream_t s, const uint64_t* tile_info, const uint64_t* seg_in, const uint32_t* seg_len,
                    uint32_t ntiles, uint32_t tile, uint32_t* counts, uint4* zero_ptr, size_t zero_vec) {
  ProfScope ps("k_seg_hist", s);
  hipLaunchKernelGGL(k_seg_hist<RB>, dim3(ntiles), dim3(kSegBlock), 0, s, tile_info, seg_in, seg_len, tile, counts,
                     zero_ptr, zero_vec);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

template <int B, int I, int LBW, int SPLIT, int RB>
int launch_seg_pass(hipStream_t s, const SegPassArgs& a) {
  ProfScope ps("k_onesweep_seg", s);
  hipLaunchKernelGGL((k_onesweep<uint32_t, false, B, I, LBW, kModeSeg, true, 4, true, SPLIT, 16, RB>), dim3(a.ntiles),
                     dim3(B), 0, s, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)nullptr,
                     (uint32_t*)nullptr, 0u, a.counts, a.lb, a.ctr, a.ntiles, a.shift, 0u, a.tile_info, a.tin, a.tout,
                     a.len, a.gstride, a.row);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

#define PA_SEG_VARIANT(B, I, W, S, RB, Q) \
  { B, I, W, S, RB, Q, &launch_seg_pass<B, I, W, S, RB>, &launch_seg_hist<RB> }
// Entry 0 is the built-in shape: the plain sort's large-array pass shape of
// 512 x 48 keys in two staging rounds, 8-bit digits, queue up to 65,536 keys.
static const SegVariant kSegVariants[] = {
    PA_SEG_VARIANT(512, 48, 8, 2, 8, 65536), PA_SEG_VARIANT(512, 48, 8, 2, 8, 32768),
    PA_SEG_VARIANT(512, 48, 8, 2, 8, 16384), PA_SEG_VARIANT(512, 16, 4, 1, 8, 65536),
    PA_SEG_VARIANT(768, 32, 8, 2, 8, 65536), PA_SEG_VARIANT(512, 48, 8, 2, 6, 65536),
    PA_SEG_VARIANT(512, 48, 8, 2, 4, 65536), PA_SEG_VARIANT(512, 16, 4, 1, 6, 65536)};
constexpr int kSegVariantCount = sizeof(kSegVariants) / sizeof(SegVariant);

static std::atomic<int> g_seg_tuned{-1};

int seg_shape_choice() {
  static const int forced = [] {
    const char* e = getenv("PA_SEG_VARIANT");
    const int v = e ? atoi(e) : -1;
    return (v >= 0 && v < kSegVariantCount) ? v : -1;
  }();
  if (forced >= 0) return forced;
  tune_cache_autoload();
  const int t = g_seg_tuned.load(std::memory_order_relaxed);
  return t >= 0 ? t : 0;
}

static const SegVariant& seg_variant() { return kSegVariants[seg_shape_choice()]; }

static uint64_t seg_queue_max(const SegVariant& v) {
  const char* e = getenv("PA_SEG_QUEUE_MAX");
  return e ? strtoull(e, nullptr, 10) : v.queue_max;
}

// Sorts the segments described by (address, length) lists; large segments get
// ping-pong space from `work` (their tmp addresses).  ctx.mu must be held.
struct SegPlan {
  std::vector<uint64_t> small_ptr[kLdsClasses], queue_ptr, queue_tmp, large_in, large_tmp, tile_info;
  std::vector<uint32_t> small_len[kLdsClasses], queue_len, large_len, large_tiles, large_row;
  // Segments of 2^30 keys or more (past a look-back word's 30-bit count):
  // sorted one by one as the plain sort sorts such arrays, in chunks merged
  // pairwise, after the other classes.
  std::vector<uint64_t> huge_ptr, huge_tmp, huge_len;
  // The shape in force when the plan is made (a captured plan keeps it).
  const SegVariant* var = &seg_variant();
  const uint32_t tile = var->tile();
  const uint64_t queue_max = seg_queue_max(*var);
  void add(uint64_t ptr, uint64_t len, uint64_t tmp_ptr) {
    if (len <= 1) return;
    if (len > kMaxSortN) {
      huge_ptr.push_back(ptr);
      huge_tmp.push_back(tmp_ptr);
      huge_len.push_back(len);
      return;
    }
    for (int c = 0; c < kLdsClasses; ++c) {
      if (len <= kLdsLimit[c]) {
        small_ptr[c].push_back(ptr);
        small_len[c].push_back((uint32_t)len);
        return;
      }
    }
    if (len <= queue_max) {
      queue_ptr.push_back(ptr);
      queue_tmp.push_back(tmp_ptr);
      queue_len.push_back((uint32_t)len);
      return;
    }
    large_in.push_back(ptr);
    large_tmp.push_back(tmp_ptr);
    large_len.push_back((uint32_t)len);
    large_tiles.push_back((uint32_t)((len + tile - 1) / tile));
  }
  // Look-back rows (contiguous per segment) and the interleaved claim order:
  // a counting sort of the tiles by bucket, stable in (segment, tile).
  void finalize() {
    // The queue is served longest first, so the last segments taken are short.
    std::vector<uint32_t> order(queue_len.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return queue_len[a] > queue_len[b]; });
    std::vector<uint64_t> qp(order.size()), qt(order.size());
    std::vector<uint32_t> ql(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
      qp[i] = queue_ptr[order[i]];
      qt[i] = queue_tmp[order[i]];
      ql[i] = queue_len[order[i]];
    }
    // A workgroup sorts a queued segment alone, so the queue only pays when
    // the longest segment is at most twice the average load of a workgroup;
    // the longest segments that break this go to the onesweep class instead.
    size_t first = qp.size();
    uint64_t tail = 0;
    for (size_t i = qp.size(); i-- > 0;) {
      tail += ql[i];
      if ((uint64_t)ql[i] * kQueueGrid <= 2 * tail) first = i;
    }
    for (size_t i = 0; i < first; ++i) {
      large_in.push_back(qp[i]);
      large_tmp.push_back(qt[i]);
      large_len.push_back(ql[i]);
      large_tiles.push_back((uint32_t)((ql[i] + tile - 1) / tile));
    }
    queue_ptr.assign(qp.begin() + first, qp.end());
    queue_tmp.assign(qt.begin() + first, qt.end());
    queue_len.assign(ql.begin() + first, ql.end());
    const size_t nl = large_tiles.size();
    large_row.resize(nl);
    uint64_t rows = 0;
    for (size_t g = 0; g < nl; ++g) {
      large_row[g] = (uint32_t)rows;
      rows += large_tiles[g];
    }
    tile_info.assign(rows, 0);
    // PA_SEG_GROUP overrides the group size in tiles, for sweeps (0: one
    // group holding every segment).
    const char* ge = getenv("PA_SEG_GROUP");
    uint64_t budget = ge ? strtoull(ge, nullptr, 10) : kSegGroupTiles;
    if (budget == 0) budget = ~0ull;
    size_t pos = 0;
    std::vector<uint32_t> count(kSegBuckets + 1);
    for (size_t g0 = 0, g1; g0 < nl; g0 = g1) {
      uint64_t tiles = 0;
      for (g1 = g0; g1 < nl && tiles < budget; ++g1) tiles += large_tiles[g1];
      std::fill(count.begin(), count.end(), 0u);
      for (size_t g = g0; g < g1; ++g) {
        const uint64_t T = large_tiles[g];
        for (uint64_t k = 0; k < T; ++k) ++count[(2 * k + 1) * kSegBuckets / (2 * T) + 1];
      }
      count[0] = (uint32_t)pos;
      for (uint32_t b = 0; b < kSegBuckets; ++b) count[b + 1] += count[b];
      for (size_t g = g0; g < g1; ++g) {
        const uint64_t T = large_tiles[g];
        for (uint64_t k = 0; k < T; ++k) tile_info[count[(2 * k + 1) * kSegBuckets / (2 * T)]++] = ((uint64_t)g << 32) | k;
      }
      pos = count[kSegBuckets - 1];
    }
  }
};

// Device image of a plan: the table block (pointer and length lists, rows,
// tile order) followed by per-segment counts, tile counters, the look-back
// table and `extra` bytes of ping-pong space for large segments.
struct PlanLayout {
  size_t t_sp[kLdsClasses], t_sl[kLdsClasses], t_qp, t_qt, t_ql, t_li, t_lt, t_ll, t_lr, t_ti, table_bytes;
  size_t o_tab, o_cnt, o_ctr, o_lb, o_huge, o_work, total;
};

static PlanLayout plan_layout(const SegPlan& P, size_t extra) {
  const size_t nl = P.large_in.size(), nt = P.tile_info.size();
  PlanLayout Y;
  Layout T;
  for (int c = 0; c < kLdsClasses; ++c) {
    Y.t_sp[c] = T.take(P.small_ptr[c].size() * 8);
    Y.t_sl[c] = T.take(P.small_len[c].size() * 4);
  }
  const size_t nq = P.queue_len.size();
  Y.t_qp = T.take(nq * 8);
  Y.t_qt = T.take(nq * 8);
  Y.t_ql = T.take(nq * 4);
  Y.t_li = T.take(nl * 8);
  Y.t_lt = T.take(nl * 8);
  Y.t_ll = T.take(nl * 4);
  Y.t_lr = T.take(nl * 4);
  Y.t_ti = T.take(nt * 8);
  Y.table_bytes = T.off;
  Layout L;
  Y.o_tab = L.take(Y.table_bytes);
  const size_t DR = (size_t)P.var->passes() * P.var->bins();  // counts per segment
  Y.o_cnt = L.take(nl * DR * 4);
  Y.o_ctr = L.take(64);
  Y.o_lb = L.take(DR * nt * 4);
  size_t huge = 0;
  for (uint64_t len : P.huge_len) huge = std::max(huge, large_sort_u32_scratch_bytes(len));
  Y.o_huge = L.take(huge);
  Y.o_work = L.take(extra);
  Y.total = L.off;
  return Y;
}

// Queue and large segments without a caller tmp get consecutive ping-pong
// space.
static void assign_work(SegPlan& P, char* work) {
  uint64_t off = 0;
  for (size_t i = 0; i < P.queue_tmp.size(); ++i) {
    P.queue_tmp[i] = (uint64_t)(uintptr_t)work + off * 4;
    off += P.queue_len[i];
  }
  for (size_t i = 0; i < P.large_tmp.size(); ++i) {
    P.large_tmp[i] = (uint64_t)(uintptr_t)work + off * 4;
    off += P.large_len[i];
  }
  for (size_t i = 0; i < P.huge_tmp.size(); ++i) {
    P.huge_tmp[i] = (uint64_t)(uintptr_t)work + off * 4;
    off += P.huge_len[i];
  }
}

static void write_tables(const SegPlan& P, const PlanLayout& Y, char* pin) {
  for (int c = 0; c < kLdsClasses; ++c) {
    std::copy(P.small_ptr[c].begin(), P.small_ptr[c].end(), (uint64_t*)(pin + Y.t_sp[c]));
    std::copy(P.small_len[c].begin(), P.small_len[c].end(), (uint32_t*)(pin + Y.t_sl[c]));
  }
  std::copy(P.queue_ptr.begin(), P.queue_ptr.end(), (uint64_t*)(pin + Y.t_qp));
  std::copy(P.queue_tmp.begin(), P.queue_tmp.end(), (uint64_t*)(pin + Y.t_qt));
  std::copy(P.queue_len.begin(), P.queue_len.end(), (uint32_t*)(pin + Y.t_ql));
  std::copy(P.large_in.begin(), P.large_in.end(), (uint64_t*)(pin + Y.t_li));
  std::copy(P.large_tmp.begin(), P.large_tmp.end(), (uint64_t*)(pin + Y.t_lt));
  std::copy(P.large_len.begin(), P.large_len.end(), (uint32_t*)(pin + Y.t_ll));
  std::copy(P.large_row.begin(), P.large_row.end(), (uint32_t*)(pin + Y.t_lr));
  std::copy(P.tile_info.begin(), P.tile_info.end(), (uint64_t*)(pin + Y.t_ti));
}

// The kernels of one plan, on device memory laid out by plan_layout.  Only
// kernels (no memset, no copy), so the sequence can be captured in a graph.
static int launch_plan(const SegPlan& P, const PlanLayout& Y, char* dev, hipStream_t s) {
  const size_t nl = P.large_in.size(), nt = P.tile_info.size();
  const char* tab = dev + Y.o_tab;
  const uint64_t* d_li = (const uint64_t*)(tab + Y.t_li);
  const uint64_t* d_lt = (const uint64_t*)(tab + Y.t_lt);
  const uint32_t* d_ll = (const uint32_t*)(tab + Y.t_ll);
  const uint32_t* d_lr = (const uint32_t*)(tab + Y.t_lr);
  const uint64_t* d_ti = (const uint64_t*)(tab + Y.t_ti);
  uint32_t* counts = (uint32_t*)(dev + Y.o_cnt);
  uint32_t* ctr = (uint32_t*)(dev + Y.o_ctr);
  uint32_t* lbp = (uint32_t*)(dev + Y.o_lb);
  const SegVariant& V = *P.var;
  const uint32_t D = (uint32_t)V.passes(), R = V.bins();
  const size_t lb_bytes = (size_t)D * R * nt * 4;
  for (int c = 0; c < kLdsClasses; ++c) {
    const size_t ns = P.small_ptr[c].size();
    if (!ns) continue;
    const uint64_t* d_sp = (const uint64_t*)(tab + Y.t_sp[c]);
    const uint32_t* d_sl = (const uint32_t*)(tab + Y.t_sl[c]);
    ProfScope ps("k_seg_small", s);
    switch (c) {
      case 0: hipLaunchKernelGGL(k_seg_small<256>, dim3((unsigned)ns), dim3(256), 0, s, d_sp, d_sl); break;
      case 1: hipLaunchKernelGGL(k_seg_small<512>, dim3((unsigned)ns), dim3(512), 0, s, d_sp, d_sl); break;
      default: hipLaunchKernelGGL(k_seg_small<1024>, dim3((unsigned)ns), dim3(1024), 0, s, d_sp, d_sl); break;
    }
  }
  const size_t nq = P.queue_len.size();
  if (nl || nq)
    hipLaunchKernelGGL(k_seg_prepare,
                       dim3((unsigned)std::max<size_t>(1, std::min<size_t>((nl * D * R + 255) / 256, 4096))), dim3(256),
                       0, s, counts, nl * D * R, ctr);
  if (nq) {
    ProfScope ps("k_seg_queue", s);
    hipLaunchKernelGGL(k_seg_queue, dim3((unsigned)std::min<size_t>(nq, kQueueGrid)), dim3(kQueueBlock), 0, s,
                       (const uint64_t*)(tab + Y.t_qp), (const uint64_t*)(tab + Y.t_qt), (const uint32_t*)(tab + Y.t_ql),
                       (uint32_t)nq, ctr + 8);
  }
  if (nl) {
    PA_TRY(V.hist(s, d_ti, d_li, d_ll, (uint32_t)nt, V.tile(), counts, (uint4*)lbp, lb_bytes / 16));
    for (uint32_t p = 0; p < D; ++p) {
      SegPassArgs a;
      a.counts = counts + p * R;
      a.lb = lbp + (size_t)p * nt * R;
      a.ctr = ctr + p;
      a.ntiles = (uint32_t)nt;
      a.shift = (uint32_t)V.rb * p;
      a.gstride = D;
      a.tile_info = d_ti;
      a.tin = (p & 1) ? d_lt : d_li;
      a.tout = (p & 1) ? d_li : d_lt;
      a.len = d_ll;
      a.row = d_lr;
      PA_TRY(V.pass(s, a));
    }
  }
  for (size_t h = 0; h < P.huge_len.size(); ++h)
    PA_TRY(large_sort_u32_enqueue(reinterpret_cast<uint32_t*>(P.huge_ptr[h]), reinterpret_cast<uint32_t*>(P.huge_tmp[h]),
                                  P.huge_len[h], dev + Y.o_huge, s));
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// One-shot run on the stream's scratch: tables go up through pinned memory.
// `extra` bytes of ping-pong space are reserved when no caller tmp exists.
// ctx.mu must be held.
// Tables up through pinned memory, then the plan's kernels, on device memory
// b laid out by Y.  ctx.mu must be held.
static int run_plan_at(const SegPlan& P, const PlanLayout& Y, char* b, hipStream_t s, StreamCtx& ctx) {
  void* pinned;
  int slot;
  PA_TRY(acquire_pinned(ctx, Y.table_bytes, &pinned, &slot));
  write_tables(P, Y, (char*)pinned);
  PA_HIP(hipMemcpyAsync(b + Y.o_tab, pinned, Y.table_bytes, hipMemcpyHostToDevice, s));
  PA_TRY(release_pinned(ctx, slot));
  return launch_plan(P, Y, b, s);
}

static int run_plan_alloc(SegPlan& P, size_t extra, bool fill_tmp, hipStream_t s, StreamCtx& ctx) {
  if (P.tile_info.size() >= (1ull << 32)) return arg_error("segmented sort: too many tiles");
  const PlanLayout Y = plan_layout(P, extra);
  void* base;
  PA_TRY(ensure_scratch(ctx, Y.total, &base));
  char* b = (char*)base;
  if (fill_tmp) assign_work(P, b + Y.o_work);
  return run_plan_at(P, Y, b, s, ctx);
}

// The leaves of a merge sort as one segmented batch: keys[i * leaf, (i + 1) *
// leaf) for every i (the last one shorter), each with the same range of tmp
// as its ping-pong space.
static void leaves_plan(uintptr_t keys, uintptr_t tmp, size_t n, size_t leaf, SegPlan* P) {
  for (size_t o = 0; o < n; o += leaf) P->add(keys + o * 4, std::min(leaf, n - o), tmp + o * 4);
  P->finalize();
}

size_t leaves_scratch_bytes_u32(size_t n, size_t leaf) {
  SegPlan P;
  leaves_plan(0, 0, n, leaf, &P);
  return plan_layout(P, 0).total;
}

int leaves_enqueue_u32(uint32_t* keys, uint32_t* tmp, size_t n, size_t leaf, char* scratch, hipStream_t s,
                       StreamCtx& ctx) {
  SegPlan P;
  leaves_plan((uintptr_t)keys, (uintptr_t)tmp, n, leaf, &P);
  if (P.tile_info.size() >= (1ull << 32)) return arg_error("merge sort: too many tiles");
  return run_plan_at(P, plan_layout(P, 0), scratch, s, ctx);
}

// Builds the plan of a pointer-table call; returns PA_OK or an error status.
static int build_ptr_plan(uint32_t* const* keys, const uint64_t* lens, size_t num_segments, const char* who,
                          SegPlan* P, uint64_t* large_total) {
  *large_total = 0;
  for (size_t i = 0; i < num_segments; ++i) {
    if (lens[i] == 0) continue;
    if (!keys[i]) return arg_error((std::string(who) + ": null segment pointer").c_str());
    P->add((uint64_t)(uintptr_t)keys[i], lens[i], 0);
    if (lens[i] > kLdsMax) *large_total += lens[i];
  }
  for (size_t i = 0; i < num_segments; ++i)
    if (lens[i] > 1 && !is_device_ptr(keys[i]))
      return arg_error((std::string(who) + ": segments must be device memory").c_str());
  P->finalize();
  return PA_OK;
}

// Plain sort of one device array of n <= kLdsMax u32 keys in one launch: the
// smallest workgroup whose LDS tile holds the array.
int small_sort_enqueue_u32(uint32_t* keys, size_t n, hipStream_t s) {
  if (n <= 1) return PA_OK;
  ProfScope ps("k_sort_small", s);
  if (n <= kLdsLimit[0])
    hipLaunchKernelGGL(k_sort_small<256>, dim3(1), dim3(256), 0, s, keys, (uint32_t)n);
  else if (n <= kLdsLimit[1])
    hipLaunchKernelGGL(k_sort_small<512>, dim3(1), dim3(512), 0, s, keys, (uint32_t)n);
  else
    hipLaunchKernelGGL(k_sort_small<1024>, dim3(1), dim3(1024), 0, s, keys, (uint32_t)n);
  PA_HIP(hipGetLastError());
  return PA_OK;
}
size_t small_sort_max_keys() { return kLdsMax; }

static std::string seg_variant_name(int i) {
  const SegVariant& v = kSegVariants[i];
  char b[96];
  std::snprintf(b, sizeof(b), "%dx%d/lb%d/sp%d/ul16/rb%d/q%llu", v.block, v.items, v.lbw, v.split, v.rb,
                (unsigned long long)v.queue_max);
  return b;
}

const TuneTarget* seg_tune_target() {
  static const TuneTarget t = [] {
    TuneTarget x;
    x.key = "seg_u32";
    x.classes = 1;
    x.count = [] { return kSegVariantCount; };
    x.name = seg_variant_name;
    x.get = [](int) { return g_seg_tuned.load(); };
    x.set = [](int, int i) { g_seg_tuned.store(i); };
    return x;
  }();
  return &t;
}

void seg_tune_reset() { g_seg_tuned.store(-1); }

int seg_shape_info(int index, pa_sort_shape* out) {
  if (index < 0 || index >= kSegVariantCount) return arg_error("pa_sort_shape_info: index");
  const SegVariant& v = kSegVariants[index];
  out->block = v.block;
  out->items = v.items;
  out->lookback = v.lbw;
  out->split = v.split;
  out->lanes_per_unit = 16;
  out->digit_bits = v.rb;
  out->passes = v.passes();
  out->queue_max_keys = v.queue_max;
  return PA_OK;
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_segmented_sort_u32(uint32_t* keys, uint32_t* tmp, const uint64_t* offsets, size_t num_segments,
                          pa_stream_t stream) {
  if (num_segments == 0) return PA_OK;
  if (!offsets) return arg_error("pa_segmented_sort_u32: null offsets");
  for (size_t i = 0; i < num_segments; ++i)
    if (offsets[i + 1] < offsets[i]) return arg_error("pa_segmented_sort_u32: offsets must be non-decreasing");
  if (offsets[num_segments] == offsets[0]) return PA_OK;
  if (!keys) return arg_error("pa_segmented_sort_u32: null keys");
  hipStream_t s = (hipStream_t)stream;
  if (!is_device_ptr(keys) || (tmp && !is_device_ptr(tmp)))
    return arg_error("pa_segmented_sort_u32: keys and tmp must be device memory");
  SegPlan P;
  uint64_t large_total = 0;
  for (size_t i = 0; i < num_segments; ++i) {
    const uint64_t len = offsets[i + 1] - offsets[i];
    P.add((uint64_t)(uintptr_t)(keys + offsets[i]), len, tmp ? (uint64_t)(uintptr_t)(tmp + offsets[i]) : 0);
    if (len > kLdsMax) large_total += len;
  }
  P.finalize();
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  return run_plan_alloc(P, tmp ? 0 : large_total * 4, tmp == nullptr, s, ctx);
}

}  // extern "C"

namespace pa {
// Scratch of pa_segmented_sort_ptrs_u32 for these segment lengths: the same
// plan and layout, from the lengths alone (no pointer checks).
int segmented_lens_scratch_bytes(const uint64_t* lens, size_t num, size_t* bytes) {
  *bytes = 0;
  if (num == 0) return PA_OK;
  if (!lens) return arg_error("segmented sort: null length list");
  SegPlan P;
  uint64_t large_total = 0;
  for (size_t i = 0; i < num; ++i) {
    P.add(1, lens[i], 0);  // addresses do not enter the layout
    if (lens[i] > kLdsMax) large_total += lens[i];
  }
  P.finalize();
  *bytes = plan_layout(P, large_total * 4).total;
  return PA_OK;
}
}  // namespace pa

extern "C" {

int pa_segmented_sort_ptrs_u32(uint32_t* const* keys, const uint64_t* lens, size_t num_segments, pa_stream_t stream) {
  if (num_segments == 0) return PA_OK;
  if (!keys || !lens) return arg_error("pa_segmented_sort_ptrs_u32: null table");
  hipStream_t s = (hipStream_t)stream;
  SegPlan P;
  uint64_t large_total = 0;
  PA_TRY(build_ptr_plan(keys, lens, num_segments, "pa_segmented_sort_ptrs_u32", &P, &large_total));
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  return run_plan_alloc(P, large_total * 4, true, s, ctx);
}

int pa_autotune_segmented_u32(const uint64_t* lens, size_t num_segments, pa_stream_t stream) {
  if (!lens || num_segments == 0) return arg_error("pa_autotune_segmented_u32: empty length list");
  uint64_t total = 0;
  for (size_t i = 0; i < num_segments; ++i) {
    total += lens[i];
  }
  if (total == 0) return arg_error("pa_autotune_segmented_u32: no keys");
  hipStream_t s = (hipStream_t)stream;
  uint32_t* keys = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = PA_OK;
  std::vector<uint32_t*> ptrs(num_segments);
  do {
    hipError_t e;
    if ((e = hipMalloc(&keys, total * 4)) != hipSuccess) {
      st = hip_fail(e, "hipMalloc", __FILE__, __LINE__);
      break;
    }
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
      st = hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
      break;
    }
    uint64_t off = 0;
    for (size_t i = 0; i < num_segments; ++i) {
      ptrs[i] = keys + off;
      off += lens[i];
    }
    float best_ms = 0.f;
    int best = -1;
    const int before = g_seg_tuned.load();
    for (int v = 0; v < kSegVariantCount && st == PA_OK; ++v) {
      g_seg_tuned.store(v);
      float sum = 0.f;
      for (int r = 0; r < 3 && st == PA_OK; ++r) {  // one warm-up, two timed
        if ((st = pa_fill_random_u32(keys, total, 0x5e9 + r, stream)) != PA_OK) break;
        (void)hipEventRecord(e0, s);
        if ((st = pa_segmented_sort_ptrs_u32(ptrs.data(), lens, num_segments, stream)) != PA_OK) break;
        (void)hipEventRecord(e1, s);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) {
          st = hip_fail(e, "hipEventSynchronize", __FILE__, __LINE__);
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) sum += ms;
      }
      if (st == PA_OK && (best < 0 || sum < best_ms)) {
        best = v;
        best_ms = sum;
      }
    }
    g_seg_tuned.store(st == PA_OK ? best : before);
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(keys);
  if (st == PA_OK) tune_cache_autosave();
  return st;
}

struct pa_segplan {
  pa::SegPlan plan;
  pa::PlanLayout layout;
  char* mem = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void pa_segmented_plan_destroy(pa_segplan* plan) {
  if (!plan) return;
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  if (plan->mem) (void)hipFree(plan->mem);
  (void)hipGetLastError();
  delete plan;
}

pa_segplan* pa_segmented_plan_create_u32(uint32_t* const* keys, const uint64_t* lens, size_t num_segments) {
  if (!keys || !lens || num_segments == 0) {
    arg_error("pa_segmented_plan_create_u32: empty or null table");
    return nullptr;
  }
  pa_segplan* G = new pa_segplan();
  uint64_t large_total = 0;
  hipStream_t cap = nullptr;
  auto fail = [&](hipError_t e, const char* what) -> pa_segplan* {
    if (e != hipSuccess) hip_fail(e, what, __FILE__, __LINE__);
    if (cap) (void)hipStreamDestroy(cap);
    pa_segmented_plan_destroy(G);
    return nullptr;
  };
  if (build_ptr_plan(keys, lens, num_segments, "pa_segmented_plan_create_u32", &G->plan, &large_total) != PA_OK)
    return fail(hipSuccess, "");
  if (G->plan.tile_info.size() >= (1ull << 32)) {
    arg_error("pa_segmented_plan_create_u32: too many tiles");
    return fail(hipSuccess, "");
  }
  G->layout = plan_layout(G->plan, large_total * 4);
  hipError_t e = hipMalloc(&G->mem, std::max<size_t>(G->layout.total, 256));
  if (e != hipSuccess) return fail(e, "hipMalloc");
  assign_work(G->plan, G->mem + G->layout.o_work);
  std::vector<char> image(G->layout.table_bytes);
  write_tables(G->plan, G->layout, image.data());
  if (!image.empty() &&
      (e = hipMemcpy(G->mem + G->layout.o_tab, image.data(), image.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy");
  if ((e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking)) != hipSuccess) return fail(e, "hipStreamCreate");
  if ((e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return fail(e, "hipStreamBeginCapture");
  const int st = launch_plan(G->plan, G->layout, G->mem, cap);
  e = hipStreamEndCapture(cap, &G->graph);
  if (st != PA_OK) return fail(hipSuccess, "");
  if (e != hipSuccess) return fail(e, "hipStreamEndCapture");
  if ((e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0)) != hipSuccess)
    return fail(e, "hipGraphInstantiate");
  (void)hipStreamDestroy(cap);
  return G;
}

int pa_segmented_plan_run(pa_segplan* plan, pa_stream_t stream) {
  if (!plan || !plan->exec) return arg_error("pa_segmented_plan_run: null plan");
  PA_HIP(hipGraphLaunch(plan->exec, (hipStream_t)stream));
  return PA_OK;
}

}  // extern "C"
