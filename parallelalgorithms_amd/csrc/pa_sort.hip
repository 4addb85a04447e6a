// pa_sort.hip -- LSD radix sort for gfx950: one fused read that builds every
// digit histogram, then one onesweep kernel per 8-bit digit.
//
// Reference structure it replaces (DragonSpit/ParallelAlgorithms):
//   SortRadixPar / SortRadixInnerPar   RadixSortLsdParallel.h:215-387
//     per digit: per-quantum histogram (ComputeStartOfBinsPar :108-165), a
//     serial (quantum x bin) prefix, then a stable scatter per quantum with
//     64-deep per-bin write buffers (PermuteDerandomizedNew :169-212).
//   TwoPhase ("1 + D" passes)          RadixSortLsdParallel.h:43-106
//     one histogram read for all digits (HistogramByteComponentsParallel).
//
// GPU form:
//   k_sort_prepare   zero the digit counts and tile counters.
//   k_hist_tiles     one read of the keys (the TwoPhase idea): every digit's
//                    histogram, pass 0's per-tile digit counts, and the
//                    clearing of the later passes' look-back rows.
//   k_tile_sums,     pass 0's per-tile counts -> inclusive prefixes (as if
//   k_tile_scan      every pass-0 tile had resolved its look-back) and the
//                    digit bases of every pass.
//   k_onesweep       per digit: a tile of BLOCK x ITEMS keys is ranked stably
//                    in LDS, each tile's per-bin count is published to a
//                    decoupled look-back table (replacing the serial
//                    quantum x bin prefix), and keys are staged in LDS in
//                    digit order so the global writes are runs per bin (the
//                    GPU form of the 64-deep de-randomisation buffers).
//
// HBM layout: keys (n x K) ping-pong between `keys` and `tmp`; D passes with
// D = sizeof(K) is even, so the result lands back in `keys`.  Scratch per
// call: counts D x 256 u32, tile counters (8 for pass 0, one per later pass),
// look-back D x tiles x 256 u32.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_ext.h>

#include "pa_sort_kernels.h"

namespace pa {

__global__ __launch_bounds__(256) void k_sort_prepare(uint32_t* __restrict__ counts, uint32_t ncounts,
                                                      uint32_t* __restrict__ counters, uint32_t ncounters) {
  for (uint32_t i = threadIdx.x; i < ncounts; i += blockDim.x) counts[i] = 0;
  for (uint32_t i = threadIdx.x; i < ncounters; i += blockDim.x) counters[i] = 0;
}

// Pass-0 prefixes.  rows[t][d] holds digit d's count in pass-0 tile t (R =
// 2^RB digits); the two kernels below turn every row into the inclusive
// status word of its tile (count of digit d in tiles 0..t), as if every tile
// had already resolved its look-back.  Group g of blocks owns tiles
// [g * per, (g + 1) * per).  Blocks have 1024 threads: thread (q, d) =
// (tid / R, tid % R) handles digit d for one of Q = 1024 / R slices of the
// group's tiles, so each thread keeps several independent loads in flight
// instead of walking the rows one after another.
constexpr unsigned kPrefixThreads = 1024;

// Step 1: partial[g][d] = count of digit d over group g's tiles (blocks
// g < groups).  The passes' blocks after them write the digit bases of pass
// p = blockIdx.x - groups: base[p][d] = sum_{d' < d} counts[p][d'], the first
// output slot of digit d; and trivial[p] = 1 when one digit holds all n keys
// (the pass would leave the order as it is).
template <int RB>
__global__ __launch_bounds__(kPrefixThreads) void k_tile_sums(const uint32_t* __restrict__ rows, uint32_t ntiles,
                                                              uint32_t per, uint32_t* __restrict__ partial,
                                                              uint32_t groups, const uint32_t* __restrict__ counts,
                                                              uint32_t* __restrict__ base, uint32_t n,
                                                              uint32_t* __restrict__ trivial) {
  constexpr uint32_t R = 1u << RB, Q = kPrefixThreads / R;
  __shared__ uint32_t s_q[Q][R];
  __shared__ uint32_t s_full;
  const uint32_t d = threadIdx.x % R, q = threadIdx.x / R;
  if (blockIdx.x >= groups) {
    // Threads 0..R-1 scan the pass's R counts.
    const uint32_t p = blockIdx.x - groups, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t c = 0, inc = 0;
    if (threadIdx.x == 0) s_full = 0;
    if (q == 0) {
      c = counts[p * R + d];
      inc = dev::wave_incl_scan(c, lane);
      if (lane == 63) s_q[0][wave] = inc;
    }
    __syncthreads();
    if (q == 0) {
      uint32_t pre = 0;
      for (uint32_t w = 0; w < wave; ++w) pre += s_q[0][w];
      base[p * R + d] = pre + inc - c;
      if (c == n) s_full = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) trivial[p] = s_full;
    return;
  }
  const uint32_t t0 = blockIdx.x * per, t1 = min(ntiles, t0 + per);
  uint32_t acc = 0;
#pragma unroll 8
  for (uint32_t t = t0 + q; t < t1; t += Q) acc += rows[(size_t)t * R + d];
  s_q[q][d] = acc;
  __syncthreads();
  if (q == 0) {
    uint32_t sum = 0;
    for (uint32_t w = 0; w < Q; ++w) sum += s_q[w][d];
    partial[blockIdx.x * R + d] = sum;
  }
}

// Step 2: group g adds the partials of groups < g, then scans its own tiles
// in chunks of Q x 8: thread (q, d) loads 8 consecutive rows at once.  When
// roles is set, block 0 also writes the passes' roles from the trivial flags
// into the top bits of their tile counters (kRoleSkip, kRoleSwap, see
// pa_sort_kernels.h; pass 0 has kXcds counters, pass p > 0 counter
// kXcds - 1 + p).  Trivial passes are skipped, except that when the number
// of the others and the pass count D differ in parity the first trivial pass
// still runs, so the data ends in the buffer the host expects (the keys for
// an even D, the temporary array for an odd one, which the host copies
// back).
template <int RB>
__global__ __launch_bounds__(kPrefixThreads) void k_tile_scan(uint32_t* __restrict__ rows, uint32_t ntiles,
                                                              uint32_t per, const uint32_t* __restrict__ partial,
                                                              const uint32_t* __restrict__ trivial, uint32_t passes,
                                                              bool roles, uint32_t* __restrict__ counters) {
  constexpr uint32_t R = 1u << RB, Q = kPrefixThreads / R;
  __shared__ uint32_t s_q[Q][R];
  const uint32_t d = threadIdx.x % R, q = threadIdx.x / R;
  if (roles && blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t tcount = 0, keep = passes;  // keep: a trivial pass that runs anyway
    for (uint32_t p = 0; p < passes; ++p)
      if (!trivial[p])
        ++tcount;
      else if (keep == passes)
        keep = p;
    if ((tcount & 1u) == (passes & 1u)) keep = passes;  // parities agree: every trivial pass is skipped
    uint32_t cur = 0;                                   // buffer holding the data: 0 keys, 1 temporary
    for (uint32_t p = 0; p < passes; ++p) {
      uint32_t r;
      if (trivial[p] && p != keep) {
        r = kRoleSkip;
      } else {
        r = (p & 1u) != cur ? kRoleSwap : 0u;
        cur ^= 1u;
      }
      if (p == 0)
        for (uint32_t x = 0; x < kXcds; ++x) counters[x] = r;
      else
        counters[kXcds - 1 + p] = r;
    }
  }
  uint32_t acc = 0;
#pragma unroll 8
  for (uint32_t g = q; g < blockIdx.x; g += Q) acc += partial[g * R + d];
  s_q[q][d] = acc;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < Q; ++w) base += s_q[w][d];
  const uint32_t t0 = blockIdx.x * per, t1 = min(ntiles, t0 + per);
  for (uint32_t c0 = t0; c0 < t1; c0 += 8 * Q) {
    const uint32_t tq = c0 + 8 * q;
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = tq + j < t1 ? rows[(size_t)(tq + j) * R + d] : 0u;
#pragma unroll
    for (int j = 1; j < 8; ++j) v[j] += v[j - 1];
    __syncthreads();  // every thread has read s_q of the previous step
    s_q[q][d] = v[7];
    __syncthreads();
    uint32_t off = base, all = 0;
    for (uint32_t w = 0; w < Q; ++w) {
      const uint32_t x = s_q[w][d];
      if (w < q) off += x;
      all += x;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (tq + j < t1) rows[(size_t)(tq + j) * R + d] = kLbInc | (off + v[j]);
    base += all;
  }
}

// ---------------------------------------------------------------------------
// Pass variants (tile shape, look-back width, persistent blocks), selectable
// at run time; PA_SORT_VARIANT overrides the default (entry 0).
// ---------------------------------------------------------------------------
struct PassArgs {
  const void* kin;
  void* kout;
  const uint32_t* vin;
  uint32_t* vout;
  uint32_t n;
  const uint32_t* gtab;
  uint32_t* lb;
  uint32_t* ctr;
  uint32_t ntiles;
  uint32_t shift;
  uint32_t fmode;
};

struct PassVariant {
  int block, items, lbw, late, minw, split, ul, rb;
  int (*launch)(hipStream_t s, const PassArgs& a);
};

// Digit passes of a sort of K keys with RB-bit digits.
template <typename K>
constexpr int sort_passes(int rb) {
  return (8 * (int)sizeof(K) + rb - 1) / rb;
}

template <typename K, bool KV, int BLOCK, int ITEMS, int LBW, bool LATE, int MINW, int MODE, bool UNI, int SPLIT,
          int UL, int RB>
int launch_pass(hipStream_t s, const PassArgs& a) {
  // Timed launches carry their events in the dispatch (no extra packets).
  ProfScope ps("k_onesweep", s, true);
  auto kern = k_onesweep<K, KV, BLOCK, ITEMS, LBW, MODE, LATE, MINW, UNI, SPLIT, UL, RB>;
  if (ps.start) {
    hipExtLaunchKernelGGL(kern, dim3(a.ntiles), dim3(BLOCK), 0, s, ps.start, ps.stop, 0, static_cast<const K*>(a.kin),
                          static_cast<K*>(a.kout), a.vin, a.vout, a.n, a.gtab, a.lb, a.ctr, a.ntiles, a.shift, a.fmode,
                          (const uint64_t*)nullptr, (const uint64_t*)nullptr, (const uint64_t*)nullptr,
                          (const uint32_t*)nullptr, 0u, (const uint32_t*)nullptr);
    ps.attached = true;
  } else {
    hipLaunchKernelGGL(kern, dim3(a.ntiles), dim3(BLOCK), 0, s, static_cast<const K*>(a.kin), static_cast<K*>(a.kout),
                       a.vin, a.vout, a.n, a.gtab, a.lb, a.ctr, a.ntiles, a.shift, a.fmode, nullptr, nullptr, nullptr,
                       nullptr, 0u, nullptr);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// PA_VARIANT(key, payload, block, items, look-back width, late look-back, min waves/SIMD,
//            uniform-digit fast path, staging rounds, lanes per ranking unit), 8-bit digits;
// PA_VARIANT_RB(..., digit bits) for other widths.
#define PA_VARIANT_RB(K, KV, B, I, W, P, M, U, S, L, RB) \
  { B, I, W, P, M, S, L, RB, &launch_pass<K, KV, B, I, W, P, M, kModePlain, U, S, L, RB> }
#define PA_VARIANT(K, KV, B, I, W, P, M, U, S, L) PA_VARIANT_RB(K, KV, B, I, W, P, M, U, S, L, 8)

// Size classes of a sort call: n <= 2^16, <= 2^20, <= 2^24, larger.
constexpr int kClasses = 4;
inline int size_class(size_t n) { return n <= (1u << 16) ? 0 : n <= (1u << 20) ? 1 : n <= (1u << 24) ? 2 : 3; }

// Candidate passes per key type; by_class is the built-in choice per size
// class (small arrays want small tiles for parallelism, large arrays the
// split 24K-key tile).
template <typename K, bool KV>
struct Variants;
template <>
struct Variants<uint32_t, false> {
  static constexpr PassVariant table[] = {
      PA_VARIANT(uint32_t, false, 512, 48, 8, true, 4, true, 2, 32),  PA_VARIANT(uint32_t, false, 512, 24, 4, true, 6, true, 1, 32),
      PA_VARIANT(uint32_t, false, 512, 16, 4, true, 8, true, 1, 32),  PA_VARIANT(uint32_t, false, 512, 48, 16, true, 4, true, 2, 32),
      PA_VARIANT(uint32_t, false, 256, 48, 8, true, 4, true, 2, 32),  PA_VARIANT(uint32_t, false, 512, 48, 8, true, 4, true, 2, 16),
      PA_VARIANT(uint32_t, false, 512, 16, 4, true, 8, true, 1, 16),  PA_VARIANT(uint32_t, false, 512, 24, 4, true, 6, true, 1, 16),
      PA_VARIANT(uint32_t, false, 768, 32, 8, true, 6, true, 2, 16),
      // 4- to 7-bit digits (8, 7, 6, 5 passes): fewer bins per tile, more passes.
      PA_VARIANT_RB(uint32_t, false, 512, 16, 4, true, 8, true, 1, 16, 7),
      PA_VARIANT_RB(uint32_t, false, 512, 16, 4, true, 8, true, 1, 16, 6),
      PA_VARIANT_RB(uint32_t, false, 512, 16, 4, true, 8, true, 1, 16, 5),
      PA_VARIANT_RB(uint32_t, false, 512, 16, 4, true, 8, true, 1, 16, 4),
      PA_VARIANT_RB(uint32_t, false, 768, 32, 8, true, 6, true, 2, 16, 7),
      PA_VARIANT_RB(uint32_t, false, 768, 32, 8, true, 6, true, 2, 16, 6),
      PA_VARIANT_RB(uint32_t, false, 768, 32, 8, true, 6, true, 2, 16, 5),
      PA_VARIANT_RB(uint32_t, false, 768, 32, 8, true, 6, true, 2, 16, 4)};
  // 512 x 16 up to 2^20 keys; 512 x 24 with 16-lane units up to 2^24 (a sort back to back: 1.2M 74 us,
  // 2M 88, 4M 112.5, 8M 157, 16M 231.5 against 75.5, 87, 120, 174, 255 for 512 x 16; 768 x 32 is faster
  // from 8M on but slower below 2M: 87 us at 1.2M); above, the 24,576-key split tile as 768 x 32 (12
  // waves per workgroup, two workgroups per CU: 100M passes 188-227 us against 194-233 us for 512 x 48).
  static constexpr int by_class[kClasses] = {2, 2, 7, 8};
};
template <>
struct Variants<uint32_t, true> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint32_t, true, 512, 16, 4, true, 4, true, 1, 32),
                                          PA_VARIANT(uint32_t, true, 512, 20, 8, true, 4, true, 1, 32)};
  // 512 x 20 above 2^24 keys (100M: 2.157 -> 2.117 ms; at 16M it is 6% slower than 512 x 16).
  static constexpr int by_class[kClasses] = {0, 0, 0, 1};
};
template <>
struct Variants<uint64_t, false> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint64_t, false, 512, 8, 4, true, 8, true, 1, 32),
                                          PA_VARIANT(uint64_t, false, 512, 12, 4, true, 6, true, 1, 32),
                                          PA_VARIANT(uint64_t, false, 512, 16, 8, true, 4, true, 1, 16),
                                          PA_VARIANT(uint64_t, false, 512, 24, 8, true, 4, true, 2, 16),
                                          PA_VARIANT_RB(uint64_t, false, 512, 16, 8, true, 4, true, 1, 16, 7),
                                          PA_VARIANT_RB(uint64_t, false, 512, 16, 8, true, 4, true, 1, 16, 6),
                                          PA_VARIANT_RB(uint64_t, false, 512, 16, 8, true, 4, true, 1, 16, 5),
                                          PA_VARIANT_RB(uint64_t, false, 512, 16, 8, true, 4, true, 1, 16, 4)};
  // 512 x 16 with 16-lane units up to 2^20 keys (1M: 0.155 -> 0.149 ms), the 512 x 24 split with
  // 16-lane units above (4M: 0.331 -> 0.254 ms, 16M: 0.798 -> 0.632 ms, 100M: 4.77 -> 3.57 ms).
  static constexpr int by_class[kClasses] = {0, 2, 3, 3};
};
template <>
struct Variants<uint64_t, true> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint64_t, true, 512, 8, 4, true, 6, true, 1, 32),
                                          PA_VARIANT(uint64_t, true, 512, 16, 8, true, 4, true, 1, 16)};
  // 512 x 16 with 16-lane units above 2^20 keys (4M: 0.358 -> 0.284 ms, 16M: 1.10 -> 0.96 ms,
  // 100M: 7.01 -> 5.72 ms).
  static constexpr int by_class[kClasses] = {0, 0, 1, 1};
};

// Choices set by the autotuner, a tuning cache or pa_sort_shape_set (-1:
// none) for keys-only sorts: [0] u32 keys (also float and signed 32-bit),
// [1] u64 keys.
static std::atomic<int> g_tuned[2][kClasses] = {{{-1}, {-1}, {-1}, {-1}}, {{-1}, {-1}, {-1}, {-1}}};

template <typename K, bool KV>
constexpr size_t variant_count() {
  return sizeof(Variants<K, KV>::table) / sizeof(PassVariant);
}

// PA_SORT_VARIANT (an index into the key type's table) overrides every
// class; otherwise the tuned choice, otherwise the built-in one.  The first
// call loads the tuning cache named by PA_TUNE_CACHE, if any.
template <typename K, bool KV>
int variant_index(size_t n) {
  static const int forced = [] {
    const char* e = std::getenv("PA_SORT_VARIANT");
    const int v = e ? std::atoi(e) : -1;
    return (v >= 0 && (size_t)v < variant_count<K, KV>()) ? v : -1;
  }();
  if (forced >= 0) return forced;
  tune_cache_autoload();
  const int c = size_class(n);
  if constexpr (!KV) {
    const int t = g_tuned[sizeof(K) == 8][c].load(std::memory_order_relaxed);
    if (t >= 0) return t;
  }
  return Variants<K, KV>::by_class[c];
}

// Names of the candidate shapes, as stored in tuning caches.
inline std::string variant_name(const PassVariant& v) {
  char b[96];
  std::snprintf(b, sizeof(b), "%dx%d/lb%d/sp%d/ul%d/rb%d/w%d", v.block, v.items, v.lbw, v.split, v.ul, v.rb, v.minw);
  return b;
}

template <typename K, bool KV>
const PassVariant& pick_variant(size_t n) {
  return Variants<K, KV>::table[variant_index<K, KV>(n)];
}

// Look-back rows (tiles) of one pass.
inline size_t sort_rows(const PassVariant& v, size_t n) {
  const size_t tile = (size_t)v.block * v.items;
  return (n + tile - 1) / tile;
}
constexpr unsigned kPrefixGroups = 128;  // blocks of the pass-0 prefix scan

// Scratch of one sort call with pass variant var (see enqueue_sort for the
// layout).
template <typename K, bool KV>
size_t sort_scratch_bytes(const PassVariant& var, size_t n) {
  const size_t D = sort_passes<K>(var.rb), R = size_t(1) << var.rb;
  const size_t rows = sort_rows(var, n);
  Layout L;
  L.take(D * R * 4);                      // counts
  L.take(D * R * 4);                      // bases
  L.take(sort_tile_counters((int)D) * 4);  // tile counters
  L.take(D * rows * R * 4);               // look-back table
  L.take((size_t)kPrefixGroups * R * 4);  // pass-0 prefix partial sums
  L.take(D * 4);                          // trivial flags of the passes
  return L.off;
}

// The sort with RB-bit digits: D = ceil(key bits / RB) passes.  Scratch
// layout: counts [D][R], digit bases [D][R], tile counters [kXcds + D - 1]
// (pass 0 claims per XCD, the later passes from one counter each), look-back
// rows [D][tiles][R], prefix partials [groups][R], the passes' trivial flags
// [D].  With an odd D the data ends in the temporary arrays and
// is copied back (the roles keep that parity when passes are skipped).
template <typename K, bool KV, int RB>
int enqueue_sort_rb(const PassVariant& var, K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, char* scratch,
                    hipStream_t s, uint32_t xform) {
  constexpr int D = sort_passes<K>(RB);
  constexpr size_t R = size_t(1) << RB;
  const size_t rows = sort_rows(var, n);
  const size_t ncount = (size_t)D * R;
  Layout L;
  uint32_t* counts = reinterpret_cast<uint32_t*>(scratch + L.take(ncount * 4));
  uint32_t* bases = reinterpret_cast<uint32_t*>(scratch + L.take(ncount * 4));
  uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch + L.take(sort_tile_counters(D) * 4));
  const size_t lb_bytes = (size_t)D * rows * R * 4;
  uint32_t* lb = reinterpret_cast<uint32_t*>(scratch + L.take(lb_bytes));
  uint32_t* prefix_partial = reinterpret_cast<uint32_t*>(scratch + L.take((size_t)kPrefixGroups * R * 4));
  uint32_t* trivial = reinterpret_cast<uint32_t*>(scratch + L.take(D * 4));

  {
    ProfScope ps("k_sort_prepare", s);
    hipLaunchKernelGGL(k_sort_prepare, dim3(1), dim3(256), 0, s, counts, (uint32_t)ncount, ctr,
                       sort_tile_counters(D));
  }
  const size_t tile = (size_t)var.block * var.items;
  // One read: all digit histograms, pass 0's per-tile digit counts (into
  // pass 0's look-back rows) and the clearing of the other passes' rows.
  // Two small kernels then turn pass 0's rows into inclusive prefixes, so
  // pass 0 runs without look-back waits; the first also writes the digit
  // bases.
  const size_t pass_rows_bytes = rows * R * 4;
  const size_t zero_bytes = lb_bytes - pass_rows_bytes;
  // Tiles per block: about n / 512 keys per block (two blocks per CU), at
  // least one tile and at most 196,608 keys, so a block's fixed costs
  // (clearing its bins, one global flush) are spread over several tiles.
  // Measured on 100M keys in 24K-key tiles (one read alone): 2 tiles per
  // block 101 us, 4 tiles 98 us, 8 tiles 94 us.
  const size_t target = std::min<size_t>(std::max<size_t>(n / 512, tile), 196608);
  const size_t tpb = std::max<size_t>(1, std::min<size_t>((target + tile / 2) / tile, kHistTilesMaxKeys / tile));
  const size_t grid = std::max<size_t>(1, (rows + tpb - 1) / tpb);
  {
    ProfScope ps("k_hist_digits", s);
    hipLaunchKernelGGL((k_hist_tiles<K, RB>), dim3((unsigned)grid), dim3(kSortBlock), 0, s, keys, n, (uint32_t)tile,
                       lb, counts, reinterpret_cast<uint4*>(lb + rows * R), zero_bytes / 16, xform);
  }
  const uint32_t groups = (uint32_t)std::min<size_t>(rows, kPrefixGroups);
  const uint32_t per = (uint32_t)((rows + groups - 1) / groups);
  {
    ProfScope ps("k_tile_prefix", s);
    hipLaunchKernelGGL(k_tile_sums<RB>, dim3(groups + D), dim3(kPrefixThreads), 0, s, lb, (uint32_t)rows, per,
                       prefix_partial, groups, counts, bases, (uint32_t)n, trivial);
    hipLaunchKernelGGL(k_tile_scan<RB>, dim3(groups), dim3(kPrefixThreads), 0, s, lb, (uint32_t)rows, per,
                       prefix_partial, trivial, (uint32_t)D, xform == 0, ctr);
  }
  K* kin = keys;
  K* kout = ktmp;
  uint32_t* vin = vals;
  uint32_t* vout = vtmp;
  for (int p = 0; p < D; ++p) {
    uint32_t fmode = 0;
    if (p == 0) fmode |= kPrefixReady;
    if (xform && p == 0) fmode |= kFloatIn;
    if (xform && p == D - 1) fmode |= kFloatOut;
    if (xform == dev::kXformSigned) fmode |= kSignedKeys;
    PassArgs a{kin, kout, vin, vout, (uint32_t)n, bases + (size_t)p * R, lb + (size_t)p * rows * R, ctr + (p == 0 ? 0 : kXcds - 1 + p),
               (uint32_t)rows, (uint32_t)(RB * p), fmode};
    PA_TRY(var.launch(s, a));
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (D % 2) {
    // The result is in the temporary arrays.
    PA_HIP(hipMemcpyAsync(keys, ktmp, n * sizeof(K), hipMemcpyDeviceToDevice, s));
    if constexpr (KV) PA_HIP(hipMemcpyAsync(vals, vtmp, n * 4, hipMemcpyDeviceToDevice, s));
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}

template <typename K, bool KV>
int enqueue_sort(const PassVariant& var, K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, char* scratch,
                 hipStream_t s, uint32_t xform) {
  switch (var.rb) {
    case 8: return enqueue_sort_rb<K, KV, 8>(var, keys, vals, ktmp, vtmp, n, scratch, s, xform);
    case 7: return enqueue_sort_rb<K, KV, 7>(var, keys, vals, ktmp, vtmp, n, scratch, s, xform);
    case 6: return enqueue_sort_rb<K, KV, 6>(var, keys, vals, ktmp, vtmp, n, scratch, s, xform);
    case 5: return enqueue_sort_rb<K, KV, 5>(var, keys, vals, ktmp, vtmp, n, scratch, s, xform);
    case 4: return enqueue_sort_rb<K, KV, 4>(var, keys, vals, ktmp, vtmp, n, scratch, s, xform);
    default: return arg_error("sort: digits of 4 to 8 bits");
  }
}

int launch_hist_digits_u32(const uint32_t* keys, size_t n, uint32_t* counts, hipStream_t s) {
  const size_t nvec = n / 4 + 1;
  unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>((nvec + kSortBlock - 1) / kSortBlock, 2048));
  ProfScope ps("k_hist_digits", s);
  hipLaunchKernelGGL(k_hist_digits<uint32_t>, dim3(grid), dim3(kSortBlock), 0, s, keys, n, counts, (uint4*)nullptr,
                     (size_t)0, 0u);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// Elementwise key transform (forward or inverse) over a device array.
template <typename K>
__global__ __launch_bounds__(256) void k_key_xform(K* __restrict__ keys, size_t n, uint32_t xform, uint32_t inverse) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    keys[i] = inverse ? dev::key_inv<K>(keys[i], xform) : dev::key_fwd<K>(keys[i], xform);
}

// Arrays of 2^30 keys or more (past the 30-bit count of a look-back status
// word): sort as few equal chunks of at most 2^30 - 1 keys as cover the array,
// then merge adjacent sorted runs pairwise with the merge-path kernels until
// one run is left -- the shape of the reference's radix-leaf merge sort
// (parallel_merge_sort_hybrid_radix, ParallelMergeSort.h:255-295).
// Transformed orders (float, signed) are applied once before and undone once
// after, so chunks sort and merge as plain unsigned keys.  tmp holds n keys.
constexpr size_t kLargeChunk = kMaxSortN;

inline size_t large_chunk_keys(size_t n) {
  const size_t chunks = (n + kLargeChunk - 1) / kLargeChunk;
  return (n + chunks - 1) / chunks;
}

template <typename K, bool KV>
size_t large_sort_scratch_bytes(size_t n) {
  const size_t c = large_chunk_keys(n);
  return sort_scratch_bytes<K, KV>(pick_variant<K, KV>(c), c) + 256 + merge_split_bytes(n);
}

// Merges adjacent sorted runs of keys (covering [0, n) in order) pairwise,
// round after round, ping-ponging with tmp, until one run is left; the result
// ends in keys.  An odd run out is copied through its round.  Ties take the
// earlier run, so stable runs merge stably.  split: merge_split_bytes(n).
template <typename K, bool KV>
int merge_rounds(K* keys, uint32_t* vals, K* tmp, uint32_t* vtmp, size_t n,
                 std::vector<std::pair<size_t, size_t>> runs, size_t* split, hipStream_t s) {
  K* x = keys;
  K* y = tmp;
  uint32_t* vx = vals;
  uint32_t* vy = vtmp;
  while (runs.size() > 1) {
    std::vector<std::pair<size_t, size_t>> next;
    for (size_t i = 0; i + 1 < runs.size(); i += 2) {
      const auto [o1, l1] = runs[i];
      const auto [o2, l2] = runs[i + 1];
      if constexpr (KV)
        PA_TRY(merge_kv_enqueue<K>(x + o1, vx + o1, l1, x + o2, vx + o2, l2, y + o1, vy + o1, split, s));
      else
        PA_TRY(merge_enqueue<K>(x + o1, l1, x + o2, l2, y + o1, split, s));
      next.emplace_back(o1, l1 + l2);
    }
    if (runs.size() % 2) {
      const auto [o, l] = runs.back();
      PA_HIP(hipMemcpyAsync(y + o, x + o, l * sizeof(K), hipMemcpyDeviceToDevice, s));
      if constexpr (KV) PA_HIP(hipMemcpyAsync(vy + o, vx + o, l * 4, hipMemcpyDeviceToDevice, s));
      next.push_back(runs.back());
    }
    runs.swap(next);
    std::swap(x, y);
    std::swap(vx, vy);
  }
  if (x != keys) {
    PA_HIP(hipMemcpyAsync(keys, x, n * sizeof(K), hipMemcpyDeviceToDevice, s));
    if constexpr (KV) PA_HIP(hipMemcpyAsync(vals, vx, n * 4, hipMemcpyDeviceToDevice, s));
  }
  return PA_OK;
}

// Key-value sorts take the same route: the chunks' stable sorts, then stable
// merges (ties take the earlier run), so equal keys keep their input order.
template <typename K, bool KV>
int enqueue_large_sort(K* keys, uint32_t* vals, K* tmp, uint32_t* vtmp, size_t n, char* scratch, hipStream_t s,
                       uint32_t xform) {
  const size_t csize = large_chunk_keys(n);
  const PassVariant& var = pick_variant<K, KV>(csize);
  const size_t sort_bytes = sort_scratch_bytes<K, KV>(var, csize);
  size_t* split = reinterpret_cast<size_t*>(scratch + ((sort_bytes + 255) & ~size_t(255)));
  const unsigned xgrid = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
  if (xform) hipLaunchKernelGGL(k_key_xform<K>, dim3(xgrid), dim3(256), 0, s, keys, n, xform, 0u);
  std::vector<std::pair<size_t, size_t>> runs;  // (first key, length)
  for (size_t off = 0; off < n; off += csize) {
    const size_t len = std::min(csize, n - off);
    // Each chunk's result is in keys (an odd pass count copies it back).
    PA_TRY((enqueue_sort<K, KV>(var, keys + off, KV ? vals + off : nullptr, tmp + off, KV ? vtmp + off : nullptr, len,
                                scratch, s, dev::kXformNone)));
    runs.emplace_back(off, len);
  }
  PA_TRY((merge_rounds<K, KV>(keys, vals, tmp, vtmp, n, runs, split, s)));
  if (xform) hipLaunchKernelGGL(k_key_xform<K>, dim3(xgrid), dim3(256), 0, s, keys, n, xform, 1u);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

size_t large_sort_u32_scratch_bytes(size_t n) { return large_sort_scratch_bytes<uint32_t, false>(n); }

int large_sort_u32_enqueue(uint32_t* keys, uint32_t* tmp, size_t n, char* scratch, hipStream_t s) {
  return enqueue_large_sort<uint32_t, false>(keys, nullptr, tmp, nullptr, n, scratch, s, dev::kXformNone);
}

// ---------------------------------------------------------------------------
// Allocation-free u32 sort (pa_sort_u32_noalloc).  The sort's scratch (digit
// counts and bases, tile counters, look-back table, prefix partials) lives in
// a file-scope device array instead of a per-stream allocation; every word a
// sort reads is written first by its own kernels (k_sort_prepare clears the
// counters, k_hist_tiles writes pass 0's rows and clears the others), so the
// contents left by an earlier call do not matter and nothing is cleared with
// a memset.  The array's device address is looked up once per device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocMaxKeys = size_t(1) << 27;
// Covers the built-in pass shape of every size class up to kNoAllocMaxKeys
// (2^27 keys in 24,576-key tiles: 5,462 look-back rows per pass, 21.4 MB).
constexpr size_t kNoAllocScratchBytes = size_t(22) << 20;
static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// Device address of the current device's copy of noalloc_scratch.
static int noalloc_scratch_ptr(char** out) {
  int dev = 0;
  PA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return arg_error("pa_sort_u32_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    PA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    if (!p) return arg_error("pa_sort_u32_noalloc: no device address for the scratch array");
    g_noalloc_ptr[dev] = p;
  }
  *out = static_cast<char*>(g_noalloc_ptr[dev]);
  return PA_OK;
}

// Scratch layout of one sort call: the sort's own scratch, then staging for
// whichever of keys / tmp / values / value tmp are not device memory.
struct SortEntryLayout {
  size_t o_sort, o_k, o_kt, o_v, o_vt, total;
};
template <typename K, bool KV>
SortEntryLayout sort_entry_layout(size_t n, bool dk, bool dkt, bool dv, bool dvt) {
  const bool large = n > kMaxSortN;
  const PassVariant& var = pick_variant<K, KV>(large ? large_chunk_keys(n) : n);
  SortEntryLayout Y;
  Layout L;
  Y.o_sort = L.take(large ? large_sort_scratch_bytes<K, KV>(n) : sort_scratch_bytes<K, KV>(var, n));
  Y.o_k = dk ? 0 : L.take(n * sizeof(K));
  Y.o_kt = dkt ? 0 : L.take(n * sizeof(K));
  Y.o_v = dv ? 0 : L.take(n * 4);
  Y.o_vt = dvt ? 0 : L.take(n * 4);
  Y.total = L.off;
  return Y;
}

size_t sort_call_scratch_bytes(bool u64, size_t n) {
  if (n == 0) return 0;
  return u64 ? sort_entry_layout<uint64_t, false>(n, true, false, true, true).total
             : sort_entry_layout<uint32_t, false>(n, true, false, true, true).total;
}

// Host entry: classifies pointers, stages host arrays, runs, copies back.
template <typename K, bool KV>
int sort_entry(K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, hipStream_t s, uint32_t xform,
               const char* who) {
  if (n == 0) return PA_OK;
  if (!keys || (KV && !vals)) return arg_error(who);
  const bool large = n > kMaxSortN;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = is_device_ptr(keys);
  const bool dv = !KV || is_device_ptr(vals);
  const bool dkt = ktmp && is_device_ptr(ktmp);
  const bool dvt = !KV || (vtmp && is_device_ptr(vtmp));
  const PassVariant& var = pick_variant<K, KV>(large ? large_chunk_keys(n) : n);
  const SortEntryLayout Y = sort_entry_layout<K, KV>(n, dk, dkt, dv, dvt);
  const size_t o_sort = Y.o_sort, o_k = Y.o_k, o_kt = Y.o_kt, o_v = Y.o_v, o_vt = Y.o_vt;
  void* base = nullptr;
  PA_TRY(ensure_scratch(ctx, Y.total, &base));
  char* b = static_cast<char*>(base);
  K* dkeys = dk ? keys : reinterpret_cast<K*>(b + o_k);
  K* dktmp = dkt ? ktmp : reinterpret_cast<K*>(b + o_kt);
  uint32_t* dvals = KV ? (dv ? vals : reinterpret_cast<uint32_t*>(b + o_v)) : nullptr;
  uint32_t* dvtmp = KV ? (dvt ? vtmp : reinterpret_cast<uint32_t*>(b + o_vt)) : nullptr;
  if (!dk) PA_TRY(stage_to_device(ctx, dkeys, keys, n * sizeof(K)));
  if (KV && !dv) PA_TRY(stage_to_device(ctx, dvals, vals, n * 4));
  if constexpr (!KV) {
    if (std::is_same<K, uint32_t>::value && xform == dev::kXformNone && n <= small_sort_max_keys()) {
      // One workgroup sorts the whole array in LDS: one launch.
      PA_TRY(small_sort_enqueue_u32(reinterpret_cast<uint32_t*>(dkeys), n, s));
    } else if (large) {
      PA_TRY((enqueue_large_sort<K, false>(dkeys, nullptr, dktmp, nullptr, n, b + o_sort, s, xform)));
    } else {
      PA_TRY((enqueue_sort<K, KV>(var, dkeys, dvals, dktmp, dvtmp, n, b + o_sort, s, xform)));
    }
  } else if (large) {
    PA_TRY((enqueue_large_sort<K, true>(dkeys, dvals, dktmp, dvtmp, n, b + o_sort, s, xform)));
  } else {
    PA_TRY((enqueue_sort<K, KV>(var, dkeys, dvals, dktmp, dvtmp, n, b + o_sort, s, xform)));
  }
  if (!dk) PA_TRY(stage_to_host(ctx, keys, dkeys, n * sizeof(K)));
  if (KV && !dv) PA_TRY(stage_to_host(ctx, vals, dvals, n * 4));
  return PA_OK;
}

// Merge sort: the leaves keys[i * leaf, (i + 1) * leaf) radix-sorted (u32:
// as one segmented batch; u64: one onesweep sort each), then merged upward
// pairwise with merge-path merges -- the shape of the reference's
// parallel_merge_sort_hybrid_radix (ParallelMergeSort.h:255-295: radix-sorted
// leaves, merge_parallel_L5 on the way up).  Host arrays are staged.
template <typename K>
int merge_sort_entry(K* keys, K* tmp, size_t n, size_t leaf, hipStream_t s, const char* who) {
  if (n == 0) return PA_OK;
  if (!keys) return arg_error(who);
  if (leaf == 0) leaf = std::max<size_t>(size_t(1) << 16, (n + 7) / 8);
  leaf = std::min(leaf, n);
  if (sizeof(K) == 8) leaf = std::min(leaf, kMaxSortN);
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = is_device_ptr(keys);
  const bool dkt = tmp && is_device_ptr(tmp);
  size_t leaf_bytes;
  if constexpr (sizeof(K) == 4)
    leaf_bytes = leaves_scratch_bytes_u32(n, leaf);
  else
    leaf_bytes = sort_scratch_bytes<K, false>(pick_variant<K, false>(leaf), leaf);
  Layout L;
  const size_t o_leaf = L.take(leaf_bytes);
  const size_t o_split = L.take(merge_split_bytes(n));
  const size_t o_k = dk ? 0 : L.take(n * sizeof(K));
  const size_t o_kt = dkt ? 0 : L.take(n * sizeof(K));
  void* base = nullptr;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* b = static_cast<char*>(base);
  K* dkeys = dk ? keys : reinterpret_cast<K*>(b + o_k);
  K* dtmp = dkt ? tmp : reinterpret_cast<K*>(b + o_kt);
  if (!dk) PA_TRY(stage_to_device(ctx, dkeys, keys, n * sizeof(K)));
  std::vector<std::pair<size_t, size_t>> runs;
  for (size_t o = 0; o < n; o += leaf) runs.emplace_back(o, std::min(leaf, n - o));
  if constexpr (sizeof(K) == 4) {
    PA_TRY(leaves_enqueue_u32(reinterpret_cast<uint32_t*>(dkeys), reinterpret_cast<uint32_t*>(dtmp), n, leaf,
                              b + o_leaf, s, ctx));
  } else {
    for (const auto& r : runs) {
      const PassVariant& var = pick_variant<K, false>(r.second);
      PA_TRY((enqueue_sort<K, false>(var, dkeys + r.first, nullptr, dtmp + r.first, nullptr, r.second, b + o_leaf, s,
                                     dev::kXformNone)));
    }
  }
  PA_TRY((merge_rounds<K, false>(dkeys, nullptr, dtmp, nullptr, n, runs, reinterpret_cast<size_t*>(b + o_split), s)));
  PA_HIP(hipGetLastError());
  if (!dk) PA_TRY(stage_to_host(ctx, keys, dkeys, n * sizeof(K)));
  return PA_OK;
}

// Times every candidate shape of K keys at one size per class (up to max_n
// keys) on a fresh random array and records the fastest per class.  Each
// timing covers the whole sort, including the copy back of an odd pass count.
template <typename K>
int autotune_keys(size_t max_n, hipStream_t s) {
  if (max_n < 2) return arg_error("pa_autotune_sort: need max_n >= 2");
  static const size_t rep[kClasses] = {size_t(1) << 16, size_t(1) << 20, size_t(1) << 24, size_t(100000000)};
  const size_t nmax = std::min(max_n, std::min(rep[kClasses - 1], kMaxSortN));
  constexpr size_t NV = variant_count<K, false>();
  const int target = sizeof(K) == 8;
  size_t scratch_bytes = 0;
  for (size_t v = 0; v < NV; ++v)
    scratch_bytes = std::max(scratch_bytes, sort_scratch_bytes<K, false>(Variants<K, false>::table[v], nmax));
  K *pristine = nullptr, *keys = nullptr, *tmp = nullptr;
  char* scratch = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = PA_OK;
  auto fail = [&](hipError_t e, const char* what) {
    st = hip_fail(e, what, __FILE__, __LINE__);
    return st;
  };
  do {
    hipError_t e;
    if ((e = hipMalloc(&pristine, nmax * sizeof(K))) != hipSuccess || (e = hipMalloc(&keys, nmax * sizeof(K))) != hipSuccess ||
        (e = hipMalloc(&tmp, nmax * sizeof(K))) != hipSuccess || (e = hipMalloc(&scratch, scratch_bytes)) != hipSuccess) {
      fail(e, "hipMalloc");
      break;
    }
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
      fail(e, "hipEventCreate");
      break;
    }
    if constexpr (sizeof(K) == 8)
      st = pa_fill_random_u64(reinterpret_cast<uint64_t*>(pristine), nmax, 0x5eed, s);
    else
      st = pa_fill_random_u32(reinterpret_cast<uint32_t*>(pristine), nmax, 0x5eed, s);
    if (st != PA_OK) break;
    size_t prev_n = 0;
    int prev_best = -1;
    for (int c = 0; c < kClasses && st == PA_OK; ++c) {
      const size_t n = std::min(rep[c], nmax);
      if (n == prev_n) {
        g_tuned[target][c].store(prev_best);
        continue;
      }
      float best_ms = 0.f;
      int best = -1;
      for (size_t v = 0; v < NV && st == PA_OK; ++v) {
        const PassVariant& var = Variants<K, false>::table[v];
        float total = 0.f;
        for (int rep_i = 0; rep_i < 4 && st == PA_OK; ++rep_i) {  // first run warms up, three are timed
          if ((e = hipMemcpyAsync(keys, pristine, n * sizeof(K), hipMemcpyDeviceToDevice, s)) != hipSuccess) {
            fail(e, "hipMemcpyAsync");
            break;
          }
          (void)hipEventRecord(e0, s);
          if ((st = enqueue_sort<K, false>(var, keys, nullptr, tmp, nullptr, n, scratch, s, dev::kXformNone)) != PA_OK)
            break;
          (void)hipEventRecord(e1, s);
          if ((e = hipEventSynchronize(e1)) != hipSuccess) {
            fail(e, "hipEventSynchronize");
            break;
          }
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (rep_i > 0) total += ms;
        }
        if (st == PA_OK && (best < 0 || total < best_ms)) {
          best_ms = total;
          best = (int)v;
        }
      }
      if (st == PA_OK) {
        g_tuned[target][c].store(best);
        prev_n = n;
        prev_best = best;
      }
    }
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(pristine);
  (void)hipFree(keys);
  (void)hipFree(tmp);
  (void)hipFree(scratch);
  return st;
}

// The keys-only sort targets of the tuning state and caches.
template <typename K>
static TuneTarget keys_target(const char* key) {
  TuneTarget t;
  t.key = key;
  t.classes = kClasses;
  t.count = [] { return (int)variant_count<K, false>(); };
  t.name = [](int i) { return variant_name(Variants<K, false>::table[i]); };
  t.get = [](int c) { return g_tuned[sizeof(K) == 8][c].load(); };
  t.set = [](int c, int i) { g_tuned[sizeof(K) == 8][c].store(i); };
  return t;
}

const TuneTarget* sort_tune_target(int target) {
  static const TuneTarget t32 = keys_target<uint32_t>("u32");
  static const TuneTarget t64 = keys_target<uint64_t>("u64");
  return target == PA_TUNE_U32 ? &t32 : target == PA_TUNE_U64 ? &t64 : nullptr;
}

template <typename K>
static void fill_shape(const PassVariant& v, pa_sort_shape* out) {
  out->block = v.block;
  out->items = v.items;
  out->lookback = v.lbw;
  out->split = v.split;
  out->lanes_per_unit = v.ul;
  out->digit_bits = v.rb;
  out->passes = sort_passes<K>(v.rb);
  out->queue_max_keys = 0;
}

int sort_shape_info(int target, int index, pa_sort_shape* out) {
  if (target == PA_TUNE_U32) {
    if (index < 0 || (size_t)index >= variant_count<uint32_t, false>()) return arg_error("pa_sort_shape_info: index");
    fill_shape<uint32_t>(Variants<uint32_t, false>::table[index], out);
  } else {
    if (index < 0 || (size_t)index >= variant_count<uint64_t, false>()) return arg_error("pa_sort_shape_info: index");
    fill_shape<uint64_t>(Variants<uint64_t, false>::table[index], out);
  }
  return PA_OK;
}

int sort_shape_choice(int target, size_t n) {
  return target == PA_TUNE_U64 ? variant_index<uint64_t, false>(n) : variant_index<uint32_t, false>(n);
}

int autotune_sort(int target, size_t max_n, hipStream_t s) {
  return target == PA_TUNE_U64 ? autotune_keys<uint64_t>(max_n, s) : autotune_keys<uint32_t>(max_n, s);
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_sort_u32(uint32_t* keys, uint32_t* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, false>(keys, nullptr, tmp, nullptr, n, (hipStream_t)stream, dev::kXformNone, "pa_sort_u32");
}

int pa_sort_u32_noalloc(uint32_t* keys, uint32_t* tmp, size_t n, pa_stream_t stream) {
  if (n == 0) return PA_OK;
  if (!keys || !tmp) return arg_error("pa_sort_u32_noalloc: keys and tmp are required");
  if (n > kNoAllocMaxKeys) {
    set_last_error("pa_sort_u32_noalloc: n must be <= 2^27 (134,217,728) keys");
    return PA_ERR_TOO_LARGE;
  }
  if (!is_device_ptr(keys) || !is_device_ptr(tmp))
    return arg_error("pa_sort_u32_noalloc: keys and tmp must be device memory");
  hipStream_t s = (hipStream_t)stream;
  ProfSuppress no_timing;  // timing events would be allocations
  if (n <= small_sort_max_keys()) return small_sort_enqueue_u32(keys, n, s);
  const PassVariant& var = pick_variant<uint32_t, false>(n);
  if (sort_scratch_bytes<uint32_t, false>(var, n) > kNoAllocScratchBytes) {
    set_last_error("pa_sort_u32_noalloc: the pass shape chosen for n needs more than the 22 MiB static scratch");
    return PA_ERR_TOO_LARGE;
  }
  char* scratch = nullptr;
  PA_TRY(noalloc_scratch_ptr(&scratch));
  return enqueue_sort<uint32_t, false>(var, keys, nullptr, tmp, nullptr, n, scratch, s, dev::kXformNone);
}

int pa_sort_u64(uint64_t* keys, uint64_t* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint64_t, false>(keys, nullptr, tmp, nullptr, n, (hipStream_t)stream, dev::kXformNone, "pa_sort_u64");
}

int pa_sort_f32(float* keys, float* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, false>(reinterpret_cast<uint32_t*>(keys), nullptr, reinterpret_cast<uint32_t*>(tmp),
                                     nullptr, n, (hipStream_t)stream, dev::kXformFloat, "pa_sort_f32");
}

int pa_sort_f64(double* keys, double* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint64_t, false>(reinterpret_cast<uint64_t*>(keys), nullptr, reinterpret_cast<uint64_t*>(tmp),
                                     nullptr, n, (hipStream_t)stream, dev::kXformFloat, "pa_sort_f64");
}

int pa_sort_i32(int32_t* keys, int32_t* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, false>(reinterpret_cast<uint32_t*>(keys), nullptr, reinterpret_cast<uint32_t*>(tmp),
                                     nullptr, n, (hipStream_t)stream, dev::kXformSigned, "pa_sort_i32");
}

int pa_sort_i64(int64_t* keys, int64_t* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint64_t, false>(reinterpret_cast<uint64_t*>(keys), nullptr, reinterpret_cast<uint64_t*>(tmp),
                                     nullptr, n, (hipStream_t)stream, dev::kXformSigned, "pa_sort_i64");
}

int pa_sort_kv_u32_u32(uint32_t* keys, uint32_t* vals, uint32_t* ktmp, uint32_t* vtmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, true>(keys, vals, ktmp, vtmp, n, (hipStream_t)stream, dev::kXformNone, "pa_sort_kv_u32_u32");
}

int pa_sort_kv_u64_u32(uint64_t* keys, uint32_t* vals, uint64_t* ktmp, uint32_t* vtmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint64_t, true>(keys, vals, ktmp, vtmp, n, (hipStream_t)stream, dev::kXformNone, "pa_sort_kv_u64_u32");
}

int pa_autotune_sort_u32(size_t max_n, pa_stream_t stream) {
  return pa_autotune_sort(PA_TUNE_U32, max_n, stream);
}

int pa_sort_variant_u32(size_t n) { return variant_index<uint32_t, false>(n); }

int pa_autotune_reset(void) {
  for (int t = 0; t < 2; ++t)
    for (int c = 0; c < kClasses; ++c) g_tuned[t][c].store(-1);
  seg_tune_reset();
  return PA_OK;
}

int pa_merge_sort_u32(uint32_t* keys, uint32_t* tmp, size_t n, size_t leaf_keys, pa_stream_t stream) {
  return merge_sort_entry<uint32_t>(keys, tmp, n, leaf_keys, (hipStream_t)stream, "pa_merge_sort_u32");
}

int pa_merge_sort_u64(uint64_t* keys, uint64_t* tmp, size_t n, size_t leaf_keys, pa_stream_t stream) {
  return merge_sort_entry<uint64_t>(keys, tmp, n, leaf_keys, (hipStream_t)stream, "pa_merge_sort_u64");
}

int pa_sort_two_phase_u32(uint32_t* a, uint32_t* b, size_t n, pa_stream_t stream) {
  if (n == 0) return PA_OK;
  if (!a || !b) return arg_error("pa_sort_two_phase_u32: null pointer");
  hipStream_t s = (hipStream_t)stream;
  PA_TRY(pa_sort_u32(a, b, n, stream));
  const bool da = is_device_ptr(a), db = is_device_ptr(b);
  hipMemcpyKind kind = da ? (db ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost)
                          : (db ? hipMemcpyHostToDevice : hipMemcpyHostToHost);
  PA_HIP(hipMemcpyAsync(b, a, n * sizeof(uint32_t), kind, s));
  if (!da || !db) PA_HIP(hipStreamSynchronize(s));
  return PA_OK;
}

}  // extern "C"
