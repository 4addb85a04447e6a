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
//   k_hist_digits    one read of the keys -> D x 256 counts (the TwoPhase
//                    idea); LDS-privatised bins, wave-uniform fast path for
//                    constant/presorted runs; also zeroes the look-back table.
//   k_onesweep       per digit: a tile of 512 x ITEMS keys is ranked stably in
//                    LDS with wave64 ballot matching, each tile's per-bin count
//                    is published to a decoupled look-back table (replacing the
//                    serial quantum x bin prefix), and keys are staged in LDS in
//                    digit order so the global writes are runs per bin (the GPU
//                    form of the 64-deep de-randomisation buffers).
//
// HBM layout: keys (n x K) ping-pong between `keys` and `tmp`; D passes with
// D = sizeof(K) is even, so the result lands back in `keys`.  Scratch per
// call: counts D x 256 u32, tile counters D u32, look-back D x tiles x 256 u32.
#include <algorithm>
#include <atomic>
#include <type_traits>
#include <cstdlib>
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

// Pass-0 prefixes.  rows[t][d] holds digit d's count in pass-0 tile t; the
// two kernels below turn every row into the inclusive status word of its tile
// (count of digit d in tiles 0..t), as if every tile had already resolved its
// look-back.  Group g of blocks owns tiles [g * per, (g + 1) * per).  Blocks
// have 1024 threads: thread (q, d) = (tid >> 8, tid & 255) handles digit d for
// a quarter of the group's tiles, so each thread keeps several independent
// loads in flight instead of walking the rows one after another.
constexpr unsigned kPrefixThreads = 1024;

// Step 1: partial[g][d] = count of digit d over group g's tiles (blocks
// g < groups).  The D blocks after them write the digit bases of pass p =
// blockIdx.x - groups: base[p][d] = sum_{d' < d} counts[p][d'], the first
// output slot of digit d.
__global__ __launch_bounds__(kPrefixThreads) void k_tile_sums(const uint32_t* __restrict__ rows, uint32_t ntiles,
                                                              uint32_t per, uint32_t* __restrict__ partial,
                                                              uint32_t groups, const uint32_t* __restrict__ counts,
                                                              uint32_t* __restrict__ base) {
  __shared__ uint32_t s_q[4][256];
  const uint32_t d = threadIdx.x & 255, q = threadIdx.x >> 8;
  if (blockIdx.x >= groups) {
    // Threads 0..255 (waves 0..3) scan the pass's 256 counts.
    const uint32_t p = blockIdx.x - groups, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t c = 0, inc = 0;
    if (q == 0) {
      c = counts[p * 256 + d];
      inc = dev::wave_incl_scan(c, lane);
      if (lane == 63) s_q[0][wave] = inc;
    }
    __syncthreads();
    if (q == 0) {
      uint32_t pre = 0;
      for (uint32_t w = 0; w < wave; ++w) pre += s_q[0][w];
      base[p * 256 + d] = pre + inc - c;
    }
    return;
  }
  const uint32_t t0 = blockIdx.x * per, t1 = min(ntiles, t0 + per);
  uint32_t acc = 0;
#pragma unroll 8
  for (uint32_t t = t0 + q; t < t1; t += 4) acc += rows[(size_t)t * 256 + d];
  s_q[q][d] = acc;
  __syncthreads();
  if (q == 0) partial[blockIdx.x * 256 + d] = s_q[0][d] + s_q[1][d] + s_q[2][d] + s_q[3][d];
}

// Step 2: group g adds the partials of groups < g, then scans its own tiles
// in chunks of 4 x 8: thread (q, d) loads 8 consecutive rows at once.
__global__ __launch_bounds__(kPrefixThreads) void k_tile_scan(uint32_t* __restrict__ rows, uint32_t ntiles,
                                                              uint32_t per, const uint32_t* __restrict__ partial) {
  __shared__ uint32_t s_q[4][256];
  const uint32_t d = threadIdx.x & 255, q = threadIdx.x >> 8;
  uint32_t acc = 0;
#pragma unroll 8
  for (uint32_t g = q; g < blockIdx.x; g += 4) acc += partial[g * 256 + d];
  s_q[q][d] = acc;
  __syncthreads();
  uint32_t base = s_q[0][d] + s_q[1][d] + s_q[2][d] + s_q[3][d];
  const uint32_t t0 = blockIdx.x * per, t1 = min(ntiles, t0 + per);
  for (uint32_t c0 = t0; c0 < t1; c0 += 32) {
    const uint32_t tq = c0 + 8 * q;
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = tq + j < t1 ? rows[(size_t)(tq + j) * 256 + d] : 0u;
#pragma unroll
    for (int j = 1; j < 8; ++j) v[j] += v[j - 1];
    __syncthreads();  // every thread has read s_q of the previous step
    s_q[q][d] = v[7];
    __syncthreads();
    uint32_t off = base;
    for (uint32_t w = 0; w < q; ++w) off += s_q[w][d];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (tq + j < t1) rows[(size_t)(tq + j) * 256 + d] = kLbInc | (off + v[j]);
    base += s_q[0][d] + s_q[1][d] + s_q[2][d] + s_q[3][d];
  }
}

// ---------------------------------------------------------------------------
// Chained passes: per-chain digit counts, bases and the interleaved claim order
// ---------------------------------------------------------------------------
// Counts row p (p < D) is [kChains][256]: chain-major counts of digit p, where
// the chain of a key is its position >> log2s for p = 0 (contiguous position
// ranges of the input) and digit p-1 >> kChainShift for p >= 1.  Pass p's input
// is sorted by digit p-1, so the keys of value chain c are the contiguous range
// of that input holding digits [32c, 32c + 32) of digit p-1.  Every row depends
// only on the multiset of keys, so one read of the input yields all of them.
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_hist_chains(const K* __restrict__ keys, size_t n, uint32_t log2s,
                                                            uint32_t* __restrict__ partial,
                                                            uint4* __restrict__ zero_ptr, size_t zero_vec,
                                                            uint32_t xform) {
  constexpr int D = sizeof(K);
  constexpr int PER = 16 / sizeof(K);
  constexpr int ROW = kChains * 256;
  __shared__ uint32_t s_hist[D * ROW];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < D * ROW; i += kSortBlock) s_hist[i] = 0;
  const size_t gstride = (size_t)gridDim.x * kSortBlock;
  for (size_t i = (size_t)blockIdx.x * kSortBlock + tid; i < zero_vec; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();

  auto count = [&](const K (&k)[PER], const size_t (&pos)[PER]) {
    uint32_t idx[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) idx[j] = (uint32_t)(pos[j] >> log2s) * 256 + dev::digit8<K>(k[j], 0);
    dev::lds_count<PER>(s_hist, idx);
#pragma unroll
    for (int d = 1; d < D; ++d) {
#pragma unroll
      for (int j = 0; j < PER; ++j)
        idx[j] = (dev::digit8<K>(k[j], 8 * (d - 1)) >> kChainShift) * 256 + dev::digit8<K>(k[j], 8 * d);
      dev::lds_count<PER>(s_hist + d * ROW, idx);
    }
  };

  size_t head = ((16 - ((uintptr_t)keys & 15)) & 15) / sizeof(K);
  if (head > n) head = n;
  const size_t nvec = (n - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(keys + head);
  constexpr int U = 4;  // 16-byte loads in flight per thread
  const size_t v0 = (size_t)blockIdx.x * kSortBlock + tid;
  for (size_t base = 0; base < nvec; base += gstride * U) {
    uint4 x[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t v = base + u * gstride + v0;
      ok[u] = v < nvec;
      if (ok[u]) x[u] = vk[v];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      K k[PER];
      size_t pos[PER];
      __builtin_memcpy(k, &x[u], 16);
      const size_t p0 = head + (base + u * gstride + v0) * PER;
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        pos[j] = p0 + j;
        if (xform) k[j] = dev::key_fwd<K>(k[j], xform);
      }
      count(k, pos);
    }
  }
  if (blockIdx.x == 0) {  // head and tail elements outside whole vectors
    const size_t tail_start = head + nvec * PER;
    const size_t extra = head + (n - tail_start);
    for (size_t e = tid; e < extra; e += kSortBlock) {
      const size_t idx = e < head ? e : tail_start + (e - head);
      K kk = keys[idx];
      if (xform) kk = dev::key_fwd<K>(kk, xform);
      atomicAdd(&s_hist[(uint32_t)(idx >> log2s) * 256 + dev::digit8<K>(kk, 0)], 1u);
#pragma unroll
      for (int d = 1; d < D; ++d)
        atomicAdd(&s_hist[d * ROW + (dev::digit8<K>(kk, 8 * (d - 1)) >> kChainShift) * 256 + dev::digit8<K>(kk, 8 * d)],
                  1u);
    }
  }
  __syncthreads();
  uint32_t* out = partial + (size_t)blockIdx.x * D * ROW;
  for (uint32_t i = tid; i < D * ROW; i += kSortBlock) out[i] = s_hist[i];
}

// counts[i] = sum over the histogram blocks of partial[b][i]; blockIdx.y takes
// a share of the blocks and adds it atomically.
__global__ __launch_bounds__(256) void k_reduce_counts(const uint32_t* __restrict__ partial, uint32_t nparts,
                                                       uint32_t ncounts, uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ncounts) return;
  const uint32_t per = (nparts + gridDim.y - 1) / gridDim.y;
  const uint32_t b0 = blockIdx.y * per, b1 = min(nparts, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; ++b) sum += partial[(size_t)b * ncounts + i];
  if (sum) atomicAdd(&counts[i], sum);
}

// One block per pass p: chain bases, chain ranges and the claim order.
//   bases[p][c][d] = G_p[d] + sum_{c' < c} counts[p][c'][d], where G_p[d] is the
//   number of keys whose digit p is below d: chain c's keys of digit d follow
//   those of every earlier chain (pass p's input holds the chains in order and
//   the pass is stable).
//   info[p][4c..4c+2] = (first key, keys, first look-back row) of chain c;
//   info[p][32] = tiles in the pass.
//   tiles[p][r] = (c << 24) | k: tile k of chain c is claimed r-th, in the order
//   of (2k + 1) / (2 T_c) (T_c = tiles of chain c), ties by chain, so every
//   chain's tiles are spread evenly over the pass.
__global__ __launch_bounds__(1024) void k_chain_plan(const uint32_t* __restrict__ counts, uint32_t n, uint32_t log2s,
                                                     uint32_t tile, uint32_t* __restrict__ bases,
                                                     uint32_t* __restrict__ info, uint32_t* __restrict__ tiles,
                                                     uint32_t max_tiles) {
  constexpr uint32_t ROW = kChains * 256;
  __shared__ uint32_t s_prev[257];  // exclusive digit starts of digit p-1, [256] = n
  __shared__ uint32_t s_wsum[2][4];
  __shared__ uint32_t s_T[kChains], s_rb[kChains + 1];
  const uint32_t p = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t* cnt = counts + (size_t)p * ROW;
  uint32_t tot = 0, tot_prev = 0, inc = 0, inc_prev = 0;
  if (tid < 256) {
    for (int c = 0; c < kChains; ++c) tot += cnt[c * 256 + tid];
    if (p > 0) {
      const uint32_t* prev = counts + (size_t)(p - 1) * ROW;
      for (int c = 0; c < kChains; ++c) tot_prev += prev[c * 256 + tid];
    }
    inc = dev::wave_incl_scan(tot, lane);
    inc_prev = dev::wave_incl_scan(tot_prev, lane);
    if (lane == 63) {
      s_wsum[0][wave] = inc;
      s_wsum[1][wave] = inc_prev;
    }
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t pre = 0, pre_prev = 0;
    for (uint32_t w = 0; w < wave; ++w) {
      pre += s_wsum[0][w];
      pre_prev += s_wsum[1][w];
    }
    uint32_t b = pre + inc - tot;  // G_p[tid]
    for (int c = 0; c < kChains; ++c) {
      bases[(size_t)p * ROW + c * 256 + tid] = b;
      b += cnt[c * 256 + tid];
    }
    s_prev[tid] = pre_prev + inc_prev - tot_prev;
    if (tid == 0) s_prev[256] = (uint32_t)n;
  }
  __syncthreads();
  if (tid < kChains) {
    uint32_t first, len;
    if (p == 0) {
      const uint64_t f = (uint64_t)tid << log2s;
      const uint64_t cap = (uint64_t)1 << log2s;
      first = f < n ? (uint32_t)f : n;
      len = (uint64_t)n - first < cap ? n - first : (uint32_t)cap;
    } else {
      first = s_prev[tid << kChainShift];
      len = s_prev[(tid + 1) << kChainShift] - first;
    }
    info[(size_t)p * 36 + 4 * tid] = first;
    info[(size_t)p * 36 + 4 * tid + 1] = len;
    s_T[tid] = (len + tile - 1) / tile;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t r = 0;
    for (int c = 0; c < kChains; ++c) {
      s_rb[c] = r;
      info[(size_t)p * 36 + 4 * c + 2] = r;
      info[(size_t)p * 36 + 4 * c + 3] = 0;
      r += s_T[c];
    }
    s_rb[kChains] = r;
    info[(size_t)p * 36 + 32] = r;
  }
  __syncthreads();
  const uint32_t total = s_rb[kChains];
  uint32_t* order = tiles + (size_t)p * max_tiles;
  for (uint32_t t = tid; t < total; t += blockDim.x) {
    uint32_t c = 0;
    while (t >= s_rb[c + 1]) ++c;
    const uint32_t k = t - s_rb[c];
    const uint64_t Tc = s_T[c];
    uint32_t rank = k;
    for (uint32_t c2 = 0; c2 < (uint32_t)kChains; ++c2) {
      if (c2 == c || s_T[c2] == 0) continue;
      // tiles k2 of chain c2 with (2k2 + 1) * Tc < (2k + 1) * T_c2, plus the
      // tie (2k2 + 1) * Tc == (2k + 1) * T_c2 when c2 < c
      const uint64_t X = (uint64_t)(2 * k + 1) * s_T[c2];
      uint64_t before = ((X + Tc - 1) / Tc) / 2;
      if (c2 < c && X % Tc == 0 && ((X / Tc) & 1)) ++before;
      rank += before < s_T[c2] ? (uint32_t)before : s_T[c2];
    }
    order[rank] = (c << 24) | k;
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
  uint32_t ntiles;  // plain: tiles; chained: launch bound (tiles + kChains)
  uint32_t shift;
  uint32_t fmode;
  const uint32_t* chain_info;
  const uint32_t* chain_tiles;
};

struct PassVariant {
  int block, items, lbw, late, minw, chained;
  int (*launch)(hipStream_t s, const PassArgs& a);
};

template <typename K, bool KV, int BLOCK, int ITEMS, int LBW, bool LATE, int MINW, int MODE, bool UNI, int SPLIT,
          int UL>
int launch_pass(hipStream_t s, const PassArgs& a) {
  // Timed launches carry their events in the dispatch (no extra packets).
  ProfScope ps("k_onesweep", s, true);
  auto kern = k_onesweep<K, KV, BLOCK, ITEMS, LBW, MODE, LATE, MINW, UNI, SPLIT, UL>;
  if (ps.start) {
    hipExtLaunchKernelGGL(kern, dim3(a.ntiles), dim3(BLOCK), 0, s, ps.start, ps.stop, 0, static_cast<const K*>(a.kin),
                          static_cast<K*>(a.kout), a.vin, a.vout, a.n, a.gtab, a.lb, a.ctr, a.ntiles, a.shift, a.fmode,
                          (const uint64_t*)nullptr, (const uint64_t*)nullptr, (const uint64_t*)nullptr,
                          (const uint32_t*)nullptr, 0u, a.chain_info, a.chain_tiles);
    ps.attached = true;
  } else {
    hipLaunchKernelGGL(kern, dim3(a.ntiles), dim3(BLOCK), 0, s, static_cast<const K*>(a.kin), static_cast<K*>(a.kout),
                       a.vin, a.vout, a.n, a.gtab, a.lb, a.ctr, a.ntiles, a.shift, a.fmode, nullptr, nullptr, nullptr,
                       nullptr, 0u, a.chain_info, a.chain_tiles);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// PA_VARIANT(key, payload, block, items, look-back width, late look-back, min waves/SIMD, chained,
//            uniform-digit fast path, staging rounds, lanes per ranking unit)
#define PA_VARIANT(K, KV, B, I, W, P, M, C, U, S, L) \
  { B, I, W, P, M, C, &launch_pass<K, KV, B, I, W, P, M, (C) ? kModeChain : kModePlain, U, S, L> }

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
      PA_VARIANT(uint32_t, false, 512, 48, 8, true, 4, 0, true, 2, 32),  PA_VARIANT(uint32_t, false, 512, 24, 4, true, 6, 0, true, 1, 32),
      PA_VARIANT(uint32_t, false, 512, 16, 4, true, 8, 0, true, 1, 32),  PA_VARIANT(uint32_t, false, 512, 48, 8, true, 4, 1, true, 2, 32),
      PA_VARIANT(uint32_t, false, 512, 48, 16, true, 4, 0, true, 2, 32), PA_VARIANT(uint32_t, false, 256, 48, 8, true, 4, 0, true, 2, 32),
      PA_VARIANT(uint32_t, false, 512, 48, 8, true, 4, 0, true, 2, 16), PA_VARIANT(uint32_t, false, 512, 16, 4, true, 8, 0, true, 1, 16),
      PA_VARIANT(uint32_t, false, 512, 24, 4, true, 6, 0, true, 1, 16), PA_VARIANT(uint32_t, false, 768, 32, 8, true, 6, 0, true, 2, 16)};
  // 512 x 16 up to 2^24 keys (4M: 0.209 -> 0.180 ms against 512 x 48; even at 16M); above, the 24,576-key
  // split tile as 768 x 32 (12 waves per workgroup, two workgroups per CU: 100M passes 188-227 us against
  // 194-233 us for 512 x 48).
  static constexpr int by_class[kClasses] = {2, 2, 7, 9};
};
template <>
struct Variants<uint32_t, true> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint32_t, true, 512, 16, 4, true, 4, 0, true, 1, 32),
                                          PA_VARIANT(uint32_t, true, 512, 16, 4, true, 4, 1, true, 1, 32),
                                          PA_VARIANT(uint32_t, true, 512, 20, 8, true, 4, 0, true, 1, 32)};
  // 512 x 20 above 2^24 keys (100M: 2.157 -> 2.117 ms; at 16M it is 6% slower than 512 x 16).
  static constexpr int by_class[kClasses] = {0, 0, 0, 2};
};
template <>
struct Variants<uint64_t, false> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint64_t, false, 512, 8, 4, true, 8, 0, true, 1, 32),
                                          PA_VARIANT(uint64_t, false, 512, 12, 4, true, 6, 0, true, 1, 32),
                                          PA_VARIANT(uint64_t, false, 512, 8, 4, true, 8, 1, true, 1, 32),
                                          PA_VARIANT(uint64_t, false, 512, 16, 8, true, 4, 0, true, 1, 16),
                                          PA_VARIANT(uint64_t, false, 512, 24, 8, true, 4, 0, true, 2, 16)};
  // 512 x 16 with 16-lane units up to 2^20 keys (1M: 0.155 -> 0.149 ms), the 512 x 24 split with
  // 16-lane units above (4M: 0.331 -> 0.254 ms, 16M: 0.798 -> 0.632 ms, 100M: 4.77 -> 3.57 ms).
  static constexpr int by_class[kClasses] = {0, 3, 4, 4};
};
template <>
struct Variants<uint64_t, true> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint64_t, true, 512, 8, 4, true, 6, 0, true, 1, 32),
                                          PA_VARIANT(uint64_t, true, 512, 8, 4, true, 6, 1, true, 1, 32),
                                          PA_VARIANT(uint64_t, true, 512, 16, 8, true, 4, 0, true, 1, 16)};
  // 512 x 16 with 16-lane units above 2^20 keys (4M: 0.358 -> 0.284 ms, 16M: 1.10 -> 0.96 ms,
  // 100M: 7.01 -> 5.72 ms).
  static constexpr int by_class[kClasses] = {0, 0, 2, 2};
};

// Choices recorded by the autotuner (-1: none) for u32 keys-only sorts.
static std::atomic<int> g_tuned_u32[kClasses] = {{-1}, {-1}, {-1}, {-1}};

template <typename K, bool KV>
constexpr size_t variant_count() {
  return sizeof(Variants<K, KV>::table) / sizeof(PassVariant);
}

// PA_SORT_VARIANT (an index into the key type's table) overrides every
// class; otherwise the tuned choice, otherwise the built-in one.
template <typename K, bool KV>
int variant_index(size_t n) {
  static const int forced = [] {
    const char* e = std::getenv("PA_SORT_VARIANT");
    const int v = e ? std::atoi(e) : -1;
    return (v >= 0 && (size_t)v < variant_count<K, KV>()) ? v : -1;
  }();
  if (forced >= 0) return forced;
  const int c = size_class(n);
  if (std::is_same<K, uint32_t>::value && !KV) {
    const int t = g_tuned_u32[c].load(std::memory_order_relaxed);
    if (t >= 0) return t;
  }
  return Variants<K, KV>::by_class[c];
}

template <typename K, bool KV>
const PassVariant& pick_variant(size_t n) {
  return Variants<K, KV>::table[variant_index<K, KV>(n)];
}

// Look-back rows (and launch bound) of one pass: every chain rounds up to
// whole tiles.
inline size_t sort_rows(const PassVariant& v, size_t n) {
  const size_t tile = (size_t)v.block * v.items;
  return (n + tile - 1) / tile + (v.chained ? kChains : 0);
}

// Pass-0 chains are position ranges of 2^log2s keys (at least 2^16, at most
// kChains of them).
inline uint32_t chain_log2s(size_t n) {
  uint32_t l = 16;
  while ((size_t(1) << l) * kChains < n) ++l;
  return l;
}

constexpr unsigned kHistChainBlocks = 512;
constexpr unsigned kPrefixGroups = 128;  // blocks of the pass-0 prefix scan

// Scratch of one sort call with pass variant var (see enqueue_sort for the
// layout).
template <typename K, bool KV>
size_t sort_scratch_bytes(const PassVariant& var, size_t n) {
  constexpr int D = sizeof(K);
  const size_t rows = sort_rows(var, n);
  Layout L;
  L.take((size_t)D * kChains * 256 * 4);  // counts
  L.take((size_t)D * kChains * 256 * 4);  // bases
  L.take(D * 4);                          // tile counters
  L.take((size_t)D * rows * 256 * 4);     // look-back table
  L.take((size_t)kPrefixGroups * 256 * 4);  // pass-0 prefix partial sums
  if (var.chained) {
    L.take((size_t)D * 36 * 4);                                  // chain tables
    L.take((size_t)D * rows * 4);                                // claim orders
    L.take((size_t)kHistChainBlocks * D * kChains * 256 * 4);    // histogram partials
  }
  return L.off;
}

template <typename K, bool KV>
int enqueue_sort(const PassVariant& var, K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, char* scratch,
                 hipStream_t s, uint32_t xform) {
  constexpr int D = sizeof(K);
  const size_t rows = sort_rows(var, n);
  const size_t ncount = (size_t)D * kChains * 256;
  Layout L;
  uint32_t* counts = reinterpret_cast<uint32_t*>(scratch + L.take(ncount * 4));
  uint32_t* bases = reinterpret_cast<uint32_t*>(scratch + L.take(ncount * 4));
  uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch + L.take(D * 4));
  const size_t lb_bytes = (size_t)D * rows * 256 * 4;
  uint32_t* lb = reinterpret_cast<uint32_t*>(scratch + L.take(lb_bytes));
  uint32_t* prefix_partial = reinterpret_cast<uint32_t*>(scratch + L.take((size_t)kPrefixGroups * 256 * 4));
  uint32_t* chain_info = nullptr;
  uint32_t* chain_tiles = nullptr;
  uint32_t* partial = nullptr;
  if (var.chained) {
    chain_info = reinterpret_cast<uint32_t*>(scratch + L.take((size_t)D * 36 * 4));
    chain_tiles = reinterpret_cast<uint32_t*>(scratch + L.take((size_t)D * rows * 4));
    partial = reinterpret_cast<uint32_t*>(scratch + L.take((size_t)kHistChainBlocks * ncount * 4));
  }

  {
    ProfScope ps("k_sort_prepare", s);
    hipLaunchKernelGGL(k_sort_prepare, dim3(1), dim3(256), 0, s, counts, (uint32_t)ncount, ctr, (uint32_t)D);
  }
  const size_t tile = (size_t)var.block * var.items;
  if (var.chained) {
    const uint32_t log2s = chain_log2s(n);
    {
      ProfScope ps("k_hist_digits", s);
      hipLaunchKernelGGL(k_hist_chains<K>, dim3(kHistChainBlocks), dim3(kSortBlock), 0, s, keys, n, log2s, partial,
                         reinterpret_cast<uint4*>(lb), lb_bytes / 16, xform);
    }
    hipLaunchKernelGGL(k_reduce_counts, dim3((unsigned)((ncount + 255) / 256), 16), dim3(256), 0, s, partial,
                       kHistChainBlocks, (uint32_t)ncount, counts);
    hipLaunchKernelGGL(k_chain_plan, dim3(D), dim3(1024), 0, s, counts, (uint32_t)n, log2s, (uint32_t)tile, bases,
                       chain_info, chain_tiles, (uint32_t)rows);
  } else {
    // One read: all digit histograms, pass 0's per-tile digit counts (into
    // pass 0's look-back rows) and the clearing of the other passes' rows.
    // Two small kernels then turn pass 0's rows into inclusive prefixes, so
    // pass 0 runs without look-back waits; the first also writes the digit
    // bases.
    const size_t pass_rows_bytes = rows * 256 * 4;
    const size_t zero_bytes = lb_bytes - pass_rows_bytes;
    // Tiles per block: about n / 512 keys per block (two blocks per CU), at
    // least one tile and at most 49,152 keys, so a small block's fixed costs
    // (clearing its bins, one global flush) are spread over several tiles.
    // Measured per histogram read: 16M keys in 8K-key tiles 48 -> 31 us,
    // 100M keys in 24K-key tiles 126 -> 122 us, 1M keys unchanged.
    const size_t target = std::min<size_t>(std::max<size_t>(n / 512, tile), 49152);
    const size_t tpb = std::max<size_t>(1, std::min<size_t>((target + tile / 2) / tile, kHistTilesMaxKeys / tile));
    const size_t grid = std::max<size_t>(1, (rows + tpb - 1) / tpb);
    {
      ProfScope ps("k_hist_digits", s);
      hipLaunchKernelGGL(k_hist_tiles<K>, dim3((unsigned)grid), dim3(kSortBlock), 0, s, keys, n, (uint32_t)tile, lb,
                         counts, reinterpret_cast<uint4*>(lb + rows * 256), zero_bytes / 16, xform);
    }
    const uint32_t groups = (uint32_t)std::min<size_t>(rows, kPrefixGroups);
    const uint32_t per = (uint32_t)((rows + groups - 1) / groups);
    ProfScope ps("k_tile_prefix", s);
    hipLaunchKernelGGL(k_tile_sums, dim3(groups + D), dim3(kPrefixThreads), 0, s, lb, (uint32_t)rows, per, prefix_partial,
                       groups, counts, bases);
    hipLaunchKernelGGL(k_tile_scan, dim3(groups), dim3(kPrefixThreads), 0, s, lb, (uint32_t)rows, per, prefix_partial);
  }
  K* kin = keys;
  K* kout = ktmp;
  uint32_t* vin = vals;
  uint32_t* vout = vtmp;
  for (int p = 0; p < D; ++p) {
    uint32_t fmode = 0;
    if (!var.chained && p == 0) fmode |= kPrefixReady;
    if (xform && p == 0) fmode |= kFloatIn;
    if (xform && p == D - 1) fmode |= kFloatOut;
    if (xform == dev::kXformSigned) fmode |= kSignedKeys;
    PassArgs a{kin, kout, vin, vout, (uint32_t)n, bases + (size_t)p * (var.chained ? kChains * 256 : 256),
               lb + (size_t)p * rows * 256, ctr + p, (uint32_t)rows, (uint32_t)(8 * p), fmode,
               var.chained ? chain_info + (size_t)p * 36 : nullptr,
               var.chained ? chain_tiles + (size_t)p * rows : nullptr};
    PA_TRY(var.launch(s, a));
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
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

template <typename K>
size_t large_sort_scratch_bytes(size_t n) {
  const size_t c = large_chunk_keys(n);
  return sort_scratch_bytes<K, false>(pick_variant<K, false>(c), c) + 256 + merge_split_bytes(n);
}

template <typename K>
int enqueue_large_sort(K* keys, K* tmp, size_t n, char* scratch, hipStream_t s, uint32_t xform) {
  const size_t csize = large_chunk_keys(n);
  const PassVariant& var = pick_variant<K, false>(csize);
  const size_t sort_bytes = sort_scratch_bytes<K, false>(var, csize);
  size_t* split = reinterpret_cast<size_t*>(scratch + ((sort_bytes + 255) & ~size_t(255)));
  const unsigned xgrid = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
  if (xform) hipLaunchKernelGGL(k_key_xform<K>, dim3(xgrid), dim3(256), 0, s, keys, n, xform, 0u);
  std::vector<std::pair<size_t, size_t>> runs;  // (first key, length)
  for (size_t off = 0; off < n; off += csize) {
    const size_t len = std::min(csize, n - off);
    // An even number of digit passes leaves each chunk's result in keys.
    PA_TRY((enqueue_sort<K, false>(var, keys + off, nullptr, tmp + off, nullptr, len, scratch, s, dev::kXformNone)));
    runs.emplace_back(off, len);
  }
  K* x = keys;
  K* y = tmp;
  while (runs.size() > 1) {
    std::vector<std::pair<size_t, size_t>> next;
    for (size_t i = 0; i + 1 < runs.size(); i += 2) {
      const auto [o1, l1] = runs[i];
      const auto [o2, l2] = runs[i + 1];
      PA_TRY(merge_enqueue<K>(x + o1, l1, x + o2, l2, y + o1, split, s));
      next.emplace_back(o1, l1 + l2);
    }
    if (runs.size() % 2) {
      const auto [o, l] = runs.back();
      PA_HIP(hipMemcpyAsync(y + o, x + o, l * sizeof(K), hipMemcpyDeviceToDevice, s));
      next.push_back(runs.back());
    }
    runs.swap(next);
    std::swap(x, y);
  }
  if (x != keys) PA_HIP(hipMemcpyAsync(keys, x, n * sizeof(K), hipMemcpyDeviceToDevice, s));
  if (xform) hipLaunchKernelGGL(k_key_xform<K>, dim3(xgrid), dim3(256), 0, s, keys, n, xform, 1u);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// Host entry: classifies pointers, stages host arrays, runs, copies back.
template <typename K, bool KV>
int sort_entry(K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, hipStream_t s, uint32_t xform,
               const char* who) {
  if (n == 0) return PA_OK;
  if (!keys || (KV && !vals)) return arg_error(who);
  const bool large = n > kMaxSortN;
  if (KV && large) {
    set_last_error(std::string(who) + ": key-value sorts take n < 2^30");
    return PA_ERR_TOO_LARGE;
  }
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = is_device_ptr(keys);
  const bool dv = !KV || is_device_ptr(vals);
  const bool dkt = ktmp && is_device_ptr(ktmp);
  const bool dvt = !KV || (vtmp && is_device_ptr(vtmp));
  const PassVariant& var = pick_variant<K, KV>(large ? large_chunk_keys(n) : n);
  Layout L;
  const size_t o_sort = L.take(large ? large_sort_scratch_bytes<K>(n) : sort_scratch_bytes<K, KV>(var, n));
  const size_t o_k = dk ? 0 : L.take(n * sizeof(K));
  const size_t o_kt = dkt ? 0 : L.take(n * sizeof(K));
  const size_t o_v = dv ? 0 : L.take(n * 4);
  const size_t o_vt = dvt ? 0 : L.take(n * 4);
  void* base = nullptr;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
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
      PA_TRY(enqueue_large_sort<K>(dkeys, dktmp, n, b + o_sort, s, xform));
    } else {
      PA_TRY((enqueue_sort<K, KV>(var, dkeys, dvals, dktmp, dvtmp, n, b + o_sort, s, xform)));
    }
  } else {
    PA_TRY((enqueue_sort<K, KV>(var, dkeys, dvals, dktmp, dvtmp, n, b + o_sort, s, xform)));
  }
  if (!dk) PA_TRY(stage_to_host(ctx, keys, dkeys, n * sizeof(K)));
  if (KV && !dv) PA_TRY(stage_to_host(ctx, vals, dvals, n * 4));
  return PA_OK;
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_sort_u32(uint32_t* keys, uint32_t* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, false>(keys, nullptr, tmp, nullptr, n, (hipStream_t)stream, dev::kXformNone, "pa_sort_u32");
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
  if (max_n < 2) return arg_error("pa_autotune_sort_u32: need max_n >= 2");
  hipStream_t s = (hipStream_t)stream;
  static const size_t rep[kClasses] = {size_t(1) << 16, size_t(1) << 20, size_t(1) << 24, size_t(100000000)};
  const size_t nmax = std::min(max_n, std::min(rep[kClasses - 1], kMaxSortN));
  constexpr size_t NV = variant_count<uint32_t, false>();
  size_t scratch_bytes = 0;
  for (size_t v = 0; v < NV; ++v)
    scratch_bytes = std::max(scratch_bytes, sort_scratch_bytes<uint32_t, false>(Variants<uint32_t, false>::table[v], nmax));
  uint32_t *pristine = nullptr, *keys = nullptr, *tmp = nullptr;
  char* scratch = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = PA_OK;
  auto fail = [&](hipError_t e, const char* what) {
    st = hip_fail(e, what, __FILE__, __LINE__);
    return st;
  };
  do {
    hipError_t e;
    if ((e = hipMalloc(&pristine, nmax * 4)) != hipSuccess || (e = hipMalloc(&keys, nmax * 4)) != hipSuccess ||
        (e = hipMalloc(&tmp, nmax * 4)) != hipSuccess || (e = hipMalloc(&scratch, scratch_bytes)) != hipSuccess) {
      fail(e, "hipMalloc");
      break;
    }
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
      fail(e, "hipEventCreate");
      break;
    }
    if ((st = pa_fill_random_u32(pristine, nmax, 0x5eed, stream)) != PA_OK) break;
    size_t prev_n = 0;
    int prev_best = -1;
    for (int c = 0; c < kClasses && st == PA_OK; ++c) {
      const size_t n = std::min(rep[c], nmax);
      if (n == prev_n) {
        g_tuned_u32[c].store(prev_best);
        continue;
      }
      float best_ms = 0.f;
      int best = -1;
      for (size_t v = 0; v < NV && st == PA_OK; ++v) {
        const PassVariant& var = Variants<uint32_t, false>::table[v];
        float total = 0.f;
        for (int rep_i = 0; rep_i < 4 && st == PA_OK; ++rep_i) {  // first run warms up, three are timed
          if ((e = hipMemcpyAsync(keys, pristine, n * 4, hipMemcpyDeviceToDevice, s)) != hipSuccess) {
            fail(e, "hipMemcpyAsync");
            break;
          }
          (void)hipEventRecord(e0, s);
          if ((st = enqueue_sort<uint32_t, false>(var, keys, nullptr, tmp, nullptr, n, scratch, s, dev::kXformNone)) != PA_OK) break;
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
        g_tuned_u32[c].store(best);
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

int pa_sort_variant_u32(size_t n) { return variant_index<uint32_t, false>(n); }

int pa_autotune_reset(void) {
  for (int c = 0; c < kClasses; ++c) g_tuned_u32[c].store(-1);
  return PA_OK;
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
