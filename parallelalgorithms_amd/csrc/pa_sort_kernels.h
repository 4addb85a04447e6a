// pa_sort_kernels.h -- device kernels of the LSD radix sort, shared by the
// plain sort (pa_sort.hip) and the segmented batch sort (pa_segsort.hip).
#pragma once

#include "pa_internal.h"

namespace pa {

constexpr int kSortBlock = 512;  // histogram kernel block

// Look-back status word: 2 flag bits + 30-bit count.
constexpr uint32_t kLbAgg = 1u << 30;
constexpr uint32_t kLbInc = 2u << 30;
constexpr uint32_t kLbMask = (1u << 30) - 1;
constexpr size_t kMaxSortN = (size_t(1) << 30) - 1;

enum : uint32_t { kFloatIn = 1u, kFloatOut = 2u };


// One read of the keys -> D digit histograms (D = sizeof(K)).
template <typename K>
// With nchains > 1 the keys are split into nchains contiguous chains of
// chain_keys keys (the last may be shorter); block b counts part of chain
// b % nchains, and counts are laid out [D][nchains][256].  gridDim.x must be
// a multiple of nchains.
__global__ __launch_bounds__(kSortBlock) void k_hist_digits(const K* __restrict__ keys, size_t n,
                                                            uint32_t* __restrict__ counts, uint4* __restrict__ zero_ptr,
                                                            size_t zero_vec, uint32_t float_keys,
                                                            uint32_t nchains = 1, size_t chain_keys = 0) {
  constexpr int D = sizeof(K);
  constexpr int PER = 16 / sizeof(K);
  __shared__ uint32_t s_hist[D * 256];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < D * 256; i += kSortBlock) s_hist[i] = 0;

  // Clear the look-back table used by the following onesweep passes.
  const size_t gstride = (size_t)gridDim.x * kSortBlock;
  for (size_t i = (size_t)blockIdx.x * kSortBlock + tid; i < zero_vec; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();

  // This block's chain and its share of it.
  const uint32_t chain = blockIdx.x % nchains;
  const uint32_t part = blockIdx.x / nchains, parts = gridDim.x / nchains;
  if (nchains > 1) {
    const size_t lo = (size_t)chain * chain_keys;
    keys += lo;
    n = lo < n ? std::min(chain_keys, n - lo) : 0;
  }
  const size_t pstride = (size_t)parts * kSortBlock;
  size_t head = ((16 - ((uintptr_t)keys & 15)) & 15) / sizeof(K);
  if (head > n) head = n;
  const size_t nvec = (n - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(keys + head);
  // U 16-byte loads per thread are issued before any counting, so each wave
  // keeps U KiB in flight.
  constexpr int U = 4;
  const size_t v0 = (size_t)part * kSortBlock + tid;
  for (size_t base = 0; base < nvec; base += pstride * U) {
    uint4 x[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t v = base + u * pstride + v0;
      ok[u] = v < nvec;
      if (ok[u]) x[u] = vk[v];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      K k[PER];
      __builtin_memcpy(k, &x[u], 16);
      if (sizeof(K) == 4 && float_keys) {
#pragma unroll
        for (int j = 0; j < PER; ++j) k[j] = (K)dev::f32_fwd((uint32_t)k[j]);
      }
      // Whole-key fast path: every active lane holds PER copies of one key
      // (constant input) -> one lane adds the wave's count to each digit.
      bool same = true;
#pragma unroll
      for (int j = 1; j < PER; ++j) same = same && (k[j] == k[0]);
      const uint32_t f_lo = __builtin_amdgcn_readfirstlane((uint32_t)k[0]);
      const uint32_t f_hi = sizeof(K) == 8 ? __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)k[0] >> 32)) : 0u;
      const K first = sizeof(K) == 8 ? (K)(((uint64_t)f_hi << 32) | f_lo) : (K)f_lo;
      const uint64_t active = __ballot(true);
      if (__ballot(same && k[0] == first) == active) {
        if (dev::count_below(active) == 0) {
          const uint32_t c = (uint32_t)(PER * __popcll(active));
#pragma unroll
          for (int d = 0; d < D; ++d) atomicAdd(&s_hist[d * 256 + dev::digit8<K>(first, 8 * d)], c);
        }
        continue;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) {
        uint32_t dg[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) dg[j] = dev::digit8<K>(k[j], 8 * d);
        dev::lds_count<PER>(s_hist + d * 256, dg);
      }
    }
  }
  // Head and tail elements that do not fill a 16-byte vector.
  if (part == 0) {
    const size_t tail_start = head + nvec * PER;
    const size_t extra = head + (n - tail_start);
    for (size_t e = tid; e < extra; e += kSortBlock) {
      size_t idx = e < head ? e : tail_start + (e - head);
      K kk = keys[idx];
      if (sizeof(K) == 4 && float_keys) kk = (K)dev::f32_fwd((uint32_t)kk);
#pragma unroll
      for (int d = 0; d < D; ++d) atomicAdd(&s_hist[d * 256 + dev::digit8<K>(kk, 8 * d)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < D * 256; i += kSortBlock) {
    uint32_t c = s_hist[i];
    if (c) atomicAdd(&counts[((i >> 8) * nchains + chain) * 256 + (i & 255)], c);
  }
}


// One stable digit pass over a tile of BLOCK x ITEMS keys.  Tiles are claimed
// in order from a global counter so every tile a block waits on in the
// look-back is already held by a running block.  The look-back reads LBW
// predecessor status words per round trip.
//
// Chain mode (!SEG): the keys are split into nchains contiguous chains of
// tpc tiles, each with its own look-back chain; claimed index i is tile i /
// nchains of chain i % nchains, so all chains advance together.  gtab holds
// this pass's base[chain][digit] output offsets.  Independent chains keep the
// look-back short: each chain sees 1/nchains of the tile throughput.
//
// Segmented mode (SEG): many independent segments, each with its own input
// and output address (seg_in / seg_out, device addresses as u64) and length
// (seg_len); kin/kout are unused.  tile_info[t] = (segment << 32) | tile index
// within the segment, gtab holds per-segment digit counts strided by
// 256 * gstride per segment, and
// the look-back stops at the segment's first tile.  Tiles never span segments.
// MINW: minimum waves per SIMD the register allocation must allow (1 = no
// constraint); 6 keeps a 512-thread block at 3 resident blocks per CU.
// COOP > 0: before the per-digit look-back, the whole block loads the status
// words of the COOP preceding tiles (all 256 digits) into LDS in one round
// trip; each digit then walks that window in LDS and polls memory only if the
// window holds no inclusive prefix yet.
template <typename K, bool KV, int BLOCK, int ITEMS, int LBW, bool SEG = false, int RANK = 1, bool EARLY = false,
          int MINW = 1, int COOP = 0>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MINW))) void k_onesweep(const K* __restrict__ kin, K* __restrict__ kout,
                                                    const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                    uint32_t n, const uint32_t* __restrict__ gtab,
                                                    uint32_t* __restrict__ lb, uint32_t* __restrict__ tile_counter,
                                                    uint32_t shift, uint32_t fmode, uint32_t nchains, uint32_t tpc,
                                                    const uint64_t* __restrict__ tile_info = nullptr,
                                                    const uint64_t* __restrict__ seg_in = nullptr,
                                                    const uint64_t* __restrict__ seg_out = nullptr,
                                                    const uint32_t* __restrict__ seg_len = nullptr,
                                                    uint32_t gstride = 0) {
  static_assert(BLOCK >= 256 && BLOCK % 64 == 0, "one thread per digit is needed");
  constexpr int WAVES = BLOCK / 64;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int WAVE_SPAN = 64 * ITEMS;
  __shared__ __align__(16) K s_keys[TILE];
  __shared__ uint32_t s_whist[WAVES * 256];
  __shared__ uint32_t s_goff[256];
  __shared__ uint32_t s_cnt[EARLY ? 256 : 1];
  __shared__ uint32_t s_wsum[2][4];
  __shared__ uint32_t s_tile;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;

  if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (uint32_t i = tid; i < WAVES * 256; i += BLOCK) s_whist[i] = 0;
  if constexpr (EARLY) {
    if (tid < 256) s_cnt[tid] = 0;
  }
  if constexpr (RANK != 0) {
    static_assert(TILE * sizeof(K) >= WAVES * 256 * sizeof(uint64_t), "match slots must fit the key staging area");
    uint64_t* m = reinterpret_cast<uint64_t*>(s_keys);
    for (uint32_t i = tid; i < WAVES * 256; i += BLOCK) m[i] = 0ull;
  }
  __syncthreads();
  uint32_t tile = s_tile;
  uint32_t local;  // tile index within its chain or segment
  uint32_t chain = 0;
  if constexpr (!SEG) {
    chain = tile % nchains;
    local = tile / nchains;
    tile = chain * tpc + local;  // position of the tile in the key array
    if (tile * (uint32_t)TILE >= n) return;  // past the end of a short last chain
    gtab += chain * 256;
  } else {
    const uint64_t info = tile_info[tile];
    const uint32_t seg = (uint32_t)(info >> 32);
    local = (uint32_t)info;
    static_assert(!(SEG && KV), "segmented mode sorts keys only");
    kin = reinterpret_cast<const K*>(seg_in[seg]);
    kout = reinterpret_cast<K*>(seg_out[seg]);
    n = seg_len[seg];
    gtab += (size_t)seg * gstride * 256;
  }
  const uint32_t first_tile = tile - local;  // look-back stops here
  const uint32_t tile_base = (SEG ? local : tile) * (uint32_t)TILE;
  const uint32_t valid = min((uint32_t)TILE, n - tile_base);
  const bool full = valid == (uint32_t)TILE;

  // Wave-striped load: wave w owns [tile_base + w*WAVE_SPAN, +WAVE_SPAN); item
  // i of lane l is element i*64 + l of that span, so (item, lane) order is the
  // input order and every load instruction reads 64 consecutive keys.
  const uint32_t wbase = tile_base + wave * WAVE_SPAN + lane;
  K key[ITEMS];
  uint32_t val[KV ? ITEMS : 1];
  if (full) {
    // One base address per lane; the item offsets are compile-time constants.
    const K* src = kin + wbase;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) key[i] = src[i * 64];
    if constexpr (KV) {
      const uint32_t* vsrc = vin + wbase;
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) val[i] = vsrc[i * 64];
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint32_t idx = wbase + i * 64;
      key[i] = idx < n ? kin[idx] : K(0);
    }
    if constexpr (KV) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        uint32_t idx = wbase + i * 64;
        val[i] = idx < n ? vin[idx] : 0u;
      }
    }
  }
  if (sizeof(K) == 4 && (fmode & kFloatIn)) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) key[i] = (K)dev::f32_fwd((uint32_t)key[i]);
  }

  // EARLY: count the tile's digits first, publish the aggregate and put the
  // predecessor's status word in flight, so the look-back latency overlaps
  // the ranking instead of following it.
  uint32_t cnt = 0;  // this thread's digit count for the tile (threads < 256)
  uint32_t prefetched = 0;
  if constexpr (EARLY) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (full || (wbase + i * 64 < n)) {
        const uint32_t dg[1] = {dev::digit8<K>(key[i], shift)};
        dev::lds_count<1>(s_cnt, dg);
      }
    }
    __syncthreads();
    if (tid < 256) {
      cnt = s_cnt[tid];
      dev::atomic_store_agent(&lb[(size_t)tile * 256 + tid], (local == 0 ? kLbInc : kLbAgg) | cnt);
      if (local > 0) prefetched = dev::atomic_load_agent(&lb[(size_t)(tile - 1) * 256 + tid]);
    }
  }

  // Stable rank within the wave.  The lanes holding the same digit (the
  // "peers") are found either by 8 ballots (RANK 0) or by every lane OR-ing
  // its lane bit into the wave's 64-bit LDS slot for its digit and reading the
  // slot back (RANK 1: 3 LDS operations instead of ~70 VALU instructions; OR
  // commutes, so the result does not depend on the order the LDS applies the
  // lanes).  The lowest peer bumps the wave's private digit counter and clears
  // the slot.  LDS operations of one wave execute in program order.
  uint32_t rank[ITEMS];
  uint32_t* wh = s_whist + wave * 256;
  // The match slots alias the key staging area, which is not used until the
  // ranking is over.
  uint64_t* wm = reinterpret_cast<uint64_t*>(s_keys) + wave * 256;
#pragma unroll 4
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t d = dev::digit8<K>(key[i], shift);
    const bool ok = full || (wbase + i * 64 < n);
    uint64_t peers;
    if constexpr (RANK == 0) {
      peers = __ballot(ok);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
      }
    } else {
      if (ok) atomicOr(reinterpret_cast<unsigned long long*>(&wm[d]), 1ull << lane);
      peers = ok ? wm[d] : 0ull;
    }
    const uint32_t below = dev::count_below(peers);
    const uint32_t base = wh[d];
    rank[i] = base + below;
    if (ok && below == 0) {
      wh[d] = base + (uint32_t)__popcll(peers);
      if constexpr (RANK != 0) wm[d] = 0ull;
    }
  }
  __syncthreads();

  // Per digit: exclusive prefix over waves, tile total, publish the aggregate
  // (unless EARLY already did).
  uint32_t gval = 0;
  if (tid < 256) {
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
      uint32_t c = s_whist[w * 256 + tid];
      s_whist[w * 256 + tid] = tot;
      tot += c;
    }
    if constexpr (!EARLY) {
      cnt = tot;
      dev::atomic_store_agent(&lb[(size_t)tile * 256 + tid], (local == 0 ? kLbInc : kLbAgg) | cnt);
    }
    if constexpr (SEG) gval = gtab[tid];
  }
  // Exclusive scans over the 256 digits of (tile count, global count).
  uint32_t cinc = 0, ginc = 0;
  if (tid < 256) {
    cinc = dev::wave_incl_scan(cnt, lane);
    ginc = dev::wave_incl_scan(gval, lane);
    if (lane == 63) {
      s_wsum[0][wave] = cinc;
      s_wsum[1][wave] = ginc;
    }
  }
  // Cooperative look-back window: status words of tiles tile-1 .. tile-COOP
  // for every digit, win[j][d] (the key staging area is idle until the
  // scatter below).  Words before the chain's first tile read as "inclusive 0".
  uint32_t* const win = reinterpret_cast<uint32_t*>(s_keys);
  if constexpr (COOP > 0) {
    static_assert(COOP * 256 * 4 <= TILE * (int)sizeof(K), "window must fit the key staging area");
    static_assert((COOP * 256) % BLOCK == 0, "whole window rows per thread");
    if (local > 0) {
      uint32_t w[COOP * 256 / BLOCK];
#pragma unroll
      for (int r = 0; r < COOP * 256 / BLOCK; ++r) {
        const uint32_t q = tid + r * BLOCK;
        const int64_t pt = (int64_t)tile - 1 - (int64_t)(q >> 8);
        w[r] = pt >= (int64_t)first_tile ? dev::atomic_load_agent(&lb[(size_t)pt * 256 + (q & 255)]) : kLbInc;
      }
#pragma unroll
      for (int r = 0; r < COOP * 256 / BLOCK; ++r) win[tid + r * BLOCK] = w[r];
    }
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t cpre = 0, gpre = 0;
    for (uint32_t w = 0; w < wave; ++w) {
      cpre += s_wsum[0][w];
      gpre += s_wsum[1][w];
    }
    const uint32_t tstart = cpre + cinc - cnt;  // tile-local start of this digit
    // First output slot of this digit for the chain (or segment).
    const uint32_t gstart = SEG ? gpre + ginc - gval : gtab[tid];

    // Decoupled look-back over earlier tiles of the segment for this digit, LBW
    // status words per round trip; words before the first tile read as
    // "inclusive 0".
    uint32_t excl = 0;
    if (local > 0) {
      int64_t p = (int64_t)tile - 1;
      bool resolved = false;
      if constexpr (COOP > 0) {
        for (int j = 0; j < COOP && !resolved; ++j) {
          const uint32_t wv = win[j * 256 + tid];
          const uint32_t flag = wv & ~kLbMask;
          if (flag == 0) break;  // not published when the window was read
          excl += wv & kLbMask;
          resolved = flag == kLbInc;
          --p;
        }
      } else if constexpr (EARLY) {
        const uint32_t flag = prefetched & ~kLbMask;
        if (flag != 0) {
          excl += prefetched & kLbMask;
          resolved = flag == kLbInc;
          --p;
        }
      }
      while (!resolved) {
        uint32_t w[LBW];
#pragma unroll
        for (int j = 0; j < LBW; ++j)
          w[j] = (p - j >= (int64_t)first_tile) ? dev::atomic_load_agent(&lb[(size_t)(p - j) * 256 + tid]) : kLbInc;
        int used = 0;
        bool stop = false, done = false;
#pragma unroll
        for (int j = 0; j < LBW; ++j) {
          if (!stop) {
            const uint32_t flag = w[j] & ~kLbMask;
            if (flag == 0) {
              stop = true;  // not published yet: poll again from here
            } else {
              excl += w[j] & kLbMask;
              ++used;
              if (flag == kLbInc) stop = done = true;
            }
          }
        }
        if (done) break;
        p -= used;
        if (used == 0) __builtin_amdgcn_s_sleep(1);
      }
      dev::atomic_store_agent(&lb[(size_t)tile * 256 + tid], kLbInc | (excl + cnt));
    }
    s_goff[tid] = gstart + excl - tstart;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) s_whist[w * 256 + tid] += tstart;
  }
  __syncthreads();

  // Stage the tile in LDS in digit order.
  uint32_t lpos[ITEMS];
#pragma unroll 4
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t d = dev::digit8<K>(key[i], shift);
    lpos[i] = wh[d] + rank[i];
    if (full || (wbase + i * 64 < n)) s_keys[lpos[i]] = key[i];
  }
  __syncthreads();

  // Coalesced write-out: consecutive slots of one digit go to consecutive
  // addresses of that digit's bin.
  uint32_t gpos[KV ? ITEMS : 1];
#pragma unroll 4
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t j = tid + i * BLOCK;
    if (j < valid) {
      K k = s_keys[j];
      const uint32_t p = s_goff[dev::digit8<K>(k, shift)] + j;
      if (sizeof(K) == 4 && (fmode & kFloatOut)) k = (K)dev::f32_inv((uint32_t)k);
      kout[p] = k;
      if constexpr (KV) gpos[i] = p;
    }
  }
  if constexpr (KV) {
    __syncthreads();
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (full || (wbase + i * 64 < n)) s_vals[lpos[i]] = val[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t j = tid + i * BLOCK;
      if (j < valid) vout[gpos[i]] = s_vals[j];
    }
  }
}

}  // namespace pa
