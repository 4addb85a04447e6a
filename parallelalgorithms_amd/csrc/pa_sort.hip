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

#include "pa_internal.h"

namespace pa {

constexpr int kSortBlock = 512;
constexpr int kSortWaves = kSortBlock / 64;

// Look-back status word: 2 flag bits + 30-bit count.
constexpr uint32_t kLbAgg = 1u << 30;
constexpr uint32_t kLbInc = 2u << 30;
constexpr uint32_t kLbMask = (1u << 30) - 1;
constexpr size_t kMaxSortN = (size_t(1) << 30) - 1;

enum : uint32_t { kFloatIn = 1u, kFloatOut = 2u };

template <typename K>
struct SortTile;
template <>
struct SortTile<uint32_t> {
  static constexpr int items = 16;
};
template <>
struct SortTile<uint64_t> {
  static constexpr int items = 8;
};

__global__ __launch_bounds__(256) void k_sort_prepare(uint32_t* __restrict__ counts, uint32_t ncounts,
                                                      uint32_t* __restrict__ counters, uint32_t ncounters) {
  for (uint32_t i = threadIdx.x; i < ncounts; i += blockDim.x) counts[i] = 0;
  for (uint32_t i = threadIdx.x; i < ncounters; i += blockDim.x) counters[i] = 0;
}

// Adds the digit contributions of PER keys held by one lane to the LDS bins of
// digit row `row`.  When every active lane of the wave holds PER keys of the
// same digit (constant input, long presorted runs), one lane adds the whole
// wave's count instead of 64 lanes hitting one LDS address.
template <int PER>
__device__ __forceinline__ void hist_add(uint32_t* __restrict__ row, const uint32_t (&dg)[PER]) {
  bool same = true;
#pragma unroll
  for (int j = 1; j < PER; ++j) same = same && (dg[j] == dg[0]);
  const uint32_t first = __builtin_amdgcn_readfirstlane(dg[0]);
  const uint64_t active = __ballot(true);
  const uint64_t uni = __ballot(same && dg[0] == first);
  if (uni == active) {
    if (dev::count_below(active) == 0) atomicAdd(&row[first], (uint32_t)(PER * __popcll(active)));
  } else if (same) {
    atomicAdd(&row[dg[0]], (uint32_t)PER);
  } else {
#pragma unroll
    for (int j = 0; j < PER; ++j) atomicAdd(&row[dg[j]], 1u);
  }
}

// One read of the keys -> D digit histograms (D = sizeof(K)).
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_hist_digits(const K* __restrict__ keys, size_t n,
                                                            uint32_t* __restrict__ counts, uint4* __restrict__ zero_ptr,
                                                            size_t zero_vec, uint32_t float_keys) {
  constexpr int D = sizeof(K);
  constexpr int PER = 16 / sizeof(K);
  __shared__ uint32_t s_hist[D * 256];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < D * 256; i += kSortBlock) s_hist[i] = 0;

  // Clear the look-back table used by the following onesweep passes.
  const size_t gstride = (size_t)gridDim.x * kSortBlock;
  for (size_t i = (size_t)blockIdx.x * kSortBlock + tid; i < zero_vec; i += gstride) zero_ptr[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();

  size_t head = ((16 - ((uintptr_t)keys & 15)) & 15) / sizeof(K);
  if (head > n) head = n;
  const size_t nvec = (n - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(keys + head);
  for (size_t v = (size_t)blockIdx.x * kSortBlock + tid; v < nvec; v += gstride) {
    uint4 x = vk[v];
    K k[PER];
    __builtin_memcpy(k, &x, 16);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      uint32_t dg[PER];
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        K kk = k[j];
        if (sizeof(K) == 4 && float_keys) kk = (K)dev::f32_fwd((uint32_t)kk);
        dg[j] = dev::digit8<K>(kk, 8 * d);
      }
      hist_add<PER>(s_hist + d * 256, dg);
    }
  }
  // Head and tail elements that do not fill a 16-byte vector.
  if (blockIdx.x == 0) {
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
    if (c) atomicAdd(&counts[i], c);
  }
}

// One stable digit pass.  Tiles are claimed in order from a global counter so
// every tile a block waits on in the look-back is already held by a running
// block.
template <typename K, bool KV, int ITEMS>
__global__ __launch_bounds__(kSortBlock) void k_onesweep(const K* __restrict__ kin, K* __restrict__ kout,
                                                         const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                         uint32_t n, const uint32_t* __restrict__ gcount,
                                                         uint32_t* __restrict__ lb, uint32_t* __restrict__ tile_counter,
                                                         uint32_t shift, uint32_t fmode) {
  constexpr int TILE = kSortBlock * ITEMS;
  constexpr int WAVE_SPAN = 64 * ITEMS;
  __shared__ K s_keys[TILE];
  __shared__ uint32_t s_whist[kSortWaves * 256];
  __shared__ uint32_t s_goff[256];
  __shared__ uint32_t s_wsum[2][4];
  __shared__ uint32_t s_tile;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;

  if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (uint32_t i = tid; i < kSortWaves * 256; i += kSortBlock) s_whist[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t tile_base = tile * (uint32_t)TILE;
  const uint32_t valid = min((uint32_t)TILE, n - tile_base);
  const bool full = valid == (uint32_t)TILE;

  // Wave-striped load: wave w owns [tile_base + w*WAVE_SPAN, +WAVE_SPAN); item
  // i of lane l is element i*64 + l of that span, so (item, lane) order is the
  // input order and every load instruction reads 64 consecutive keys.
  const uint32_t wbase = tile_base + wave * WAVE_SPAN + lane;
  K key[ITEMS];
  uint32_t val[KV ? ITEMS : 1];
  if (full) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) key[i] = kin[wbase + i * 64];
    if constexpr (KV) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) val[i] = vin[wbase + i * 64];
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

  // Stable rank within the wave: lanes holding the same digit are found with
  // 8 ballots; the lowest such lane bumps the wave's private counter.
  uint32_t rank[ITEMS];
  uint32_t* wh = s_whist + wave * 256;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t d = dev::digit8<K>(key[i], shift);
    const bool ok = full || (wbase + i * 64 < n);
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t below = dev::count_below(peers);
    const uint32_t base = wh[d];
    rank[i] = base + below;
    if (ok && below == 0) wh[d] = base + (uint32_t)__popcll(peers);
  }
  __syncthreads();

  // Per digit: exclusive prefix over waves, tile total, publish the aggregate.
  uint32_t cnt = 0, gval = 0;
  if (tid < 256) {
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
      uint32_t c = s_whist[w * 256 + tid];
      s_whist[w * 256 + tid] = cnt;
      cnt += c;
    }
    dev::atomic_store_agent(&lb[(size_t)tile * 256 + tid], (tile == 0 ? kLbInc : kLbAgg) | cnt);
    gval = gcount[tid];
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
  __syncthreads();
  if (tid < 256) {
    uint32_t cpre = 0, gpre = 0;
    for (uint32_t w = 0; w < wave; ++w) {
      cpre += s_wsum[0][w];
      gpre += s_wsum[1][w];
    }
    const uint32_t tstart = cpre + cinc - cnt;   // tile-local start of this digit
    const uint32_t gstart = gpre + ginc - gval;  // global start of this digit

    // Decoupled look-back over earlier tiles for this digit.
    uint32_t excl = 0;
    if (tile > 0) {
      int64_t p = (int64_t)tile - 1;
      while (p >= 0) {
        const uint32_t w = dev::atomic_load_agent(&lb[(size_t)p * 256 + tid]);
        const uint32_t flag = w & ~kLbMask;
        if (flag == 0) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += w & kLbMask;
        if (flag == kLbInc) break;
        --p;
      }
      dev::atomic_store_agent(&lb[(size_t)tile * 256 + tid], kLbInc | (excl + cnt));
    }
    s_goff[tid] = gstart + excl - tstart;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) s_whist[w * 256 + tid] += tstart;
  }
  __syncthreads();

  // Stage the tile in LDS in digit order.
  uint32_t lpos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t d = dev::digit8<K>(key[i], shift);
    lpos[i] = wh[d] + rank[i];
    if (full || (wbase + i * 64 < n)) s_keys[lpos[i]] = key[i];
  }
  __syncthreads();

  // Coalesced write-out: consecutive slots of one digit go to consecutive
  // addresses of that digit's bin.
  uint32_t gpos[KV ? ITEMS : 1];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t j = tid + i * kSortBlock;
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
      const uint32_t j = tid + i * kSortBlock;
      if (j < valid) vout[gpos[i]] = s_vals[j];
    }
  }
}

// Enqueues the full sort on device buffers.  scratch must hold
// sort_scratch_bytes<K>(n) bytes.
template <typename K>
size_t sort_scratch_bytes(size_t n) {
  constexpr int D = sizeof(K);
  const size_t tile = (size_t)kSortBlock * SortTile<K>::items;
  const size_t tiles = (n + tile - 1) / tile;
  Layout L;
  L.take(D * 256 * 4);
  L.take(D * 4);
  L.take((size_t)D * tiles * 256 * 4);
  return L.off;
}

template <typename K, bool KV>
int enqueue_sort(K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, char* scratch, hipStream_t s,
                 bool float_keys) {
  constexpr int D = sizeof(K);
  constexpr int ITEMS = SortTile<K>::items;
  const size_t tile = (size_t)kSortBlock * ITEMS;
  const size_t tiles = (n + tile - 1) / tile;
  Layout L;
  uint32_t* counts = reinterpret_cast<uint32_t*>(scratch + L.take(D * 256 * 4));
  uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch + L.take(D * 4));
  const size_t lb_bytes = (size_t)D * tiles * 256 * 4;
  uint32_t* lb = reinterpret_cast<uint32_t*>(scratch + L.take(lb_bytes));

  {
    ProfScope ps("k_sort_prepare", s);
    hipLaunchKernelGGL(k_sort_prepare, dim3(1), dim3(256), 0, s, counts, (uint32_t)(D * 256), ctr, (uint32_t)D);
  }
  {
    const size_t nvec = n / (16 / sizeof(K)) + 1;
    unsigned grid = (unsigned)std::min<size_t>((nvec + kSortBlock - 1) / kSortBlock, 2048);
    grid = std::max<unsigned>(grid, (unsigned)std::min<size_t>(lb_bytes / 16 / kSortBlock + 1, 2048));
    ProfScope ps("k_hist_digits", s);
    hipLaunchKernelGGL(k_hist_digits<K>, dim3(grid), dim3(kSortBlock), 0, s, keys, n, counts,
                       reinterpret_cast<uint4*>(lb), lb_bytes / 16, (uint32_t)float_keys);
  }
  K* kin = keys;
  K* kout = ktmp;
  uint32_t* vin = vals;
  uint32_t* vout = vtmp;
  for (int p = 0; p < D; ++p) {
    uint32_t fmode = 0;
    if (float_keys && p == 0) fmode |= kFloatIn;
    if (float_keys && p == D - 1) fmode |= kFloatOut;
    ProfScope ps("k_onesweep", s);
    hipLaunchKernelGGL((k_onesweep<K, KV, ITEMS>), dim3((unsigned)tiles), dim3(kSortBlock), 0, s, kin, kout, vin, vout,
                       (uint32_t)n, counts + p * 256, lb + (size_t)p * tiles * 256, ctr + p, (uint32_t)(8 * p), fmode);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}

// Host entry: classifies pointers, stages host arrays, runs, copies back.
template <typename K, bool KV>
int sort_entry(K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, hipStream_t s, bool float_keys,
               const char* who) {
  if (n == 0) return PA_OK;
  if (!keys || (KV && !vals)) return arg_error(who);
  if (n > kMaxSortN) {
    set_last_error(std::string(who) + ": n exceeds 2^30-1");
    return PA_ERR_TOO_LARGE;
  }
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = is_device_ptr(keys);
  const bool dv = !KV || is_device_ptr(vals);
  const bool dkt = ktmp && is_device_ptr(ktmp);
  const bool dvt = !KV || (vtmp && is_device_ptr(vtmp));
  Layout L;
  const size_t o_sort = L.take(sort_scratch_bytes<K>(n));
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
  if (!dk) PA_HIP(hipMemcpyAsync(dkeys, keys, n * sizeof(K), hipMemcpyHostToDevice, s));
  if (KV && !dv) PA_HIP(hipMemcpyAsync(dvals, vals, n * 4, hipMemcpyHostToDevice, s));
  PA_TRY((enqueue_sort<K, KV>(dkeys, dvals, dktmp, dvtmp, n, b + o_sort, s, float_keys)));
  if (!dk) PA_HIP(hipMemcpyAsync(keys, dkeys, n * sizeof(K), hipMemcpyDeviceToHost, s));
  if (KV && !dv) PA_HIP(hipMemcpyAsync(vals, dvals, n * 4, hipMemcpyDeviceToHost, s));
  if (!dk || (KV && !dv)) PA_HIP(hipStreamSynchronize(s));
  return PA_OK;
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_sort_u32(uint32_t* keys, uint32_t* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, false>(keys, nullptr, tmp, nullptr, n, (hipStream_t)stream, false, "pa_sort_u32");
}

int pa_sort_u64(uint64_t* keys, uint64_t* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint64_t, false>(keys, nullptr, tmp, nullptr, n, (hipStream_t)stream, false, "pa_sort_u64");
}

int pa_sort_f32(float* keys, float* tmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, false>(reinterpret_cast<uint32_t*>(keys), nullptr, reinterpret_cast<uint32_t*>(tmp),
                                     nullptr, n, (hipStream_t)stream, true, "pa_sort_f32");
}

int pa_sort_kv_u32_u32(uint32_t* keys, uint32_t* vals, uint32_t* ktmp, uint32_t* vtmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint32_t, true>(keys, vals, ktmp, vtmp, n, (hipStream_t)stream, false, "pa_sort_kv_u32_u32");
}

int pa_sort_kv_u64_u32(uint64_t* keys, uint32_t* vals, uint64_t* ktmp, uint32_t* vtmp, size_t n, pa_stream_t stream) {
  return sort_entry<uint64_t, true>(keys, vals, ktmp, vtmp, n, (hipStream_t)stream, false, "pa_sort_kv_u64_u32");
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
