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
#include <cstdlib>

#include "pa_sort_kernels.h"

namespace pa {

__global__ __launch_bounds__(256) void k_sort_prepare(uint32_t* __restrict__ counts, uint32_t ncounts,
                                                      uint32_t* __restrict__ counters, uint32_t ncounters) {
  for (uint32_t i = threadIdx.x; i < ncounts; i += blockDim.x) counts[i] = 0;
  for (uint32_t i = threadIdx.x; i < ncounters; i += blockDim.x) counters[i] = 0;
}

// Per pass p (one block each): base[p][d] = sum_{d' < d} counts[p][d'], the
// first output slot of digit d.
__global__ __launch_bounds__(256) void k_digit_bases(const uint32_t* __restrict__ counts,
                                                     uint32_t* __restrict__ base) {
  __shared__ uint32_t s_wsum[4];
  const uint32_t d = threadIdx.x, lane = d & 63, wave = d >> 6;
  const uint32_t c = counts[blockIdx.x * 256 + d];
  const uint32_t inc = dev::wave_incl_scan(c, lane);
  if (lane == 63) s_wsum[wave] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (uint32_t w = 0; w < wave; ++w) pre += s_wsum[w];
  base[blockIdx.x * 256 + d] = pre + inc - c;
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
  int block, items, lbw, pipe, minw;
  int (*launch)(hipStream_t s, const PassArgs& a);
};

// Resident blocks of a kernel on the current device (persistent launches).
template <typename F>
unsigned resident_blocks(F kernel, int block) {
  int dev = 0, cus = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
  (void)hipGetLastError();
  return (unsigned)std::max(1, cus * std::max(1, per_cu));
}

template <typename K, bool KV, int BLOCK, int ITEMS, int LBW, bool PIPE, int MINW>
int launch_pass(hipStream_t s, const PassArgs& a) {
  auto kern = k_onesweep<K, KV, BLOCK, ITEMS, LBW, false, PIPE, MINW>;
  unsigned grid = a.ntiles;
  if (PIPE) {
    static const unsigned resident = resident_blocks(kern, BLOCK);  // one device per process is assumed
    grid = std::min(grid, resident);
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, s, static_cast<const K*>(a.kin), static_cast<K*>(a.kout), a.vin,
                     a.vout, a.n, a.gtab, a.lb, a.ctr, a.ntiles, a.shift, a.fmode, nullptr, nullptr, nullptr, nullptr,
                     0u);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

#define PA_VARIANT(K, KV, B, I, W, P, M) \
  { B, I, W, P, M, &launch_pass<K, KV, B, I, W, P, M> }

template <typename K, bool KV>
struct Variants;
template <>
struct Variants<uint32_t, false> {
  static constexpr PassVariant table[] = {
      PA_VARIANT(uint32_t, false, 512, 24, 4, false, 6), PA_VARIANT(uint32_t, false, 512, 24, 8, false, 6),
      PA_VARIANT(uint32_t, false, 512, 24, 2, false, 6), PA_VARIANT(uint32_t, false, 512, 20, 4, false, 6),
      PA_VARIANT(uint32_t, false, 256, 32, 4, false, 4), PA_VARIANT(uint32_t, false, 256, 48, 4, false, 3),
      PA_VARIANT(uint32_t, false, 512, 32, 4, false, 4), PA_VARIANT(uint32_t, false, 512, 24, 4, true, 4),
  };
};
template <>
struct Variants<uint32_t, true> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint32_t, true, 512, 16, 8, false, 5),
                                          PA_VARIANT(uint32_t, true, 512, 16, 8, false, 4)};
};
template <>
struct Variants<uint64_t, false> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint64_t, false, 512, 8, 8, false, 8),
                                          PA_VARIANT(uint64_t, false, 512, 12, 8, false, 6),
                                          PA_VARIANT(uint64_t, false, 512, 8, 8, true, 4)};
};
template <>
struct Variants<uint64_t, true> {
  static constexpr PassVariant table[] = {PA_VARIANT(uint64_t, true, 512, 8, 8, false, 6),
                                          PA_VARIANT(uint64_t, true, 512, 8, 8, true, 4)};
};

template <typename K, bool KV>
const PassVariant& pick_variant() {
  constexpr size_t count = sizeof(Variants<K, KV>::table) / sizeof(PassVariant);
  static int idx = [] {
    const char* e = std::getenv("PA_SORT_VARIANT");
    int v = e ? std::atoi(e) : 0;
    return (v >= 0 && (size_t)v < count) ? v : 0;
  }();
  return Variants<K, KV>::table[idx];
}

template <typename K, bool KV>
size_t sort_tiles(size_t n) {
  const PassVariant& v = pick_variant<K, KV>();
  const size_t tile = (size_t)v.block * v.items;
  return (n + tile - 1) / tile;
}

// Enqueues the full sort on device buffers.  scratch must hold
// sort_scratch_bytes<K, KV>(n) bytes.
template <typename K, bool KV>
size_t sort_scratch_bytes(size_t n) {
  constexpr int D = sizeof(K);
  Layout L;
  L.take((size_t)D * 256 * 4);  // counts
  L.take((size_t)D * 256 * 4);  // bases
  L.take(D * 4);                // tile counters
  L.take((size_t)D * sort_tiles<K, KV>(n) * 256 * 4);
  return L.off;
}

template <typename K, bool KV>
int enqueue_sort(K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, size_t n, char* scratch, hipStream_t s,
                 bool float_keys) {
  constexpr int D = sizeof(K);
  const PassVariant& var = pick_variant<K, KV>();
  const size_t tiles = sort_tiles<K, KV>(n);
  const size_t ncount = (size_t)D * 256;
  Layout L;
  uint32_t* counts = reinterpret_cast<uint32_t*>(scratch + L.take(ncount * 4));
  uint32_t* bases = reinterpret_cast<uint32_t*>(scratch + L.take(ncount * 4));
  uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch + L.take(D * 4));
  const size_t lb_bytes = (size_t)D * tiles * 256 * 4;
  uint32_t* lb = reinterpret_cast<uint32_t*>(scratch + L.take(lb_bytes));

  {
    ProfScope ps("k_sort_prepare", s);
    hipLaunchKernelGGL(k_sort_prepare, dim3(1), dim3(256), 0, s, counts, (uint32_t)ncount, ctr, (uint32_t)D);
  }
  {
    const size_t nvec = n / (16 / sizeof(K)) + 1;
    const size_t grid = std::max<size_t>(std::min<size_t>((nvec + kSortBlock - 1) / kSortBlock, 2048),
                                         std::min<size_t>(lb_bytes / 16 / kSortBlock + 1, 2048));
    ProfScope ps("k_hist_digits", s);
    hipLaunchKernelGGL(k_hist_digits<K>, dim3((unsigned)grid), dim3(kSortBlock), 0, s, keys, n, counts,
                       reinterpret_cast<uint4*>(lb), lb_bytes / 16, (uint32_t)float_keys);
  }
  {
    ProfScope ps("k_digit_bases", s);
    hipLaunchKernelGGL(k_digit_bases, dim3(D), dim3(256), 0, s, counts, bases);
  }
  K* kin = keys;
  K* kout = ktmp;
  uint32_t* vin = vals;
  uint32_t* vout = vtmp;
  for (int p = 0; p < D; ++p) {
    uint32_t fmode = 0;
    if (float_keys && p == 0) fmode |= kFloatIn;
    if (float_keys && p == D - 1) fmode |= kFloatOut;
    PassArgs a{kin, kout, vin, vout, (uint32_t)n, bases + (size_t)p * 256, lb + (size_t)p * tiles * 256,
               ctr + p, (uint32_t)tiles, (uint32_t)(8 * p), fmode};
    ProfScope ps("k_onesweep", s);
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
  const size_t o_sort = L.take(sort_scratch_bytes<K, KV>(n));
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
