// pa_select.hip -- radix select (k-th smallest u32 or u64 key) for gfx950.
//
// Reference: SelectRadix / RadixSelectiontNonRecursiveInner (RadixSelect.h:83-168).
// The reference histograms the current digit over the candidate range
// (HistogramOneComponentOpt, Histogram.h:103-133), finds the bin holding rank
// k, moves that bin's keys together (MoveOutsideOfKthBinIn, :36-55), narrows
// the range to the bin and repeats for digits 24, 16, 8, 0 (same digit order
// here, so the bin chosen at every pass is the reference's).  u64 keys use the
// same scheme over digits 56 .. 0 (an extension; the reference is 32-bit).
//
// GPU form (the caller's array is never modified):
//   k_select_first  one read of all keys builds the D digit histograms.
//   k_select_walk   (one block) walks the digits from the top: while the
//                   chosen bin still holds every candidate (constant input,
//                   shared high bytes) the next digit's histogram over the same
//                   set is already known, so no further read is needed.
//   k_select_pass   (each remaining digit; returns at once when the answer is
//                   known) histograms the next digit over the keys matching the
//                   fixed high digits.  When the candidates fit the buffer
//                   (half the input) they are compacted in the same read, so
//                   later passes read only the candidates.
//   k_select_pick   (one block) picks that digit's bin.
// Blocks add their LDS counts to global histograms with atomics and the
// one-block kernels read them after the launch boundary: no grid-wide fence
// or last-block election.  Control state lives in device memory; the whole
// select is enqueued without a host round trip.
#include <algorithm>

#include "pa_internal.h"

namespace pa {

constexpr int kSelBlock = 512;
constexpr uint32_t kStage = 4096;  // LDS staging for compaction (keys)

struct SelState {
  uint64_t k;          // rank within the current candidate set
  uint64_t cand;       // number of keys matching (prefix, pmask)
  uint64_t buf_count;  // keys in the compaction buffer
  uint64_t cap;        // capacity of the compaction buffer
  uint64_t prefix;     // fixed high digits of the answer
  uint64_t pmask;      // bits fixed so far
  uint64_t result;
  int32_t shift;       // digit of the next pass (-8 when done)
  uint32_t done;
  uint32_t src;          // 0: read the caller's keys, 1: read the buffer
  uint32_t compact_now;  // the next pass reading the keys compacts candidates
};

__global__ void k_select_init(SelState* st, uint64_t k, uint64_t n, uint64_t cap, int32_t top_shift, uint32_t* zero,
                              uint32_t nzero) {
  if (threadIdx.x == 0) {
    SelState s{};
    s.k = k;
    s.cand = n;
    s.cap = cap;
    s.shift = top_shift;
    *st = s;
  }
  for (uint32_t i = threadIdx.x; i < nzero; i += blockDim.x) zero[i] = 0;
}

// Finds the bin b of `row` (256 counts, in s_row) that holds rank k; returns it
// through s_pick[0] (bin) and s_pick64 (count below the bin).
__device__ void pick_bin(const uint32_t* s_row, uint64_t k, uint32_t* s_wsum, uint32_t* s_pick, uint64_t* s_below) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t c = 0, inc = 0;
  if (tid < 256) {
    c = s_row[tid];
    inc = dev::wave_incl_scan(c, lane);
    if (lane == 63) s_wsum[wave] = inc;
  }
  __syncthreads();
  if (tid < 256) {
    uint64_t pre = 0;
    for (uint32_t w = 0; w < wave; ++w) pre += s_wsum[w];
    const uint64_t below = pre + inc - c;
    if (c > 0 && k >= below && k < below + c) {
      s_pick[0] = tid;
      *s_below = below;
    }
  }
  __syncthreads();
}

// Adds a block's LDS counts into the global histogram (zero bins skipped).
__device__ __forceinline__ void flush_counts(const uint32_t* s_hist, uint32_t* ghist, uint32_t nbins) {
  for (uint32_t i = threadIdx.x; i < nbins; i += kSelBlock)
    if (s_hist[i]) atomicAdd(&ghist[i], s_hist[i]);
}

// One read of all keys: the D digit histograms.  Four 16-byte loads per thread
// are in flight per step.
template <typename K>
__global__ __launch_bounds__(kSelBlock) void k_select_first(const K* __restrict__ keys, size_t n,
                                                            uint32_t* __restrict__ ghist) {
  constexpr int D = sizeof(K);
  constexpr int PER = 16 / sizeof(K);
  constexpr int U = 4;
  __shared__ uint32_t s_hist[D * 256];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < D * 256; i += kSelBlock) s_hist[i] = 0;
  __syncthreads();
  size_t head = ((16 - ((uintptr_t)keys & 15)) & 15) / sizeof(K);
  if (head > n) head = n;
  const size_t nvec = (n - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(keys + head);
  const size_t stride = (size_t)gridDim.x * kSelBlock;
  for (size_t v0 = (size_t)blockIdx.x * kSelBlock + tid; v0 < nvec; v0 += stride * U) {
    uint4 x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (v0 + u * stride < nvec) x[u] = vk[v0 + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (v0 + u * stride >= nvec) continue;
      K kk[PER];
      __builtin_memcpy(kk, &x[u], 16);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        uint32_t dg[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) dg[j] = dev::digit8<K>(kk[j], 8 * d);
        dev::lds_count<PER>(s_hist + d * 256, dg);
      }
    }
  }
  if (blockIdx.x == 0) {
    const size_t tail = head + nvec * PER;
    const size_t extra = head + (n - tail);
    for (size_t e = tid; e < extra; e += kSelBlock) {
      const K key = keys[e < head ? e : tail + (e - head)];
#pragma unroll
      for (int d = 0; d < D; ++d) atomicAdd(&s_hist[d * 256 + dev::digit8<K>(key, 8 * d)], 1u);
    }
  }
  __syncthreads();
  flush_counts(s_hist, ghist, D * 256);
}

// After the first read (one block): walk the digits from the top while the
// chosen bin still holds every candidate (constant input, shared high bytes),
// so no further read is needed for those digits.  Clears the histograms.
template <typename K>
__global__ __launch_bounds__(256) void k_select_walk(SelState* __restrict__ st, uint32_t* __restrict__ ghist) {
  constexpr int D = sizeof(K);
  __shared__ uint32_t s_hist[D * 256];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_pick[1];
  __shared__ uint64_t s_below;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < D * 256; i += 256) {
    s_hist[i] = ghist[i];
    ghist[i] = 0;
  }
  __syncthreads();
  uint64_t k = st->k, cand = st->cand;
  uint64_t prefix = 0, pmask = 0;
  int shift = 8 * (D - 1);
  while (shift >= 0) {
    const int d = shift / 8;
    pick_bin(s_hist + d * 256, k, s_wsum, s_pick, &s_below);
    const uint32_t b = s_pick[0];
    const uint64_t c = s_hist[d * 256 + b];
    prefix |= (uint64_t)b << shift;
    pmask |= (uint64_t)0xFF << shift;
    k -= s_below;
    shift -= 8;
    if (c < cand) {
      cand = c;
      break;
    }
  }
  if (tid == 0) {
    st->k = k;
    st->cand = cand;
    st->prefix = prefix;
    st->pmask = pmask;
    st->shift = shift;
    st->done = shift < 0;
    st->result = prefix;
    st->src = 0;
    st->compact_now = cand <= st->cap;
  }
}

// One filtered histogram pass at st->shift over the keys matching the fixed
// digits; when st->compact_now is set, the matching keys are also appended to
// buf.  Each wave stages its matches in its own LDS slice and flushes the
// slice with one global atomic when it fills and at the end, so no block-wide
// barrier sits inside the loop; one wave scan places a step's matches.
template <typename K>
__global__ __launch_bounds__(kSelBlock) void k_select_pass(const K* __restrict__ keys, size_t n,
                                                           K* __restrict__ buf, const SelState* __restrict__ st,
                                                           uint32_t* __restrict__ ghist,
                                                           unsigned long long* __restrict__ buf_ctr) {
  constexpr int PER = 16 / sizeof(K);
  constexpr int U = 4;
  constexpr int WAVES = kSelBlock / 64;
  constexpr uint32_t STAGE = 64 * U * PER;  // one step's worst case per wave
  __shared__ uint32_t s_hist[256];
  __shared__ K s_stage[WAVES][STAGE];
  if (st->done) return;  // uniform for the whole grid
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const K prefix = (K)st->prefix, pmask = (K)st->pmask;
  const uint32_t shift = (uint32_t)st->shift;
  const bool from_buf = st->src != 0;
  const bool compact = !from_buf && st->compact_now;
  const K* src = from_buf ? buf : keys;
  const size_t m = from_buf ? (size_t)st->buf_count : n;
  if (tid < 256) s_hist[tid] = 0;
  __syncthreads();
  K* stage = s_stage[wave];
  uint32_t staged = 0;  // wave-uniform
  auto flush = [&]() {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(buf_ctr, (unsigned long long)staged);
    base = __shfl(base, 0, 64);
    for (uint32_t j = lane; j < staged; j += 64) buf[base + j] = stage[j];
    staged = 0;
  };
  // Appends this lane's matches (bit u*PER+j of hits: key j of vector u) to
  // the wave's slice, flushing first if the slice cannot hold the step.
  auto take = [&](const uint4 (&x)[U], uint32_t hits) {
    const uint32_t nm = (uint32_t)__popc(hits);
    const uint32_t inc = dev::wave_incl_scan(nm, lane);
    const uint32_t wtot = __shfl(inc, 63, 64);
    if (wtot == 0) return;
    if (staged + wtot > STAGE) flush();
    __builtin_amdgcn_wave_barrier();
    uint32_t pos = staged + inc - nm;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      K xs[PER];
      __builtin_memcpy(xs, &x[u], 16);
#pragma unroll
      for (int j = 0; j < PER; ++j)
        if ((hits >> (u * PER + j)) & 1u) stage[pos++] = xs[j];
    }
    __builtin_amdgcn_wave_barrier();
    staged += wtot;
  };
  size_t head = ((16 - ((uintptr_t)src & 15)) & 15) / sizeof(K);
  if (head > m) head = m;
  const size_t nvec = (m - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(src + head);
  const size_t stride = (size_t)gridDim.x * kSelBlock;
  // Every lane of a wave runs the same iterations (the bound uses the wave's
  // first index), so the wave-level staging stays convergent.
  const size_t wfirst = (size_t)blockIdx.x * kSelBlock + wave * 64;
  for (size_t w0 = wfirst; w0 < nvec; w0 += stride * U) {
    const size_t v0 = w0 + lane;
    uint4 x[U];
    uint32_t hits = 0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (v0 + u * stride < nvec) x[u] = vk[v0 + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (v0 + u * stride >= nvec) continue;
      K xs[PER];
      __builtin_memcpy(xs, &x[u], 16);
      uint32_t h = 0;
#pragma unroll
      for (int j = 0; j < PER; ++j) h |= (uint32_t)((xs[j] & pmask) == prefix) << j;
      if (h == (1u << PER) - 1) {  // wave-aggregated counting for runs (presorted, shared digits)
        uint32_t dg[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) dg[j] = dev::digit8<K>(xs[j], shift);
        dev::lds_count<PER>(s_hist, dg);
      } else {
#pragma unroll
        for (int j = 0; j < PER; ++j)
          if ((h >> j) & 1u) atomicAdd(&s_hist[dev::digit8<K>(xs[j], shift)], 1u);
      }
      hits |= h << (u * PER);
    }
    if (compact) take(x, hits);
  }
  // Head and tail elements outside whole vectors (at most 2 * PER - 2): wave 0 of block 0.
  if (blockIdx.x == 0 && wave == 0) {
    const size_t tail = head + nvec * PER;
    const size_t extra = head + (m - tail);
    for (size_t e0 = 0; e0 < extra; e0 += 64) {
      const size_t e = e0 + lane;
      uint4 x[U] = {};
      uint32_t hits = 0;
      if (e < extra) {
        const K key = src[e < head ? e : tail + (e - head)];
        __builtin_memcpy(&x[0], &key, sizeof(K));
        if ((key & pmask) == prefix) {
          atomicAdd(&s_hist[dev::digit8<K>(key, shift)], 1u);
          hits = 1;
        }
      }
      if (compact) take(x, hits);
    }
  }
  if (compact && staged) flush();
  __syncthreads();
  flush_counts(s_hist, ghist, 256);
}

// After a pass (one block): picks the bin of digit st->shift holding rank k,
// fixes that digit, and switches later passes to the compacted buffer when
// this pass compacted.  Clears the histogram and the buffer counter.
template <typename K>
__global__ __launch_bounds__(256) void k_select_pick(SelState* __restrict__ st, uint32_t* __restrict__ ghist,
                                                     unsigned long long* __restrict__ buf_ctr) {
  __shared__ uint32_t s_hist[256];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_pick[1];
  __shared__ uint64_t s_below;
  if (st->done) return;
  const uint32_t tid = threadIdx.x;
  s_hist[tid] = ghist[tid];
  ghist[tid] = 0;
  __syncthreads();
  const uint64_t k = st->k;
  const uint32_t shift = (uint32_t)st->shift;
  pick_bin(s_hist, k, s_wsum, s_pick, &s_below);
  if (tid == 0) {
    const uint32_t b = s_pick[0];
    const bool compacted = !st->src && st->compact_now;
    st->k = k - s_below;
    st->cand = s_hist[b];
    st->prefix |= (uint64_t)b << shift;
    st->pmask |= (uint64_t)0xFF << shift;
    st->shift = (int32_t)shift - 8;
    st->done = shift == 0;
    st->result = st->prefix;
    if (compacted) {
      st->src = 1;
      st->buf_count = *buf_ctr;
      *buf_ctr = 0;
    }
    st->compact_now = !st->src && st->cand <= st->cap;
  }
}

// a[k] = the selected value (the reference's post-condition, RadixSelect.h:149).
template <typename K>
__global__ void k_select_place(const SelState* st, K* dst) { *dst = (K)st->result; }

template <typename K>
__global__ void k_select_result(const SelState* st, K* dst) { *dst = (K)st->result; }

// Select on device keys into caller-provided scratch (no allocation, no
// locking): used by the partition.  The selected key is written to
// *value_out (device memory).
// Zeroed control words: D digit histograms, one pass histogram, and the 64-bit
// compaction counter.
template <typename K>
static constexpr size_t select_zero_words() { return sizeof(K) * 256 + 256 + 2; }

template <typename K>
static size_t select_bytes(size_t n) {
  Layout L;
  L.take(sizeof(SelState));
  L.take(select_zero_words<K>() * 4);
  L.take(std::max<size_t>(n / 2, 1) * sizeof(K));
  return L.off;
}

// The launch sequence of one select, all stream-ordered with no host round
// trip: the first read with the digit walk, then for each remaining digit a
// filtered pass and its bin pick (both return at once when the answer is
// known), then the result.
template <typename K>
static void launch_select(const K* din, size_t n, size_t k, SelState* st, uint32_t* z, uint32_t nz, K* buf, size_t cap,
                          hipStream_t s) {
  constexpr int D = sizeof(K);
  uint32_t* ghistD = z;
  uint32_t* ghist = z + D * 256;
  unsigned long long* buf_ctr = reinterpret_cast<unsigned long long*>(ghist + 256);
  hipLaunchKernelGGL(k_select_init, dim3(1), dim3(256), 0, s, st, (uint64_t)k, (uint64_t)n, (uint64_t)cap,
                     (int32_t)(8 * (D - 1)), z, nz);
  const size_t nvec = n / (16 / sizeof(K));
  const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>((nvec + 4 * kSelBlock - 1) / (4 * kSelBlock), 512));
  {
    ProfScope ps("k_select_first", s);
    hipLaunchKernelGGL(k_select_first<K>, dim3(grid), dim3(kSelBlock), 0, s, din, n, ghistD);
  }
  hipLaunchKernelGGL(k_select_walk<K>, dim3(1), dim3(256), 0, s, st, ghistD);
  for (int pass = 0; pass < D - 1; ++pass) {
    ProfScope ps("k_select_pass", s);
    hipLaunchKernelGGL(k_select_pass<K>, dim3(grid), dim3(kSelBlock), 0, s, din, n, buf, st, ghist, buf_ctr);
    hipLaunchKernelGGL(k_select_pick<K>, dim3(1), dim3(256), 0, s, st, ghist, buf_ctr);
  }
}

template <typename K>
static int select_enqueue(const K* din, size_t n, size_t k, char* b, K* value_out, hipStream_t s) {
  const size_t cap = std::max<size_t>(n / 2, 1);
  const size_t nz = select_zero_words<K>();
  Layout L;
  SelState* st = (SelState*)(b + L.take(sizeof(SelState)));
  uint32_t* z = (uint32_t*)(b + L.take(nz * 4));
  K* buf = (K*)(b + L.take(cap * sizeof(K)));
  launch_select<K>(din, n, k, st, z, (uint32_t)nz, buf, cap, s);
  hipLaunchKernelGGL(k_select_result<K>, dim3(1), dim3(1), 0, s, st, value_out);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

size_t select_scratch_bytes_u32(size_t n) { return select_bytes<uint32_t>(n); }
int select_enqueue_u32(const uint32_t* din, size_t n, size_t k, char* scratch, uint32_t* value_out, hipStream_t s) {
  return select_enqueue<uint32_t>(din, n, k, scratch, value_out, s);
}

template <typename K>
static int select_impl(const K* keys, size_t n, size_t k, K* result, K* place, hipStream_t s, const char* who) {
  if (!keys || !result) return arg_error(who);
  if (n == 0 || k >= n) return arg_error("select: need 0 <= k < n");
  if (n > 0xFFFFFFFFull) return arg_error("select: n must be < 2^32");
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = is_device_ptr(keys);
  const bool dres = is_device_ptr(result);
  const size_t cap = std::max<size_t>(n / 2, 1);
  Layout L;
  const size_t o_st = L.take(sizeof(SelState));
  const size_t nz = select_zero_words<K>();
  const size_t o_z = L.take(nz * 4);
  const size_t o_res = L.take(sizeof(K));
  const size_t o_buf = L.take(cap * sizeof(K));
  const size_t o_in = dk ? 0 : L.take(n * sizeof(K));
  void* base;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* b = (char*)base;
  SelState* st = (SelState*)(b + o_st);
  uint32_t* z = (uint32_t*)(b + o_z);
  K* dres_tmp = (K*)(b + o_res);
  K* buf = (K*)(b + o_buf);
  const K* din = keys;
  if (!dk) {
    PA_TRY(stage_to_device(ctx, b + o_in, keys, n * sizeof(K)));
    din = (const K*)(b + o_in);
  }
  launch_select<K>(din, n, k, st, z, (uint32_t)nz, buf, cap, s);
  if (place && dk) hipLaunchKernelGGL(k_select_place<K>, dim3(1), dim3(1), 0, s, st, place);
  hipLaunchKernelGGL(k_select_result<K>, dim3(1), dim3(1), 0, s, st, dres ? result : dres_tmp);
  PA_HIP(hipGetLastError());
  if (!dres) PA_HIP(hipMemcpyAsync(result, dres_tmp, sizeof(K), hipMemcpyDeviceToHost, s));
  if (!dres || !dk) PA_HIP(hipStreamSynchronize(s));
  if (place && !dk) {  // host keys: store the value into the caller's array
    if (dres)
      PA_HIP(hipMemcpy(place, dres_tmp, sizeof(K), hipMemcpyDeviceToHost));
    else
      *place = *result;
  }
  return PA_OK;
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_select_u32(const uint32_t* keys, size_t n, size_t k, uint32_t* result, pa_stream_t stream) {
  return select_impl<uint32_t>(keys, n, k, result, nullptr, (hipStream_t)stream, "pa_select_u32: null pointer");
}

int pa_select_u64(const uint64_t* keys, size_t n, size_t k, uint64_t* result, pa_stream_t stream) {
  return select_impl<uint64_t>(keys, n, k, result, nullptr, (hipStream_t)stream, "pa_select_u64: null pointer");
}

int pa_select_radix_u32(uint32_t* keys, size_t n, size_t k, uint32_t* result, pa_stream_t stream) {
  return select_impl<uint32_t>(keys, n, k, result, keys ? keys + k : nullptr, (hipStream_t)stream,
                               "pa_select_radix_u32: null pointer");
}

}  // extern "C"
