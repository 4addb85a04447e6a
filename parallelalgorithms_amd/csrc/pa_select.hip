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
//   k_select_init   (16 blocks) state, cleared control words, and the top
//                   digits of an evenly strided sample of the keys, counted
//                   per block.
//   k_select_first  every block sums the sample's counts into a band of
//                   top-digit bins predicted to hold the answer, then one
//                   read of all keys builds the top digit's histogram and the
//                   keys' minimum and maximum, and compacts the keys of the
//                   band; when the answer's top digit is in the band, the
//                   later passes read only those keys.
//   k_select_pass   (launch i of D - 1; returns at once when the answer is
//                   known) first derives its state from the counts the
//                   previous launch left: launch 0 fixes the high bytes that
//                   the minimum and maximum share (every key has them; equal
//                   minimum and maximum end the select there) and, when the
//                   top byte differs, picks the top digit's bin from the
//                   histogram; launch i > 0 picks the bin of the
//                   previous pass's digit.  Every block computes the same
//                   state from the same counts.  Then it histograms the next
//                   digit over the keys matching the fixed high digits; when
//                   the candidates fit the buffer (half the input) they are
//                   compacted in the same read, so later passes read only the
//                   candidates.
//   k_select_final  (one block) picks the last digit and writes the value.
// Blocks add their LDS counts to global histograms (one per launch) with
// atomics and the next launch reads them after the launch boundary: no
// grid-wide fence or last-block election.  Control state lives in device
// memory; the whole select is enqueued without a host round trip.
#include <algorithm>

#include "pa_internal.h"

namespace pa {

constexpr int kSelBlock = 512;

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
  uint32_t src;          // 0: read the caller's keys, 1: read the buffer, 2: read block regions of it
  uint32_t compact_now;  // the next pass reading the keys compacts candidates
  // Top-digit bins [band_lo, band_hi] predicted from a sample to hold the
  // answer; when band_on the first read also compacts their keys.
  uint32_t band_lo, band_hi, band_on;
};

// Keys sampled to predict the answer's top-digit bins: kSelRuns runs of
// kSelRun consecutive keys at an even stride (single keys at that stride would
// each cost a page walk), or every key when n < kSelSample.  The sample is
// read by kSelInitBlocks blocks on as many CUs (one block alone took 8.5 us,
// its loads and page walks queued on one CU), each writing the 256 top-digit
// counts of its share; the first read sums them and picks the band.
constexpr uint32_t kSelSample = 32768;
constexpr uint32_t kSelInitBlock = 1024;
constexpr uint32_t kSelInitBlocks = 16;
constexpr uint32_t kSelRun = 128;  // two halves of 64 keys, one a wave load
constexpr uint32_t kSelRuns = kSelSample / kSelRun;
constexpr uint32_t kSelPartWords = kSelInitBlocks * 256;

// Initial state, cleared control words and the sample's per-block top-digit
// counts.  Sample element e (0 <= e < kSelSample) is key e when every key is
// sampled, else key e % kSelRun of run e / kSelRun; thread g of the grid takes
// elements g and g + kSelInitBlocks * kSelInitBlock, so the 64 lanes of a wave
// read 64 consecutive keys.  Every load is issued before any is counted; the
// counts go to lane-private sub-histograms (a repeated key costs no more than
// distinct ones), whose sums block b writes to parts[b][0..255].
template <typename K>
__global__ __launch_bounds__(kSelInitBlock) void k_select_init(SelState* st, const K* __restrict__ keys, uint64_t k,
                                                               uint64_t n, uint64_t cap, int32_t top_shift,
                                                               uint32_t* zero, uint32_t nzero,
                                                               uint32_t* __restrict__ parts) {
  using H = dev::SubHist<32, false>;
  constexpr uint32_t T = kSelInitBlocks * kSelInitBlock;
  constexpr uint32_t SPT = kSelSample / T;  // sample elements per thread
  static_assert(SPT * T == kSelSample && kSelRun % 64 == 0, "whole runs, one wave load per 64 keys");
  __shared__ __align__(16) uint32_t s_sub[H::words(256)];
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const uint32_t g = blockIdx.x * kSelInitBlock + tid;
  const bool all = n < kSelSample;
  // Run r starts at floor(r * span / (kSelRuns - 1)) = r * q + r * rem / (kSelRuns - 1)
  // (one 64-bit division per thread instead of one per run).
  const uint64_t span = all ? 0 : n - kSelRun;
  const uint64_t q = span / (kSelRuns - 1);
  const uint32_t rem = (uint32_t)(span - q * (kSelRuns - 1));
  K v[SPT];
  bool ok[SPT];
#pragma unroll
  for (uint32_t j = 0; j < SPT; ++j) {
    const uint32_t e = g + j * T, r = e / kSelRun;
    const uint64_t i = all ? e : r * q + (uint64_t)(r * rem) / (kSelRuns - 1) + e % kSelRun;
    ok[j] = i < n;
    v[j] = ok[j] ? keys[i] : K(0);
  }
  uint4* const z4 = reinterpret_cast<uint4*>(zero);  // zero is 16-byte aligned
  for (uint32_t i = g; i < nzero / 4; i += T) z4[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0)
    for (uint32_t i = nzero / 4 * 4 + tid; i < nzero; i += kSelInitBlock) zero[i] = 0;
  for (uint32_t i = tid; i < H::words(256) / 4; i += kSelInitBlock)
    reinterpret_cast<uint4*>(s_sub)[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < SPT; ++j)
    if (ok[j]) H::add(s_sub, dev::digit8<K>(v[j], (uint32_t)top_shift), lane, 1u);
  __syncthreads();
  if (tid < 256) parts[blockIdx.x * 256 + tid] = H::sum(s_sub, tid);
  if (g == 0) {
    SelState t{};
    t.k = k;
    t.cand = n;
    t.cap = cap;
    t.shift = top_shift;
    *st = t;  // band_* are written by the first read's block 0
  }
}

// The band, computed by every block of the first read from the init's
// per-block counts pc (thread tid < 256 holds bin tid's sum): the answer's
// rank k maps to sample rank k * S / n, whose sampling error has a standard
// deviation of at most sqrt(S) / 2, so the top-digit bins holding sample ranks
// within 4 of those deviations of it (a contiguous range) hold the answer with
// near certainty.  The band is used only when the sample puts at most half of
// the keys in it, so its keys fit the compaction buffer (cap = n / 2) unless
// the sample misjudged them; a miss or an overflow only costs the second read
// the select would do anyway.  (A band straddling the boundary of two large
// bins, as a presorted input's middle does, holds a third of the keys.)
// Returns (lo bin, hi bin, on) in out[0..2]; every thread of the block calls it.
__device__ void sample_band(uint32_t pc, uint64_t k, uint64_t n, uint32_t* s_wsum, uint64_t* s_band, uint32_t* out) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool all = n < kSelSample;
  const uint64_t S = all ? n : kSelSample;  // keys sampled
  uint32_t inc = 0;
  if (tid < 4) s_band[tid] = tid == 1 ? 255u : (tid == 3 ? S : 0u);
  if (tid < 256) {
    inc = dev::wave_incl_scan(pc, lane);
    if (lane == 63) s_wsum[wave] = inc;
  }
  __syncthreads();
  const uint64_t ks = k * S / n;
  uint32_t delta = (uint32_t)sqrtf((float)S);  // then corrected to ceil(sqrt(S))
  while ((uint64_t)delta * delta < S) ++delta;
  while (delta > 1 && (uint64_t)(delta - 1) * (delta - 1) >= S) --delta;
  delta = 2 * delta + 1;  // 4 deviations of at most sqrt(S) / 2
  const uint64_t lo_r = ks > delta ? ks - delta : 0, hi_r = ks + delta < S ? ks + delta : S - 1;
  if (tid < 256) {
    uint32_t pre = 0;
    for (uint32_t w = 0; w < wave; ++w) pre += s_wsum[w];
    const uint64_t below = pre + inc - pc;
    if (pc > 0 && lo_r >= below && lo_r < below + pc) {
      s_band[0] = tid;
      s_band[2] = below;
    }
    if (pc > 0 && hi_r >= below && hi_r < below + pc) {
      s_band[1] = tid;
      s_band[3] = below + pc;
    }
  }
  __syncthreads();
  out[0] = (uint32_t)s_band[0];
  out[1] = (uint32_t)s_band[1];
  out[2] = !all && (s_band[3] - s_band[2]) * 2 <= S;
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

// Global histograms are kept as kSelCopies copies of the 256 bins: block b
// adds its counts into copy b % kSelCopies, so the adds of a launch's 768
// blocks meet 48 to an address rather than 768 (adds to one address are
// applied one after another), and a reader sums the copies.
constexpr uint32_t kSelCopies = 16;
constexpr uint32_t kSelHistWords = 256 * kSelCopies;

__device__ __forceinline__ uint32_t hist_read(const uint32_t* __restrict__ g, uint32_t bin) {
  uint32_t c[kSelCopies];
#pragma unroll
  for (uint32_t i = 0; i < kSelCopies; ++i) c[i] = g[i * 256 + bin];
  uint32_t t = 0;
#pragma unroll
  for (uint32_t i = 0; i < kSelCopies; ++i) t += c[i];
  return t;
}

// Adds a block's LDS counts into its copy of a global histogram (zero bins
// skipped).
__device__ __forceinline__ void flush_counts(const uint32_t* s_hist, uint32_t* ghist) {
  uint32_t* const g = ghist + (blockIdx.x % kSelCopies) * 256;
  for (uint32_t i = threadIdx.x; i < 256; i += kSelBlock)
    if (s_hist[i]) atomicAdd(&g[i], s_hist[i]);
}

template <typename K>
__device__ __forceinline__ K umax(K a, K b) { return a > b ? a : b; }

// Per-wave compaction of matching keys into buf.  Each wave stages its
// matches in its own LDS slice and flushes the slice with one global atomic
// when a step's matches would not fit; finish() sends every wave's last slice
// out under one counter reservation per block (every wave reaches it at once,
// and thousands of same-address atomics at the end of a launch would
// serialise).  Keys whose slot is past cap are dropped; the counter still
// counts them, so a reader sees the overflow.  All lanes of a wave call take()
// together (the loops bound their iterations by the wave's first index).
template <typename K, int U>
struct WaveCompactor {
  static constexpr int PER = 16 / sizeof(K);
  static constexpr uint32_t STAGE = 64 * U * PER;  // one step's worst case per wave
  K* stage;
  K* buf;
  uint64_t cap;
  unsigned long long* ctr;
  uint32_t lane;
  uint32_t staged = 0;  // wave-uniform

  __device__ void put(unsigned long long base) {
    for (uint32_t j = lane; j < staged; j += 64)
      if (base + j < cap) buf[base + j] = stage[j];
    staged = 0;
  }
  __device__ void flush() {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(ctr, (unsigned long long)staged);
    put(__shfl(base, 0, 64));
  }
  // Appends this lane's matches (bit u * PER + j of hits: key j of vector u).
  __device__ void take(const uint4 (&x)[U], uint32_t hits) {
    if (!__any(hits != 0)) return;
    const uint32_t nm = (uint32_t)__popc(hits);
    const uint32_t inc = dev::wave_incl_scan(nm, lane);
    const uint32_t wtot = __shfl(inc, 63, 64);
    if (wtot == 0) return;
    if (staged + wtot > STAGE) flush();
    __builtin_amdgcn_wave_barrier();
    uint32_t pos = staged + inc - nm;
    // Per vector, each lane stores its hits lowest first, one a round, the
    // key picked from the vector's PER keys by selects: hits are sparse, so
    // a round in which any lane of the wave has a second hit in one vector
    // is rare, against PER vote-guarded stores per vector.
#pragma unroll
    for (int u = 0; u < U; ++u) {
      K xs[PER];
      __builtin_memcpy(xs, &x[u], 16);
      uint32_t h = (hits >> (u * PER)) & ((1u << PER) - 1u);
      while (__any(h != 0)) {
        if (h) {
          const uint32_t c = (uint32_t)__builtin_ctz(h);
          K v = xs[0];
#pragma unroll
          for (int j = 1; j < PER; ++j) v = c == (uint32_t)j ? xs[j] : v;
          stage[pos++] = v;
          h &= h - 1u;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    staged += wtot;
  }
  // Block-wide (every thread calls it): the waves' last slices.
  template <int WAVES>
  __device__ void finish(uint32_t* s_wcnt, unsigned long long* s_bbase, uint32_t tid, uint32_t wave) {
    if (lane == 0) s_wcnt[wave] = staged;
    __syncthreads();
    if (tid == 0) {
      uint32_t total = 0;
      for (int w = 0; w < WAVES; ++w) total += s_wcnt[w];
      *s_bbase = total ? atomicAdd(ctr, (unsigned long long)total) : 0ull;
    }
    __syncthreads();
    unsigned long long base = *s_bbase;
    for (uint32_t w = 0; w < wave; ++w) base += s_wcnt[w];
    put(base);
  }
};

// One read of all keys: the top digit's histogram, and each block's minimum
// and maximum key (kept as max(~key) and max(key)) in mm[2 * block] and
// mm[2 * block + 1].  When the band is on, block b also compacts the keys of
// its share whose top digit lies in the band into its own region of buf
// (cap / grid keys at buf + b * (cap / grid); band_counts[b] keys, more than
// the region when it overflowed), so no counter is shared between blocks.
// Two 16-byte loads per thread per step, issued a step ahead.
template <typename K>
__global__ __launch_bounds__(kSelBlock) void k_select_first(const K* __restrict__ keys, size_t n,
                                                            uint32_t* __restrict__ ghist, K* __restrict__ mm,
                                                            SelState* __restrict__ S, K* __restrict__ buf,
                                                            uint32_t* __restrict__ band_counts,
                                                            const uint32_t* __restrict__ parts) {
  constexpr int D = sizeof(K);
  constexpr int PER = 16 / sizeof(K);
  // Two loads per step keep the LDS at 48 KB (three blocks per CU).
  constexpr int U = 2;
  constexpr int WAVES = kSelBlock / 64;
  using C = WaveCompactor<K, U>;
  using H = dev::SubHist<32, false>;
  __shared__ __align__(16) uint32_t s_sub[H::words(256)];
  __shared__ K s_mm[WAVES][2];
  __shared__ K s_stage[WAVES][C::STAGE];
  __shared__ uint32_t s_wcnt[WAVES];
  __shared__ unsigned long long s_bbase;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t cap = S->cap;
  size_t head = ((16 - ((uintptr_t)keys & 15)) & 15) / sizeof(K);
  if (head > n) head = n;
  const size_t nvec = (n - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(keys + head);
  const size_t stride = (size_t)gridDim.x * kSelBlock;
  // Every lane of a wave runs the same iterations (the bound uses the wave's
  // first index), so the wave-level compaction stays convergent.
  const size_t wfirst = (size_t)blockIdx.x * kSelBlock + wave * 64;
  // The loop is software-pipelined: a step's loads are issued before the
  // previous step's keys are counted and compacted.
  // The sample's per-block counts are loaded before the first step's keys, so
  // the band is computed while those keys are in flight.
  uint32_t pv[kSelInitBlocks] = {};
  if (tid < 256) {
#pragma unroll
    for (uint32_t b = 0; b < kSelInitBlocks; ++b) pv[b] = parts[b * 256 + tid];
  }
  uint4 nx[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (wfirst + lane + u * stride < nvec) nx[u] = vk[wfirst + lane + u * stride];
  __shared__ uint32_t s_bwsum[4];
  __shared__ uint64_t s_band[4];
  uint32_t bd[3];
  {
    uint32_t pc = 0;
#pragma unroll
    for (uint32_t b = 0; b < kSelInitBlocks; ++b) pc += pv[b];
    sample_band(pc, S->k, n, s_bwsum, s_band, bd);
  }
  if (blockIdx.x == 0 && tid == 0) {  // for the passes
    S->band_lo = bd[0];
    S->band_hi = bd[1];
    S->band_on = bd[2];
  }
  const uint32_t band = bd[2];
  const uint32_t band_lo = bd[0], band_w = bd[1] - bd[0];
  __shared__ unsigned long long s_ctr;
  if (tid == 0) s_ctr = 0;
  const uint64_t region = cap / gridDim.x;
  C comp{s_stage[wave], buf + (size_t)blockIdx.x * region, region, &s_ctr, lane};
  for (uint32_t i = tid; i < H::words(256) / 4; i += kSelBlock) reinterpret_cast<uint4*>(s_sub)[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  K mn = ~K(0), hi = 0;  // min(key), max(key) over this thread's keys
  const uint32_t lane_c = lane & 31;
  for (size_t w0 = wfirst; w0 < nvec; w0 += stride * U) {
    const size_t v0 = w0 + lane;
    uint4 x[U];
    uint32_t hits = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = nx[u];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (v0 + (u + U) * stride < nvec) nx[u] = vk[v0 + (u + U) * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (v0 + u * stride >= nvec) continue;
      K kk[PER];
      __builtin_memcpy(kk, &x[u], 16);
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        const uint32_t dg = dev::digit8<K>(kk[j], 8 * (D - 1));
        mn = kk[j] < mn ? kk[j] : mn;
        hi = umax<K>(hi, kk[j]);
        hits |= (uint32_t)(dg - band_lo <= band_w) << (u * PER + j);
        // Copy lane % 32 of bin dg: word dg * 32 + lane % 32 (the SubHist<32, false> layout).
        atomicAdd(&s_sub[(dg << 5) | lane_c], 1u);
      }
    }
    if (band) comp.take(x, hits);
  }
  // Head and tail elements outside whole vectors: wave 0 of block 0.
  if (blockIdx.x == 0 && wave == 0) {
    const size_t tail = head + nvec * PER;
    const size_t extra = head + (n - tail);
    for (size_t e0 = 0; e0 < extra; e0 += 64) {
      const size_t e = e0 + lane;
      uint4 x[U] = {};
      uint32_t hits = 0;
      if (e < extra) {
        const K key = keys[e < head ? e : tail + (e - head)];
        __builtin_memcpy(&x[0], &key, sizeof(K));
        const uint32_t dg = dev::digit8<K>(key, 8 * (D - 1));
        H::add(s_sub, dg, lane, 1u);
        mn = key < mn ? key : mn;
        hi = umax<K>(hi, key);
        hits = dg - band_lo <= band_w;
      }
      if (band) comp.take(x, hits);
    }
  }
  if (band) {
    comp.template finish<WAVES>(s_wcnt, &s_bbase, tid, wave);
    __syncthreads();
    if (tid == 0) band_counts[blockIdx.x] = (uint32_t)(s_ctr < 0xFFFFFFFFull ? s_ctr : 0xFFFFFFFFull);
  }
  // Wave maxima, then the block's, stored without atomics (the next launch
  // reduces the blocks' values); the minimum is kept as the maximum of ~key.
  K lo_c = (K)~mn;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo_c = umax<K>(lo_c, (K)__shfl_xor(lo_c, off, 64));
    hi = umax<K>(hi, (K)__shfl_xor(hi, off, 64));
  }
  if (lane == 0) {
    s_mm[wave][0] = lo_c;
    s_mm[wave][1] = hi;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < WAVES; ++w) {
      lo_c = umax<K>(lo_c, s_mm[w][0]);
      hi = umax<K>(hi, s_mm[w][1]);
    }
    mm[2 * blockIdx.x] = lo_c;
    mm[2 * blockIdx.x + 1] = hi;
  }
  if (tid < 256) {
    const uint32_t c = H::sum(s_sub, tid);
    if (c) atomicAdd(&ghist[(blockIdx.x % kSelCopies) * 256 + tid], c);
  }
}

// The state a pass works with, computed by every block of that pass from the
// histogram the previous step left (every block reads the same counts and
// picks the same bin, so no separate pick launch is needed):
//   step 0: fix the high bytes the keys' minimum and maximum share, then
//           pick the top digit's bin if the top byte differs (walk_state);
//   step i: pick the bin of the previous pass's digit.
// Block 0 stores the result in S[i + 1] for the next step.  Every thread of
// the block must call these (they synchronise the block).
struct SelShared {
  uint32_t wsum[4];
  uint32_t pick[1];
  uint64_t below;
  uint64_t mm[kSelBlock / 64][2];
};

// Step 0: the high bytes shared by the minimum mn and maximum mx are the
// answer's (every key lies between them); if they are equal the select is
// over.  When the top byte differs, its bin is picked from the first read's
// histogram (the candidates are all keys), else the passes start at the
// first differing byte.
template <typename K>
__device__ void walk_state(const uint32_t* s_top, K mn, K mx, SelState& t, SelShared& sh) {
  constexpr int D = sizeof(K);
  const uint64_t diff = (uint64_t)(mn ^ mx);
  int hd = -1;  // highest differing byte
  for (int d = D - 1; d >= 0; --d)
    if ((diff >> (8 * d)) & 0xFF) {
      hd = d;
      break;
    }
  t.src = 0;
  if (hd < 0) {  // constant input
    t.prefix = t.result = (uint64_t)mn;
    t.pmask = ~0ull;
    t.shift = -8;
    t.done = 1;
    t.compact_now = 0;
    return;
  }
  const uint64_t above = hd == 7 ? 0 : (~0ull << (8 * (hd + 1)));
  t.prefix = (uint64_t)mn & above;
  t.pmask = above;
  if (hd == D - 1) {
    pick_bin(s_top, t.k, sh.wsum, sh.pick, &sh.below);
    const uint32_t b = sh.pick[0];
    t.k -= sh.below;
    t.cand = s_top[b];
    t.prefix |= (uint64_t)b << (8 * hd);
    t.pmask |= (uint64_t)0xFF << (8 * hd);
    t.shift = 8 * hd - 8;
    __syncthreads();  // sh is reused by the caller
  } else {
    t.shift = 8 * hd;
  }
  t.done = 0;  // the differing byte hd (or the bytes below the top pick) remain
  t.result = t.prefix;
  t.compact_now = t.cand <= t.cap;
}

// prev: the state the previous pass used; hist: that pass's histogram of
// digit prev.shift; compacted: that pass compacted the candidates, buf_count
// of them.
__device__ void pick_state(const uint32_t* s_row, SelState& t, bool compacted, uint64_t buf_count, SelShared& sh) {
  const uint32_t shift = (uint32_t)t.shift;
  pick_bin(s_row, t.k, sh.wsum, sh.pick, &sh.below);
  const uint32_t b = sh.pick[0];
  t.k -= sh.below;
  t.cand = s_row[b];
  t.prefix |= (uint64_t)b << shift;
  t.pmask |= (uint64_t)0xFF << shift;
  t.shift = (int32_t)shift - 8;
  t.done = shift == 0;
  t.result = t.prefix;
  if (compacted) {
    t.src = 1;
    t.buf_count = buf_count;
  }
  t.compact_now = !t.src && t.cand <= t.cap;
  __syncthreads();  // sh is reused by the caller
}

// The state of step `step` (see above), from S[step] and the histogram the
// previous step left; block 0 stores it in S[step + 1].
template <typename K>
__device__ SelState step_state(int step, const SelState* __restrict__ S, const uint32_t* __restrict__ ghist_first,
                               const uint32_t* __restrict__ ghist_prev,
                               const unsigned long long* __restrict__ ctr_prev,
                               const uint32_t* __restrict__ band_counts, SelState* __restrict__ S_next,
                               uint32_t* s_hist, SelShared& sh) {
  SelState t = S[step];
  if (step == 0) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_hist[i] = hist_read(ghist_first, i);
    __syncthreads();
    // Minimum and maximum over the first read's per-block values (that
    // launch had as many blocks as this one).
    const K* mm = reinterpret_cast<const K*>(ghist_first + kSelHistWords);
    K lo_c = 0, hi = 0;
    for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
      lo_c = umax<K>(lo_c, mm[2 * i]);
      hi = umax<K>(hi, mm[2 * i + 1]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo_c = umax<K>(lo_c, (K)__shfl_xor(lo_c, off, 64));
      hi = umax<K>(hi, (K)__shfl_xor(hi, off, 64));
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
      sh.mm[wave][0] = lo_c;
      sh.mm[wave][1] = hi;
    }
    __syncthreads();
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
      lo_c = umax<K>(lo_c, (K)sh.mm[w][0]);
      hi = umax<K>(hi, (K)sh.mm[w][1]);
    }
    walk_state<K>(s_hist, (K)~lo_c, hi, t, sh);
    // The first read compacted the keys of the band's top-digit bins into
    // block regions: when the answer's top digit is among them and no region
    // overflowed, every candidate is in the buffer and no later pass reads
    // the whole input (block b of each pass reads region b).
    constexpr uint32_t top = 8 * (sizeof(K) - 1);
    const uint32_t region = (uint32_t)(t.cap / gridDim.x);
    __shared__ uint32_t s_over;
    if (threadIdx.x == 0) s_over = 0;
    __syncthreads();
    if (t.band_on)
      for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x)
        if (band_counts[i] > region) s_over = 1;
    __syncthreads();
    if (!t.done && t.band_on && !s_over && ((t.pmask >> top) & 0xFF) == 0xFF) {
      const uint32_t b = (uint32_t)(t.prefix >> top) & 0xFF;
      if (b >= t.band_lo && b <= t.band_hi) {
        t.src = 2;
        t.buf_count = region;
        t.compact_now = 0;
      }
    }
  } else if (!t.done) {
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_hist[i] = hist_read(ghist_prev, i);
    __syncthreads();
    const bool compacted = !t.src && t.compact_now;
    pick_state(s_hist, t, compacted, compacted ? *ctr_prev : 0, sh);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *S_next = t;
  return t;
}

// One filtered histogram pass (launch `step`) over the keys matching the
// fixed digits; when the state says so, the matching keys are also appended
// to buf.  Each wave stages its matches in its own LDS slice and flushes the
// slice with one global atomic when it fills and at the end, so no block-wide
// barrier sits inside the loop; one wave scan places a step's matches.
template <typename K>
__global__ __launch_bounds__(kSelBlock) void k_select_pass(const K* __restrict__ keys, size_t n,
                                                           K* __restrict__ buf, int step, SelState* __restrict__ S,
                                                           const uint32_t* __restrict__ ghist_first,
                                                           uint32_t* __restrict__ ghist,
                                                           unsigned long long* __restrict__ ctr,
                                                           const uint32_t* __restrict__ band_counts) {
  constexpr int PER = 16 / sizeof(K);
  constexpr int U = 4;
  constexpr int WAVES = kSelBlock / 64;
  using C = WaveCompactor<K, U>;
  __shared__ uint32_t s_hist[256];
  __shared__ K s_stage[WAVES][C::STAGE];
  __shared__ uint32_t s_wcnt[WAVES];
  __shared__ unsigned long long s_bbase;
  __shared__ SelShared sh;
  const SelState t = step_state<K>(step, S, ghist_first, step ? ghist + (step - 1) * kSelHistWords : nullptr,
                                   step ? ctr + (step - 1) : nullptr, band_counts, S + step + 1, s_hist, sh);
  if (t.done) return;  // uniform for the whole grid
  uint32_t* const gh = ghist + step * kSelHistWords;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const K prefix = (K)t.prefix, pmask = (K)t.pmask;
  const uint32_t shift = (uint32_t)t.shift;
  const bool from_buf = t.src != 0;
  const bool regions = t.src == 2;  // block b reads region b of the buffer
  const bool compact = !from_buf && t.compact_now;
  const K* src = regions ? buf + (size_t)blockIdx.x * t.buf_count : (from_buf ? buf : keys);
  const size_t m = regions ? band_counts[blockIdx.x] : (from_buf ? (size_t)t.buf_count : n);
  C comp{s_stage[wave], buf, t.cap, ctr + step, lane};
  __syncthreads();  // s_hist held the previous histogram until here
  if (tid < 256) s_hist[tid] = 0;
  __syncthreads();
  size_t head = ((16 - ((uintptr_t)src & 15)) & 15) / sizeof(K);
  if (head > m) head = m;
  const size_t nvec = (m - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(src + head);
  const size_t stride = regions ? (size_t)kSelBlock : (size_t)gridDim.x * kSelBlock;
  // Every lane of a wave runs the same iterations (the bound uses the wave's
  // first index), so the wave-level staging stays convergent.
  const size_t wfirst = (regions ? 0 : (size_t)blockIdx.x * kSelBlock) + wave * 64;
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
    if (compact) comp.take(x, hits);
  }
  // Head and tail elements outside whole vectors (at most 2 * PER - 2): wave 0
  // of block 0, or of every block when each reads its own region.
  if ((regions || blockIdx.x == 0) && wave == 0) {
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
      if (compact) comp.take(x, hits);
    }
  }
  if (compact) comp.template finish<WAVES>(s_wcnt, &s_bbase, tid, wave);
  __syncthreads();
  flush_counts(s_hist, gh);
}

// The selected key from the last step's state (picking from the last
// pass's histogram when the walk or the passes have not finished), written to
// dst and, when given, to place (a[k], the reference's post-condition,
// RadixSelect.h:149).  One block of 256 threads.
template <typename K>
__global__ __launch_bounds__(256) void k_select_final(const SelState* __restrict__ S_last,
                                                      const uint32_t* __restrict__ ghist_last,
                                                      const unsigned long long* __restrict__ ctr_last, K* dst,
                                                      K* place) {
  __shared__ uint32_t s_hist[256];
  __shared__ SelShared sh;
  SelState t = *S_last;
  if (!t.done) {
    s_hist[threadIdx.x] = hist_read(ghist_last, threadIdx.x);
    __syncthreads();
    const bool compacted = !t.src && t.compact_now;
    pick_state(s_hist, t, compacted, compacted ? *ctr_last : 0, sh);
  }
  if (threadIdx.x == 0) {
    *dst = (K)t.result;
    if (place) *place = (K)t.result;
  }
}

// Scratch: the D per-step states, then the zeroed control words (the top
// digit's histogram from the first read, the per-block minima and maxima,
// one histogram per pass, each histogram in kSelCopies copies, and one
// 64-bit compaction counter per pass), then the compaction buffer.
// Blocks of the first read and the passes: three per CU (the first read's
// LDS allows three), which measured 79 us for the first read of 100M keys
// against 84 us with two.
constexpr unsigned kSelMaxGrid = 768;
constexpr size_t kSelMinMaxWords = 2 * kSelMaxGrid * 2;  // two 64-bit words per block
template <typename K>
static constexpr size_t select_zero_words() {
  return kSelHistWords + kSelMinMaxWords + (sizeof(K) - 1) * kSelHistWords + 2 * (sizeof(K) - 1);
}
// The band's per-block key counts and the sample's per-block top-digit counts
// follow the cleared words (written in full before they are read, never
// cleared).
template <typename K>
static constexpr size_t select_control_words() {
  return select_zero_words<K>() + kSelMaxGrid + kSelPartWords;
}

template <typename K>
static size_t select_bytes(size_t n) {
  Layout L;
  L.take(sizeof(K) * sizeof(SelState));
  L.take(select_control_words<K>() * 4);
  L.take(std::max<size_t>(n / 2, 1) * sizeof(K));
  return L.off;
}

// The launch sequence of one select, all stream-ordered with no host round
// trip: init, the first read, D - 1 filtered passes (each deriving its state
// from the previous launch's counts, returning at once when the answer is
// known), then the final pick, which writes the value to dst (device memory)
// and, when given, to place.
template <typename K>
static void launch_select(const K* din, size_t n, size_t k, SelState* S, uint32_t* z, uint32_t nz, K* buf, size_t cap,
                          K* dst, K* place, hipStream_t s) {
  constexpr int D = sizeof(K);
  uint32_t* ghist_first = z;  // followed by the minimum and maximum words
  uint32_t* ghist = z + kSelHistWords + kSelMinMaxWords;
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(ghist + (D - 1) * kSelHistWords);
  uint32_t* band_counts = reinterpret_cast<uint32_t*>(ctr + (D - 1));
  uint32_t* parts = band_counts + kSelMaxGrid;
  hipLaunchKernelGGL(k_select_init<K>, dim3(kSelInitBlocks), dim3(kSelInitBlock), 0, s, S, din, (uint64_t)k,
                     (uint64_t)n, (uint64_t)cap, (int32_t)(8 * (D - 1)), z, nz, parts);
  const size_t nvec = n / (16 / sizeof(K));
  const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>((nvec + 4 * kSelBlock - 1) / (4 * kSelBlock), kSelMaxGrid));
  {
    ProfScope ps("k_select_first", s);
    hipLaunchKernelGGL(k_select_first<K>, dim3(grid), dim3(kSelBlock), 0, s, din, n, ghist_first,
                       reinterpret_cast<K*>(ghist_first + kSelHistWords), S, buf, band_counts, parts);
  }
  for (int step = 0; step < D - 1; ++step) {
    ProfScope ps("k_select_pass", s);
    hipLaunchKernelGGL(k_select_pass<K>, dim3(grid), dim3(kSelBlock), 0, s, din, n, buf, step, S, ghist_first, ghist,
                       ctr, band_counts);
  }
  hipLaunchKernelGGL(k_select_final<K>, dim3(1), dim3(256), 0, s, S + (D - 1), ghist + (D - 2) * kSelHistWords, ctr + (D - 2),
                     dst, place);
}

template <typename K>
static int select_enqueue(const K* din, size_t n, size_t k, char* b, K* value_out, hipStream_t s) {
  const size_t cap = std::max<size_t>(n / 2, 1);
  const size_t nz = select_zero_words<K>();
  Layout L;
  SelState* S = (SelState*)(b + L.take(sizeof(K) * sizeof(SelState)));
  uint32_t* z = (uint32_t*)(b + L.take(select_control_words<K>() * 4));
  K* buf = (K*)(b + L.take(cap * sizeof(K)));
  launch_select<K>(din, n, k, S, z, (uint32_t)nz, buf, cap, value_out, nullptr, s);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

size_t select_scratch_bytes_u32(size_t n) { return select_bytes<uint32_t>(n); }
int select_enqueue_u32(const uint32_t* din, size_t n, size_t k, char* scratch, uint32_t* value_out, hipStream_t s) {
  return select_enqueue<uint32_t>(din, n, k, scratch, value_out, s);
}

// Scratch layout of one select call: states, control words, the device
// result slot, the compaction buffer and (host keys) the staged input.
struct SelectEntryLayout {
  size_t o_st, o_z, o_res, o_buf, o_in, total;
};
template <typename K>
static SelectEntryLayout select_entry_layout(size_t n, bool dk) {
  const size_t cap = std::max<size_t>(n / 2, 1);
  SelectEntryLayout Y;
  Layout L;
  Y.o_st = L.take(sizeof(K) * sizeof(SelState));
  Y.o_z = L.take(select_control_words<K>() * 4);
  Y.o_res = L.take(sizeof(K));
  Y.o_buf = L.take(cap * sizeof(K));
  Y.o_in = dk ? 0 : L.take(n * sizeof(K));
  Y.total = L.off;
  return Y;
}

size_t select_call_scratch_bytes(bool u64, size_t n) {
  if (n == 0) return 0;
  return u64 ? select_entry_layout<uint64_t>(n, true).total : select_entry_layout<uint32_t>(n, true).total;
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
  const SelectEntryLayout Y = select_entry_layout<K>(n, dk);
  const size_t nz = select_zero_words<K>();
  const size_t o_st = Y.o_st, o_z = Y.o_z, o_res = Y.o_res, o_buf = Y.o_buf, o_in = Y.o_in;
  void* base;
  PA_TRY(ensure_scratch(ctx, Y.total, &base));
  char* b = (char*)base;
  SelState* S = (SelState*)(b + o_st);
  uint32_t* z = (uint32_t*)(b + o_z);
  K* dres_tmp = (K*)(b + o_res);
  K* buf = (K*)(b + o_buf);
  const K* din = keys;
  if (!dk) {
    PA_TRY(stage_to_device(ctx, b + o_in, keys, n * sizeof(K)));
    din = (const K*)(b + o_in);
  }
  launch_select<K>(din, n, k, S, z, (uint32_t)nz, buf, cap, dres ? result : dres_tmp, place && dk ? place : nullptr, s);
  PA_HIP(hipGetLastError());
  if (!dres) PA_HIP(hipMemcpyAsync(result, dres_tmp, sizeof(K), hipMemcpyDeviceToHost, s));
  if (!dres || !dk) PA_HIP(hipStreamSynchronize(s));
  if (place && !dk) {  // host keys: store the value into the caller's array
    if (dres)  // the value went to the caller's device result
      PA_HIP(hipMemcpy(place, result, sizeof(K), hipMemcpyDeviceToHost));
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
