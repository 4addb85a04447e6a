// pa_hist.hip -- parallel histograms for gfx950.
//
// Reference (DragonSpit/ParallelAlgorithms, HistogramParallel.h):
//   HistogramByteComponentsParallel<256,8>      :249-329  (all 4 byte digits)
//   HistogramOneByteComponentParallel<R,L>      :535-576  (one digit, R bins)
//   HistogramOneByteComponentParallel_2<R,L>    :466-533  (same intent)
//   HistogramByteComponentsQCPar<256,8>         :331-464  (per work quantum)
//   byte-array HistogramOneByteComponentParallel{,_2,_3,_4}<256> :16-247
//   ComputeStartOfBinsPar<256,8>  RadixSortLsdParallel.h:108-165
// The reference bisects the range recursively (parallel_invoke) and sums leaf
// counts on join.  Here each workgroup counts a grid-stride share of the keys
// into LDS-privatised bins (with a wave-uniform fast path so constant or
// presorted runs do not serialise on one LDS address), then adds its bins to
// global u32 counters with one atomic per non-zero bin.  A final kernel widens
// the u32 counters to the uint64 (size_t) counts the reference returns.
#include <algorithm>

#include "pa_internal.h"

namespace pa {

constexpr int kHistBlock = 512;
constexpr unsigned kHistMaxGrid = 512;

using dev::lds_count;

// Histogram of (key >> shift) & mask over u32 or u64 keys, nbins = mask + 1
// <= 4096.
template <typename K>
__global__ __launch_bounds__(kHistBlock) void k_hist_key(const K* __restrict__ keys, size_t n, uint32_t shift,
                                                         uint32_t mask, uint32_t* __restrict__ counts) {
  constexpr int PER = 16 / sizeof(K);
  extern __shared__ uint32_t s_bins[];
  const uint32_t nb = mask + 1;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < nb; i += kHistBlock) s_bins[i] = 0;
  __syncthreads();
  size_t head = ((16 - ((uintptr_t)keys & 15)) & 15) / sizeof(K);
  if (head > n) head = n;
  const size_t nvec = (n - head) / PER;
  const uint4* vk = reinterpret_cast<const uint4*>(keys + head);
  const size_t stride = (size_t)gridDim.x * kHistBlock;
  // Four 16-byte loads per thread in flight before any counting.
  constexpr int U = 4;
  for (size_t v0 = (size_t)blockIdx.x * kHistBlock + tid; v0 < nvec; v0 += stride * U) {
    uint4 x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (v0 + u * stride < nvec) x[u] = vk[v0 + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (v0 + u * stride >= nvec) continue;
      K kk[PER];
      __builtin_memcpy(kk, &x[u], 16);
      uint32_t dg[PER];
#pragma unroll
      for (int j = 0; j < PER; ++j) dg[j] = (uint32_t)(kk[j] >> shift) & mask;
      lds_count<PER>(s_bins, dg);
    }
  }
  if (blockIdx.x == 0) {
    const size_t tail = head + nvec * PER;
    const size_t extra = head + (n - tail);
    for (size_t e = tid; e < extra; e += kHistBlock) {
      size_t idx = e < head ? e : tail + (e - head);
      atomicAdd(&s_bins[(uint32_t)(keys[idx] >> shift) & mask], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < nb; i += kHistBlock) {
    uint32_t c = s_bins[i];
    if (c) atomicAdd(&counts[i], c);
  }
}

// 256-bin histogram of a byte array: 16 bytes per lane per load, four loads
// in flight per thread, counted into lane-private sub-histograms (32 copies
// of 32-bit counters, 32 KB of LDS) so the 16 adds per load never conflict.
__global__ __launch_bounds__(kHistBlock) void k_hist_u8(const uint8_t* __restrict__ data, size_t n,
                                                        uint32_t* __restrict__ counts) {
  using H = dev::SubHist<32, false>;
  __shared__ __align__(16) uint32_t s_h[H::words(256)];
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  for (uint32_t i = tid; i < H::words(256) / 4; i += kHistBlock) reinterpret_cast<uint4*>(s_h)[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  size_t head = (16 - ((uintptr_t)data & 15)) & 15;
  if (head > n) head = n;
  const size_t nvec = (n - head) / 16;
  const uint4* vk = reinterpret_cast<const uint4*>(data + head);
  const size_t stride = (size_t)gridDim.x * kHistBlock;
  constexpr int U = 4;
  for (size_t v0 = (size_t)blockIdx.x * kHistBlock + tid; v0 < nvec; v0 += stride * U) {
    uint4 x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (v0 + u * stride < nvec) x[u] = vk[v0 + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (v0 + u * stride >= nvec) continue;
      const uint32_t w[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
      uint32_t dg[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) dg[j] = (w[j >> 2] >> (8 * (j & 3))) & 0xFFu;
      H::add_all<16>(s_h, dg, lane);
    }
  }
  if (blockIdx.x == 0) {
    const size_t tail = head + nvec * 16;
    const size_t extra = head + (n - tail);
    for (size_t e = tid; e < extra; e += kHistBlock) {
      size_t idx = e < head ? e : tail + (e - head);
      H::add(s_h, data[idx], lane, 1u);
    }
  }
  __syncthreads();
  if (tid < 256) {
    const uint32_t c = H::sum(s_h, tid);
    if (c) atomicAdd(&counts[tid], c);
  }
}

// k_hist_u8 in one launch: every block adds its 256 counts into copy
// b % kHistCopies of the counts in the stream's control words (ctl[c * 256 +
// bin]) and takes a ticket from ctl[kHistCopies * 256]; the block that takes
// the last ticket reads and clears the copies (atomic exchange), writes their
// sums widened to out and resets the ticket, so the words are zero again for
// the next call and no clearing or widening launch is needed.  The blocks end
// together, and adds to one address are applied one after another: with one
// copy their flushes took about 4 us of the 26 us a call on 100M bytes.
constexpr uint32_t kHistCopies = 16;
static_assert(kHistCopies * 256 + 1 <= kStreamCtlWords, "the copies and the ticket fit the control words");
__global__ __launch_bounds__(kHistBlock) void k_hist_u8_once(const uint8_t* __restrict__ data, size_t n,
                                                             uint32_t* __restrict__ ctl, uint64_t* __restrict__ out) {
  using H = dev::SubHist<32, false>;
  __shared__ __align__(16) uint32_t s_h[H::words(256)];
  __shared__ bool s_last;
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  for (uint32_t i = tid; i < H::words(256) / 4; i += kHistBlock) reinterpret_cast<uint4*>(s_h)[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  size_t head = (16 - ((uintptr_t)data & 15)) & 15;
  if (head > n) head = n;
  const size_t nvec = (n - head) / 16;
  const uint4* vk = reinterpret_cast<const uint4*>(data + head);
  const size_t stride = (size_t)gridDim.x * kHistBlock;
  constexpr int U = 4;
  for (size_t v0 = (size_t)blockIdx.x * kHistBlock + tid; v0 < nvec; v0 += stride * U) {
    uint4 x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (v0 + u * stride < nvec) x[u] = vk[v0 + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (v0 + u * stride >= nvec) continue;
      const uint32_t w[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
      uint32_t dg[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) dg[j] = (w[j >> 2] >> (8 * (j & 3))) & 0xFFu;
      H::add_all<16>(s_h, dg, lane);
    }
  }
  if (blockIdx.x == 0) {
    const size_t tail = head + nvec * 16;
    const size_t extra = head + (n - tail);
    for (size_t e = tid; e < extra; e += kHistBlock) {
      size_t idx = e < head ? e : tail + (e - head);
      H::add(s_h, data[idx], lane, 1u);
    }
  }
  __syncthreads();
  if (tid < 256) {
    const uint32_t c = H::sum(s_h, tid);
    if (c) {
      // Device-scope atomics are applied where every CU sees them; waiting
      // for the returned value means this one has been applied before the
      // barrier and the ticket below (no fence, which would write back L2).
      const uint32_t old = atomicAdd(&ctl[(blockIdx.x % kHistCopies) * 256 + tid], c);
      asm volatile("" ::"v"(old));
    }
  }
  __syncthreads();
  if (tid == 0) s_last = atomicAdd(&ctl[kHistCopies * 256], 1u) == gridDim.x - 1;
  __syncthreads();
  if (s_last) {
    if (tid < 256) {
      uint32_t c[kHistCopies];
#pragma unroll
      for (uint32_t i = 0; i < kHistCopies; ++i) c[i] = atomicExch(&ctl[i * 256 + tid], 0u);
      uint64_t t = 0;
#pragma unroll
      for (uint32_t i = 0; i < kHistCopies; ++i) t += c[i];
      out[tid] = t;
    }
    if (tid == 0) atomicExch(&ctl[kHistCopies * 256], 0u);
  }
}

// Counting sort of bytes, the fill half (CountingSortParallel.h:30-76: after
// the histogram, bin b's value fills [start[b], start[b] + count[b])).  Every
// block scans the 256 counts in LDS, then each thread writes whole 16-byte
// words: the bin of a word's first byte by binary search over the bin starts,
// later bytes by stepping to the next non-empty bin.  Pure write traffic.
__global__ __launch_bounds__(kHistBlock) void k_counting_fill(uint8_t* __restrict__ a, size_t n,
                                                              const uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_end[256];  // s_end[b] = first index past bin b
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ uint32_t s_w[4];
  if (tid < 256) {
    const uint32_t c = counts[tid];
    const uint32_t inc = dev::wave_incl_scan(c, lane);
    if (lane == 63) s_w[wave] = inc;
    s_end[tid] = inc;  // wave-local for now
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t pre = 0;
    for (uint32_t w = 0; w < wave; ++w) pre += s_w[w];
    s_end[tid] += pre;
  }
  __syncthreads();
  // bin of index i: the first b with s_end[b] > i
  auto bin_of = [&](size_t i) {
    uint32_t lo = 0, hi = 255;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_end[mid] > i) hi = mid; else lo = mid + 1;
    }
    return lo;
  };
  const size_t head = std::min<size_t>(n, (16 - ((uintptr_t)a & 15)) & 15);
  const size_t nvec = (n - head) / 16;
  uint4* v = reinterpret_cast<uint4*>(a + head);
  const size_t stride = (size_t)gridDim.x * kHistBlock;
  for (size_t w = (size_t)blockIdx.x * kHistBlock + tid; w < nvec; w += stride) {
    const size_t i0 = head + w * 16;
    uint32_t b = bin_of(i0);
    if (s_end[b] > i0 + 15) {  // the usual case: all 16 bytes in one bin
      const uint32_t r = b * 0x01010101u;
      v[w] = make_uint4(r, r, r, r);
      continue;
    }
    uint32_t word[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      while (s_end[b] <= i0 + j) ++b;  // i0 + j < n = s_end[255], so b stays below 256
      word[j >> 2] |= b << (8 * (j & 3));
    }
    v[w] = make_uint4(word[0], word[1], word[2], word[3]);
  }
  if (blockIdx.x == 0) {
    const size_t tail = head + nvec * 16;
    for (size_t i = tid; i < head; i += kHistBlock) a[i] = (uint8_t)bin_of(i);
    for (size_t i = tail + tid; i < n; i += kHistBlock) a[i] = (uint8_t)bin_of(i);
  }
}

// Per-work-quantum counts of one byte over keys[first, last): q = index /
// quantum (absolute index).  Each block counts one chunk that lies inside a
// single quantum, so its LDS bins go to one output row.
constexpr uint32_t kQChunk = 8192;
__global__ __launch_bounds__(kHistBlock) void k_hist_quanta(const uint32_t* __restrict__ keys, size_t first,
                                                            size_t last, size_t quantum, size_t chunks_per_q,
                                                            size_t q0, uint32_t shift, uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_bins[256];
  const uint32_t tid = threadIdx.x;
  if (tid < 256) s_bins[tid] = 0;
  __syncthreads();
  const size_t q = q0 + blockIdx.x / chunks_per_q;
  const size_t c = blockIdx.x % chunks_per_q;
  size_t lo = q * quantum + c * kQChunk;
  size_t hi = std::min(std::min(lo + kQChunk, (q + 1) * quantum), last);
  lo = std::max(lo, first);
  for (size_t i = lo + tid; i < hi; i += kHistBlock) atomicAdd(&s_bins[(keys[i] >> shift) & 0xFFu], 1u);
  __syncthreads();
  if (tid < 256 && s_bins[tid]) atomicAdd(&counts[(q - q0) * 256 + tid], s_bins[tid]);
}

// Small quanta (< one chunk): one global atomic per element.
__global__ __launch_bounds__(256) void k_hist_quanta_small(const uint32_t* __restrict__ keys, size_t first,
                                                           size_t last, size_t quantum, size_t q0, uint32_t shift,
                                                           uint32_t* __restrict__ counts) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = first + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < last; i += stride)
    atomicAdd(&counts[(i / quantum - q0) * 256 + ((keys[i] >> shift) & 0xFFu)], 1u);
}

__global__ __launch_bounds__(256) void k_widen(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// start[q][b] = sum_{b'<b} total[b'] + sum_{q'<q} count[q'][b]: one thread per
// bin walks the quanta (bin-major exclusive scan, RadixSortLsdParallel.h:129-156).
__global__ __launch_bounds__(256) void k_start_of_bins(const uint32_t* __restrict__ count, size_t nq,
                                                       uint64_t* __restrict__ start) {
  __shared__ uint64_t s_tot[256];
  __shared__ uint64_t s_wsum[4];
  const uint32_t b = threadIdx.x, lane = b & 63, wave = b >> 6;
  uint64_t tot = 0;
  for (size_t q = 0; q < nq; ++q) tot += count[q * 256 + b];
  s_tot[b] = tot;
  // exclusive scan over bins of tot (64-bit, 4 waves)
  uint64_t v = tot;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint64_t t = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += t;
  }
  if (lane == 63) s_wsum[wave] = v;
  __syncthreads();
  uint64_t pre = 0;
  for (uint32_t w = 0; w < wave; ++w) pre += s_wsum[w];
  uint64_t run = pre + v - tot;
  for (size_t q = 0; q < nq; ++q) {
    start[q * 256 + b] = run;
    run += count[q * 256 + b];
  }
}

// Copies a device uint64 result array to the caller's pointer.
static int deliver_u64(const uint64_t* dsrc, uint64_t* out, size_t count, bool out_dev, hipStream_t s) {
  if (out_dev) return PA_OK;  // already written in place
  PA_HIP(hipMemcpyAsync(out, dsrc, count * 8, hipMemcpyDeviceToHost, s));
  PA_HIP(hipStreamSynchronize(s));
  return PA_OK;
}

// Stages a possibly-host input array of `bytes` into scratch at offset `off`.
static int stage_in(const void* p, size_t bytes, char* scratch_slot, bool is_dev, hipStream_t s, const void** out) {
  if (is_dev) {
    *out = p;
    return PA_OK;
  }
  PA_TRY(stage_to_device(stream_ctx(s), scratch_slot, p, bytes));
  *out = scratch_slot;
  return PA_OK;
}

static unsigned hist_grid(size_t nvec) {
  return (unsigned)std::max<size_t>(1, std::min<size_t>((nvec + kHistBlock - 1) / kHistBlock, kHistMaxGrid));
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_hist_digits_u32(const uint32_t* keys, size_t n, uint64_t* counts, pa_stream_t stream) {
  if (!counts || (n && !keys)) return arg_error("pa_hist_digits_u32: null pointer");
  if (n > 0xFFFFFFFFull) return arg_error("pa_hist_digits_u32: n must be < 2^32");
  hipStream_t s = (hipStream_t)stream;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = n == 0 || is_device_ptr(keys);
  const bool dout = is_device_ptr(counts);
  Layout L;
  size_t o_c32 = L.take(4 * 256 * 4), o_c64 = L.take(4 * 256 * 8), o_in = dk ? 0 : L.take(n * 4);
  void* base;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* b = (char*)base;
  uint32_t* c32 = (uint32_t*)(b + o_c32);
  uint64_t* c64 = dout ? counts : (uint64_t*)(b + o_c64);
  const void* din;
  PA_TRY(stage_in(keys, n * 4, b + o_in, dk, s, &din));
  {
    ProfScope ps("k_zero_u32", s);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, c32, (size_t)1024);
  }
  if (n) PA_TRY(launch_hist_digits_u32((const uint32_t*)din, n, c32, s));
  hipLaunchKernelGGL(k_widen, dim3(4), dim3(256), 0, s, c32, c64, (size_t)1024);
  PA_HIP(hipGetLastError());
  return deliver_u64(c64, counts, 1024, dout, s);
}

}  // extern "C"

namespace pa {

// Scratch layout of one pa_hist_u32 / pa_hist_u64 call: u32 and u64 counts,
// then (host keys) the staged input.
struct HistEntryLayout {
  size_t o_c32, o_c64, o_in, total;
};
static HistEntryLayout hist_entry_layout(uint32_t nb, size_t in_bytes) {
  HistEntryLayout Y;
  Layout L;
  Y.o_c32 = L.take(nb * 4);
  Y.o_c64 = L.take(nb * 8);
  Y.o_in = in_bytes ? L.take(in_bytes) : 0;
  Y.total = L.off;
  return Y;
}

size_t hist_call_scratch_bytes(unsigned bits) {
  return hist_entry_layout(1u << (bits > 12 ? 12 : bits), 0).total;
}

template <typename K>
static int hist_key_entry(const K* keys, size_t n, unsigned shift, unsigned bits, uint64_t* counts, hipStream_t s,
                          const char* who) {
  if (!counts || (n && !keys)) return arg_error((std::string(who) + ": null pointer").c_str());
  if (bits < 1 || bits > 12 || shift + 1 > 8 * sizeof(K))
    return arg_error((std::string(who) + ": need 1 <= bits <= 12 and shift < key bits").c_str());
  if (n > 0xFFFFFFFFull) return arg_error((std::string(who) + ": n must be < 2^32").c_str());
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const uint32_t nb = 1u << bits;
  const bool dk = n == 0 || is_device_ptr(keys);
  const bool dout = is_device_ptr(counts);
  const HistEntryLayout Y = hist_entry_layout(nb, dk ? 0 : n * sizeof(K));
  void* base;
  PA_TRY(ensure_scratch(ctx, Y.total, &base));
  char* b = (char*)base;
  uint32_t* c32 = (uint32_t*)(b + Y.o_c32);
  uint64_t* c64 = dout ? counts : (uint64_t*)(b + Y.o_c64);
  const void* din;
  PA_TRY(stage_in(keys, n * sizeof(K), b + Y.o_in, dk, s, &din));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, c32, (size_t)nb);
  if (n) {
    ProfScope ps(sizeof(K) == 4 ? "k_hist_u32" : "k_hist_u64", s);
    hipLaunchKernelGGL(k_hist_key<K>, dim3(hist_grid(n / (16 / sizeof(K)) + 1)), dim3(kHistBlock), nb * 4, s,
                       (const K*)din, n, (uint32_t)shift, nb - 1, c32);
  }
  hipLaunchKernelGGL(k_widen, dim3(std::max<unsigned>(1, nb / 256)), dim3(256), 0, s, c32, c64, (size_t)nb);
  PA_HIP(hipGetLastError());
  return deliver_u64(c64, counts, nb, dout, s);
}

}  // namespace pa

extern "C" {

int pa_hist_u32(const uint32_t* keys, size_t n, unsigned shift, unsigned bits, uint64_t* counts, pa_stream_t stream) {
  return hist_key_entry<uint32_t>(keys, n, shift, bits, counts, (hipStream_t)stream, "pa_hist_u32");
}

int pa_hist_u64(const uint64_t* keys, size_t n, unsigned shift, unsigned bits, uint64_t* counts, pa_stream_t stream) {
  return hist_key_entry<uint64_t>(keys, n, shift, bits, counts, (hipStream_t)stream, "pa_hist_u64");
}

int pa_hist_u8(const uint8_t* bytes, size_t n, uint64_t* counts, pa_stream_t stream) {
  if (!counts || (n && !bytes)) return arg_error("pa_hist_u8: null pointer");
  if (n > 0xFFFFFFFFull) return arg_error("pa_hist_u8: n must be < 2^32");
  hipStream_t s = (hipStream_t)stream;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = n == 0 || is_device_ptr(bytes);
  const bool dout = is_device_ptr(counts);
  Layout L;
  size_t o_c64 = L.take(256 * 8), o_in = dk ? 0 : L.take(n);
  void* base;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* b = (char*)base;
  uint64_t* c64 = dout ? counts : (uint64_t*)(b + o_c64);
  uint32_t* ctl = nullptr;
  PA_TRY(stream_ctl(ctx, &ctl));
  const void* din;
  PA_TRY(stage_in(bytes, n, b + o_in, dk, s, &din));
  {
    // n = 0 still runs one block, which writes the zero counts.
    ProfScope ps("k_hist_u8", s);
    hipLaunchKernelGGL(k_hist_u8_once, dim3(n ? hist_grid(n / 16 + 1) : 1), dim3(kHistBlock), 0, s,
                       (const uint8_t*)din, n, ctl, c64);
  }
  PA_HIP(hipGetLastError());
  return deliver_u64(c64, counts, 256, dout, s);
}

int pa_counting_sort_u8(uint8_t* bytes, size_t n, pa_stream_t stream) {
  if (n == 0) return PA_OK;
  if (!bytes) return arg_error("pa_counting_sort_u8: null pointer");
  if (n > 0xFFFFFFFFull) return arg_error("pa_counting_sort_u8: n must be < 2^32");
  hipStream_t s = (hipStream_t)stream;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = is_device_ptr(bytes);
  Layout L;
  const size_t o_c32 = L.take(256 * 4), o_in = dk ? 0 : L.take(n + 16);
  void* base;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* b = (char*)base;
  uint32_t* c32 = (uint32_t*)(b + o_c32);
  // A host array is staged with the same 16-byte phase, so the device copy is
  // laid out like the caller's.
  uint8_t* d = dk ? bytes : (uint8_t*)(b + o_in + ((uintptr_t)bytes & 15));
  if (!dk) PA_TRY(stage_to_device(ctx, d, bytes, n));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, c32, (size_t)256);
  {
    ProfScope ps("k_hist_u8", s);
    hipLaunchKernelGGL(k_hist_u8, dim3(hist_grid(n / 16 + 1)), dim3(kHistBlock), 0, s, (const uint8_t*)d, n, c32);
  }
  {
    ProfScope ps("k_counting_fill", s);
    hipLaunchKernelGGL(k_counting_fill, dim3(hist_grid(n / 16 + 1)), dim3(kHistBlock), 0, s, d, n, c32);
  }
  PA_HIP(hipGetLastError());
  if (!dk) PA_TRY(stage_to_host(ctx, bytes, d, n));
  return PA_OK;
}

// Shared by pa_hist_quanta_u32 and pa_start_of_bins_u32: u32 counts of byte
// `which_byte` for quanta rows [0, nq) over keys[first, last).
static int quanta_counts(const uint32_t* din, size_t first, size_t last, size_t quantum, size_t nq, unsigned which_byte,
                         uint32_t* c32, hipStream_t s) {
  const size_t total = nq * 256;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)std::min<size_t>((total + 255) / 256, 4096)), dim3(256), 0, s, c32,
                     total);
  if (last > first) {
    const size_t qa = first / quantum, qb = (last - 1) / quantum;
    if (quantum >= kQChunk) {
      const size_t cpq = (quantum + kQChunk - 1) / kQChunk;
      const size_t blocks = (qb - qa + 1) * cpq;
      ProfScope ps("k_hist_quanta", s);
      for (size_t done = 0; done < blocks; done += 1u << 30) {  // grid x limit
        size_t nb = std::min<size_t>(blocks - done, 1u << 30);
        hipLaunchKernelGGL(k_hist_quanta, dim3((unsigned)nb), dim3(kHistBlock), 0, s, din, first, last, quantum, cpq,
                           qa + done / cpq, 8u * which_byte, c32 + (qa + done / cpq) * 256);
      }
    } else {
      ProfScope ps("k_hist_quanta_small", s);
      unsigned grid = (unsigned)std::min<size_t>((last - first + 255) / 256, 8192);
      hipLaunchKernelGGL(k_hist_quanta_small, dim3(grid), dim3(256), 0, s, din, first, last, quantum, (size_t)0,
                         8u * which_byte, c32);
    }
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}

int pa_hist_quanta_u32(const uint32_t* keys, size_t first, size_t last, size_t quantum, size_t nq, unsigned which_byte,
                       uint64_t* counts, pa_stream_t stream) {
  if (!counts || (last > first && !keys)) return arg_error("pa_hist_quanta_u32: null pointer");
  if (quantum == 0 || which_byte > 3) return arg_error("pa_hist_quanta_u32: quantum must be > 0, which_byte <= 3");
  if (last > first && (last - 1) / quantum >= nq) return arg_error("pa_hist_quanta_u32: nq too small for the range");
  if (last > 0xFFFFFFFFull) return arg_error("pa_hist_quanta_u32: indices must be < 2^32");
  if (nq == 0) return PA_OK;
  hipStream_t s = (hipStream_t)stream;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = last <= first || is_device_ptr(keys);
  const bool dout = is_device_ptr(counts);
  const size_t total = nq * 256;
  Layout L;
  size_t o_c32 = L.take(total * 4), o_c64 = dout ? 0 : L.take(total * 8), o_in = dk ? 0 : L.take(last * 4);
  void* base;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* b = (char*)base;
  uint32_t* c32 = (uint32_t*)(b + o_c32);
  uint64_t* c64 = dout ? counts : (uint64_t*)(b + o_c64);
  const void* din = keys;
  if (!dk) {
    PA_TRY(stage_to_device(ctx, b + o_in + first * 4, keys + first, (last - first) * 4));
    din = b + o_in;
  }
  PA_TRY(quanta_counts((const uint32_t*)din, first, last, quantum, nq, which_byte, c32, s));
  hipLaunchKernelGGL(k_widen, dim3((unsigned)std::min<size_t>((total + 255) / 256, 4096)), dim3(256), 0, s, c32, c64,
                     total);
  PA_HIP(hipGetLastError());
  return deliver_u64(c64, counts, total, dout, s);
}

int pa_start_of_bins_u32(const uint32_t* keys, size_t n, size_t quantum, size_t nq, unsigned digit, uint64_t* start,
                         pa_stream_t stream) {
  if (!start || (n && !keys)) return arg_error("pa_start_of_bins_u32: null pointer");
  if (quantum == 0 || digit > 3) return arg_error("pa_start_of_bins_u32: quantum must be > 0, digit <= 3");
  if (n && (n - 1) / quantum >= nq) return arg_error("pa_start_of_bins_u32: nq too small for n");
  if (n > 0xFFFFFFFFull) return arg_error("pa_start_of_bins_u32: n must be < 2^32");
  if (nq == 0) return PA_OK;
  hipStream_t s = (hipStream_t)stream;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = n == 0 || is_device_ptr(keys);
  const bool dout = is_device_ptr(start);
  const size_t total = nq * 256;
  Layout L;
  size_t o_c32 = L.take(total * 4), o_out = dout ? 0 : L.take(total * 8), o_in = dk ? 0 : L.take(n * 4);
  void* base;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* b = (char*)base;
  uint32_t* c32 = (uint32_t*)(b + o_c32);
  uint64_t* dst = dout ? start : (uint64_t*)(b + o_out);
  const void* din;
  PA_TRY(stage_in(keys, n * 4, b + o_in, dk, s, &din));
  PA_TRY(quanta_counts((const uint32_t*)din, 0, n, quantum, nq, digit, c32, s));
  {
    ProfScope ps("k_start_of_bins", s);
    hipLaunchKernelGGL(k_start_of_bins, dim3(1), dim3(256), 0, s, c32, nq, dst);
  }
  PA_HIP(hipGetLastError());
  return deliver_u64(dst, start, total, dout, s);
}

}  // extern "C"
