// pa_internal.h -- shared internals of libpa_hip.so (host runtime + device helpers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/pa_hip.h"

namespace pa {

// ---------------------------------------------------------------------------
// Error plumbing
// ---------------------------------------------------------------------------
void set_last_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define PA_HIP(call)                                                         \
  do {                                                                       \
    hipError_t pa_e_ = (call);                                               \
    if (pa_e_ != hipSuccess) return ::pa::hip_fail(pa_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define PA_TRY(call)            \
  do {                          \
    int pa_s_ = (call);         \
    if (pa_s_ != PA_OK) return pa_s_; \
  } while (0)

int arg_error(const char* msg);

// ---------------------------------------------------------------------------
// Per-stream context: a mutex that serialises enqueues on one stream and the
// device scratch owned by that stream.  Work on one stream runs in order, so
// consecutive calls on the same stream can reuse the same scratch.
// ---------------------------------------------------------------------------
struct PinnedSlot {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;  // recorded after the copy that reads the slot
};

struct StreamCtx {
  std::mutex mu;
  hipStream_t stream = nullptr;
  int device = 0;
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  void* pinned = nullptr;  // pinned host staging for small results
  size_t pinned_bytes = 0;
  PinnedSlot ring[4];      // pinned upload buffers reused once their copy is done
  // Device control words (kStreamCtlWords u32), zero when allocated; a kernel
  // that uses them leaves them zero again when it ends, so they never need
  // clearing before a call.  Calls on one stream run in order, so one set per
  // stream suffices.
  uint32_t* ctl = nullptr;
};
constexpr size_t kStreamCtlWords = 8192;
// The stream's control words (allocated and zeroed on first use).
int stream_ctl(StreamCtx& ctx, uint32_t** out);

StreamCtx& stream_ctx(hipStream_t s);
// Ensures ctx.scratch holds >= bytes; on growth waits for the stream first,
// frees the old block and allocates scratch_alloc_bytes(bytes).
int ensure_scratch(StreamCtx& ctx, size_t bytes, void** out);
size_t scratch_alloc_bytes(size_t bytes);
// Device scratch a stream's context holds now; releasing it waits for the
// stream and frees the block (the next call allocates again).
size_t stream_scratch_held(hipStream_t s);
int release_stream_scratch(hipStream_t s);

// Scratch bytes (the ensure_scratch request) of one call of the sort,
// select, histogram and pointer-list segmented sort entry points with device
// keys, device outputs and no caller tmp: what the server admits jobs by.
size_t sort_call_scratch_bytes(bool u64, size_t n);
size_t select_call_scratch_bytes(bool u64, size_t n);
size_t hist_call_scratch_bytes(unsigned bits);
int segmented_lens_scratch_bytes(const uint64_t* lens, size_t num, size_t* bytes);
int ensure_pinned(StreamCtx& ctx, size_t bytes, void** out);
// A pinned upload buffer of >= bytes whose previous copy has finished (waits
// only if all ring slots are still in flight).  After enqueuing the copy that
// reads it, call release_pinned(ctx, slot) to record its completion event.
int acquire_pinned(StreamCtx& ctx, size_t bytes, void** out, int* slot);
int release_pinned(StreamCtx& ctx, int slot);

// Copies between pageable host memory and device memory on ctx's stream
// through the context's pinned ring: chunks are copied into pinned slots by
// several host threads while DMA moves the previous chunk.  to_device returns
// once every chunk is enqueued (the caller's buffer may be reused); to_host
// returns with the data in place.  ctx.mu must be held.
int stage_to_device(StreamCtx& ctx, void* dst, const void* src, size_t bytes);
int stage_to_host(StreamCtx& ctx, void* dst, const void* src, size_t bytes);

// True when p is device-accessible memory (hipMalloc / managed / pinned).
bool is_device_ptr(const void* p);

// Compute units of the current device (queried once per device).
int device_cu_count();

// Bump layout inside a scratch block (256-byte aligned slices).
struct Layout {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = (off + 255) & ~size_t(255);
    off = o + bytes;
    return o;
  }
};

// ---------------------------------------------------------------------------
// Kernel timing hooks (pa_profile_*): records a pair of events around a launch.
// ---------------------------------------------------------------------------
// While a ProfSuppress lives, launches on its thread are not timed (no event
// is created or recorded), for entry points that must not allocate.
struct ProfSuppress {
  ProfSuppress();
  ~ProfSuppress();
  bool prev;
};

// attach = true: the events are not recorded around the launch; the caller
// passes start/stop to hipExtLaunchKernelGGL, which stamps them in the
// dispatch itself (no extra queue packets), and sets `attached` once it did.
struct ProfScope {
  ProfScope(const char* name, hipStream_t s, bool attach = false);
  ~ProfScope();
  const char* name;
  hipStream_t stream;
  hipEvent_t start = nullptr, stop = nullptr;
  bool attach = false;
  bool attached = false;
};

// ---------------------------------------------------------------------------
// Host-staging helper: when a caller passes host arrays, they are copied into
// device scratch, the device routine runs, and results are copied back.
// ---------------------------------------------------------------------------
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Fused 4-digit histogram of u32 keys into zeroed u32 counts[4][256]
// (defined with the sort kernels).
// Radix select of device keys into caller-provided scratch (pa_select.hip).
size_t select_scratch_bytes_u32(size_t n);
int select_enqueue_u32(const uint32_t* din, size_t n, size_t k, char* scratch, uint32_t* value_out, hipStream_t s);
int launch_hist_digits_u32(const uint32_t* keys, size_t n, uint32_t* counts, hipStream_t s);
// Plain sort of one device array of at most small_sort_max_keys() u32 keys in
// one workgroup's LDS, one launch (pa_segsort.hip).
int small_sort_enqueue_u32(uint32_t* keys, size_t n, hipStream_t s);
size_t small_sort_max_keys();

// Tuning state of one target (PA_TUNE_U32, PA_TUNE_U64, PA_TUNE_SEG_U32):
// its candidate shapes by name and the chosen index per size class.
struct TuneTarget {
  const char* key;  // name in tuning caches
  int classes;
  int (*count)();
  std::string (*name)(int index);
  int (*get)(int size_class);              // -1: built-in choice
  void (*set)(int size_class, int index);  // -1: back to the built-in choice
};
const TuneTarget* sort_tune_target(int target);  // PA_TUNE_U32 / PA_TUNE_U64
const TuneTarget* seg_tune_target();
int sort_shape_info(int target, int index, pa_sort_shape* out);
int seg_shape_info(int index, pa_sort_shape* out);
int sort_shape_choice(int target, size_t n);
int seg_shape_choice();
int autotune_sort(int target, size_t max_n, hipStream_t s);
void seg_tune_reset();
// Loads the cache named by PA_TUNE_CACHE once per process (no-op without it).
void tune_cache_autoload();
// Writes the cache named by PA_TUNE_CACHE, if set, after a tuning run.
void tune_cache_autosave();
// Merge of two sorted device runs into a device output (pa_merge.hip), with
// merge_split_bytes(na + nb) bytes of caller scratch for the split table.
size_t merge_split_bytes(size_t n);
template <typename K>
int merge_enqueue(const K* a, size_t na, const K* b, size_t nb, K* out, size_t* split, hipStream_t s);
// Stable merge of key-value runs (ties take the first run), same split table.
template <typename K>
int merge_kv_enqueue(const K* a, const uint32_t* va, size_t na, const K* b, const uint32_t* vb, size_t nb, K* out,
                     uint32_t* vout, size_t* split, hipStream_t s);
// Keys-only u32 sort of any length on device memory (chunks of < 2^30 keys
// merged pairwise); tmp holds n keys, scratch large_sort_u32_scratch_bytes(n).
size_t large_sort_u32_scratch_bytes(size_t n);
int large_sort_u32_enqueue(uint32_t* keys, uint32_t* tmp, size_t n, char* scratch, hipStream_t s);
// Merge sort leaves keys[i * leaf, (i + 1) * leaf) sorted as one segmented
// batch (pa_segsort.hip), tmp as ping-pong space; ctx.mu must be held.
size_t leaves_scratch_bytes_u32(size_t n, size_t leaf);
int leaves_enqueue_u32(uint32_t* keys, uint32_t* tmp, size_t n, size_t leaf, char* scratch, hipStream_t s,
                       StreamCtx& ctx);

}  // namespace pa

// ---------------------------------------------------------------------------
// Device helpers (wave64, gfx950)
// ---------------------------------------------------------------------------
namespace pa {
namespace dev {

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// popcount(mask & lanes below this lane)
__device__ __forceinline__ uint32_t count_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Order-preserving key transforms to unsigned order and their inverses, on the
// bit patterns of K (32 or 64 bits):
//   kXformFloat  IEEE-754: a negative pattern flips every bit, a positive one
//                its sign bit, so -NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN
//   kXformSigned two's complement: flip the sign bit
enum : uint32_t { kXformNone = 0u, kXformFloat = 1u, kXformSigned = 2u };
template <typename K>
__device__ __forceinline__ K key_fwd(K u, uint32_t xform) {
  constexpr K top = K(1) << (sizeof(K) * 8 - 1);
  if (xform == kXformSigned) return u ^ top;
  const K neg = (K)((std::make_signed_t<K>)u >> (sizeof(K) * 8 - 1));  // all ones if the sign is set
  return u ^ (neg | top);
}
template <typename K>
__device__ __forceinline__ K key_inv(K v, uint32_t xform) {
  constexpr K top = K(1) << (sizeof(K) * 8 - 1);
  if (xform == kXformSigned) return v ^ top;
  return v ^ (((v >> (sizeof(K) * 8 - 1)) - K(1)) | top);  // top set: flip sign; clear: flip all
}

template <typename K>
__device__ __forceinline__ uint32_t digit8(K k, uint32_t shift) {
  return (uint32_t)(k >> shift) & 0xFFu;
}
// RB-bit digit at bit `shift` (the last digit of a key may hold fewer bits).
template <typename K, int RB>
__device__ __forceinline__ uint32_t digit_bits(K k, uint32_t shift) {
  return (uint32_t)(k >> shift) & ((1u << RB) - 1u);
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= (uint32_t)off) v += t;
  }
  return v;
}

// Wave-aggregated LDS increment of PER digits held by one lane: when every
// active lane holds PER copies of one digit (constant input, long presorted
// runs), a single lane adds the whole wave's count instead of 64 lanes
// serialising on one LDS address.
template <int PER>
__device__ __forceinline__ void lds_count(uint32_t* __restrict__ bins, const uint32_t (&dg)[PER]) {
  bool same = true;
#pragma unroll
  for (int j = 1; j < PER; ++j) same = same && (dg[j] == dg[0]);
  const uint32_t first = __builtin_amdgcn_readfirstlane(dg[0]);
  const uint64_t active = __ballot(true);
  const uint64_t uni = __ballot(same && dg[0] == first);
  if (uni == active) {
    if (count_below(active) == 0) atomicAdd(&bins[first], (uint32_t)(PER * __popcll(active)));
  } else if (same) {
    atomicAdd(&bins[dg[0]], (uint32_t)PER);
  } else {
#pragma unroll
    for (int j = 0; j < PER; ++j) atomicAdd(&bins[dg[j]], 1u);
  }
}

// Lane-private LDS sub-histograms.  C copies (32 or 16) of the bins; copy
// lane % C of bin b lives in word (b / 2) * C + lane % C with 16-bit counters
// packed two per word (half b % 2) in the packed form, or in word b * C +
// lane % C with 32-bit counters in the wide form.  With C = 32 the 32 lanes
// of a half-wave always address 32 different banks, so an add costs the
// LDS's minimum for a wave whatever the bins are (random keys, one repeated
// key or a presorted run); a single shared histogram serialises on bank
// conflicts instead.  Only lanes l and l + 32 can meet on one word, which the
// atomic resolves.  A packed counter holds at most 65535: one copy of one bin
// may count at most that many keys between flushes.
template <int C, bool PACKED>
struct SubHist {
  static constexpr uint32_t words(uint32_t nbins) { return PACKED ? nbins / 2 * C : nbins * C; }
  __device__ static __forceinline__ void add(uint32_t* h, uint32_t bin, uint32_t lane, uint32_t inc) {
    if constexpr (PACKED)
      atomicAdd(&h[(bin >> 1) * C + lane % C], inc << ((bin & 1) * 16));
    else
      atomicAdd(&h[bin * C + lane % C], inc);
  }
  // Adds PER bins held by one lane; PER equal bins take one add.
  template <int PER>
  __device__ static __forceinline__ void add_all(uint32_t* h, const uint32_t (&bins)[PER], uint32_t lane) {
    bool same = true;
#pragma unroll
    for (int j = 1; j < PER; ++j) same = same && bins[j] == bins[0];
    if (same) {
      add(h, bins[0], lane, PER);
    } else {
#pragma unroll
      for (int j = 0; j < PER; ++j) add(h, bins[j], lane, 1u);
    }
  }
  // Sum over the copies of bin b.  Copy reads start at a bin-dependent offset,
  // so the threads of a wave (consecutive bins) read different banks.
  __device__ static __forceinline__ uint32_t sum(const uint32_t* h, uint32_t b) {
    const uint32_t r = PACKED ? b >> 1 : b;
    const uint32_t sh = PACKED ? (b & 1) * 16 : 0;
    uint32_t s = 0;
#pragma unroll 8
    for (uint32_t c = 0; c < (uint32_t)C; ++c) {
      const uint32_t v = h[r * C + ((c + r) % C)] >> sh;
      s += PACKED ? (v & 0xFFFFu) : v;
    }
    return s;
  }
};

// Buffer load of one 4- or 8-byte key at byte offset off of resource rs.
template <typename K>
__device__ __forceinline__ K buffer_load(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  if constexpr (sizeof(K) == 8) {
    const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
    const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b32(rs, off + 4, 0, 0);
    return (K)(((uint64_t)hi << 32) | lo);
  } else {
    return (K)__builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
  }
}

__device__ __forceinline__ uint32_t atomic_load_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dev
}  // namespace pa
