// pa_runtime.hip -- host runtime of libpa_hip.so: errors, per-stream scratch,
// pointer classification, kernel timing hooks, copy kernel, generators.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <thread>
#include <vector>

#include "pa_internal.h"

namespace pa {

static thread_local std::string t_last_error;

void set_last_error(const std::string& msg) { t_last_error = msg; }

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  set_last_error(buf);
  (void)hipGetLastError();
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PA_ERR_OOM : PA_ERR_HIP;
}

int arg_error(const char* msg) {
  set_last_error(msg);
  return PA_ERR_ARG;
}

// ---------------------------------------------------------------------------
// Per-thread streams
// ---------------------------------------------------------------------------
// Each host thread gets its own stream per device on first use, so callers
// on different threads neither share the null stream's serialisation nor its
// scratch.  A thread's streams go back to a pool when the thread ends (no HIP
// call there) and are handed to the next thread that asks, so a process
// holds at most as many as it has threads using the library at once.
struct StreamPool {
  std::mutex mu;
  std::map<int, std::vector<hipStream_t>> free_by_device;
};
static StreamPool& stream_pool() {
  static StreamPool* p = new StreamPool();  // never destroyed: used from thread exits
  return *p;
}
struct ThreadStreams {
  std::vector<std::pair<int, hipStream_t>> held;
  ~ThreadStreams() {
    StreamPool& P = stream_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    for (auto& d : held) P.free_by_device[d.first].push_back(d.second);
  }
};
static thread_local ThreadStreams t_streams;

hipStream_t thread_stream() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  for (auto& d : t_streams.held)
    if (d.first == dev) return d.second;
  hipStream_t s = nullptr;
  {
    StreamPool& P = stream_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto& f = P.free_by_device[dev];
    if (!f.empty()) {
      s = f.back();
      f.pop_back();
    }
  }
  // Default flags: the stream is ordered after work the caller put on the
  // null stream, as the null stream itself would be.
  if (!s && hipStreamCreate(&s) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;  // the null stream
  }
  t_streams.held.emplace_back(dev, s);
  return s;
}

// ---------------------------------------------------------------------------
// Stream contexts
// ---------------------------------------------------------------------------
static std::mutex g_ctx_mu;
static std::map<std::pair<int, hipStream_t>, StreamCtx*> g_ctx;

StreamCtx& stream_ctx(hipStream_t s) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  auto key = std::make_pair(dev, s);
  auto it = g_ctx.find(key);
  if (it != g_ctx.end()) return *it->second;
  StreamCtx* c = new StreamCtx();  // lives until process exit
  c->stream = s;
  c->device = dev;
  g_ctx[key] = c;
  return *c;
}

size_t scratch_alloc_bytes(size_t bytes) {
  if (bytes == 0) bytes = 256;
  return align_up(bytes + bytes / 8, size_t(1) << 20);
}

int ensure_scratch(StreamCtx& ctx, size_t bytes, void** out) {
  if (bytes == 0) bytes = 256;
  if (ctx.scratch_bytes < bytes) {
    if (ctx.scratch) {
      PA_HIP(hipStreamSynchronize(ctx.stream));
      PA_HIP(hipFree(ctx.scratch));
      ctx.scratch = nullptr;
      ctx.scratch_bytes = 0;
    }
    const size_t want = scratch_alloc_bytes(bytes);
    PA_HIP(hipMalloc(&ctx.scratch, want));
    ctx.scratch_bytes = want;
  }
  *out = ctx.scratch;
  return PA_OK;
}

int stream_ctl(StreamCtx& ctx, uint32_t** out) {
  if (!ctx.ctl) {
    uint32_t* p = nullptr;
    PA_HIP(hipMalloc(&p, kStreamCtlWords * 4));
    // Once per stream, before any kernel uses the words: a plain clear,
    // waited for (the words then stay zero from call to call).
    if (hipMemset(p, 0, kStreamCtlWords * 4) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(p);
      return hip_fail(hipErrorUnknown, "stream control words", __FILE__, __LINE__);
    }
    ctx.ctl = p;
  }
  *out = ctx.ctl;
  return PA_OK;
}

size_t stream_scratch_held(hipStream_t s) {
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  return ctx.scratch_bytes;
}

int release_stream_scratch(hipStream_t s) {
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  if (ctx.scratch) {
    PA_HIP(hipStreamSynchronize(ctx.stream));
    PA_HIP(hipFree(ctx.scratch));
    ctx.scratch = nullptr;
    ctx.scratch_bytes = 0;
  }
  return PA_OK;
}

int ensure_pinned(StreamCtx& ctx, size_t bytes, void** out) {
  if (ctx.pinned_bytes < bytes) {
    if (ctx.pinned) {
      PA_HIP(hipStreamSynchronize(ctx.stream));
      PA_HIP(hipHostFree(ctx.pinned));
      ctx.pinned = nullptr;
      ctx.pinned_bytes = 0;
    }
    size_t want = align_up(bytes < 65536 ? 65536 : bytes, 4096);
    PA_HIP(hipHostMalloc(&ctx.pinned, want, hipHostMallocDefault));
    ctx.pinned_bytes = want;
  }
  *out = ctx.pinned;
  return PA_OK;
}

int acquire_pinned(StreamCtx& ctx, size_t bytes, void** out, int* slot) {
  const int nslots = (int)(sizeof(ctx.ring) / sizeof(ctx.ring[0]));
  int pick = -1;
  for (int i = 0; i < nslots && pick < 0; ++i) {
    PinnedSlot& r = ctx.ring[i];
    if (!r.done || hipEventQuery(r.done) == hipSuccess) pick = i;
  }
  (void)hipGetLastError();  // hipEventQuery reports hipErrorNotReady for busy slots
  if (pick < 0) {           // every slot in flight: wait for the oldest use
    pick = 0;
    PA_HIP(hipEventSynchronize(ctx.ring[0].done));
  }
  PinnedSlot& r = ctx.ring[pick];
  if (!r.done) PA_HIP(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
  if (r.bytes < bytes) {
    if (r.ptr) PA_HIP(hipHostFree(r.ptr));
    r.ptr = nullptr;
    r.bytes = 0;
    const size_t want = align_up(bytes < 65536 ? 65536 : bytes, 4096);
    PA_HIP(hipHostMalloc(&r.ptr, want, hipHostMallocDefault));
    r.bytes = want;
  }
  *out = r.ptr;
  *slot = pick;
  return PA_OK;
}

int release_pinned(StreamCtx& ctx, int slot) {
  PA_HIP(hipEventRecord(ctx.ring[slot].done, ctx.stream));
  return PA_OK;
}

// ---------------------------------------------------------------------------
// Pipelined staging of pageable host memory
// ---------------------------------------------------------------------------
constexpr size_t kStageChunk = size_t(64) << 20;   // bytes per pinned slot use
constexpr size_t kStageDirect = size_t(16) << 20;  // below this, one plain copy
constexpr int kStageThreads = 8;

// memcpy split over host threads (large pieces only).
static void par_memcpy(void* dst, const void* src, size_t bytes) {
  if (bytes < (size_t(4) << 20)) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t per = ((bytes + kStageThreads - 1) / kStageThreads + 63) & ~size_t(63);
  std::vector<std::thread> th;
  th.reserve(kStageThreads);
  size_t done = 0;  // bytes from 0 handed to threads; the rest is copied here
  try {
    for (int t = 0; t < kStageThreads; ++t) {
      const size_t lo = std::min(bytes, t * per), hi = std::min(bytes, lo + per);
      if (hi > lo) th.emplace_back([=] { std::memcpy((char*)dst + lo, (const char*)src + lo, hi - lo); });
      done = hi;
    }
  } catch (...) {  // no thread could be started: this thread copies the rest
  }
  if (done < bytes) std::memcpy((char*)dst + done, (const char*)src + done, bytes - done);
  for (auto& x : th) x.join();
}

// Slot k of the ring with at least `bytes` of pinned memory and its event,
// after its previous copy has finished.
static int ring_slot(StreamCtx& ctx, int k, size_t bytes, void** out) {
  PinnedSlot& r = ctx.ring[k];
  if (r.done) PA_HIP(hipEventSynchronize(r.done));
  else PA_HIP(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
  if (r.bytes < bytes) {
    if (r.ptr) PA_HIP(hipHostFree(r.ptr));
    r.ptr = nullptr;
    r.bytes = 0;
    PA_HIP(hipHostMalloc(&r.ptr, align_up(bytes, 4096), hipHostMallocDefault));
    r.bytes = align_up(bytes, 4096);
  }
  *out = r.ptr;
  return PA_OK;
}

int stage_to_device(StreamCtx& ctx, void* dst, const void* src, size_t bytes) {
  hipStream_t s = ctx.stream;
  if (bytes < kStageDirect) {
    PA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return PA_OK;
  }
  const int nslots = (int)(sizeof(ctx.ring) / sizeof(ctx.ring[0]));
  int k = 0;
  for (size_t off = 0; off < bytes; off += kStageChunk, k = (k + 1) % nslots) {
    const size_t len = std::min(kStageChunk, bytes - off);
    void* p;
    PA_TRY(ring_slot(ctx, k, kStageChunk, &p));  // waits for this slot's last DMA
    par_memcpy(p, (const char*)src + off, len);
    PA_HIP(hipMemcpyAsync((char*)dst + off, p, len, hipMemcpyHostToDevice, s));
    PA_HIP(hipEventRecord(ctx.ring[k].done, s));
  }
  return PA_OK;
}

int stage_to_host(StreamCtx& ctx, void* dst, const void* src, size_t bytes) {
  hipStream_t s = ctx.stream;
  if (bytes < kStageDirect) {
    PA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    PA_HIP(hipStreamSynchronize(s));
    return PA_OK;
  }
  const int nslots = (int)(sizeof(ctx.ring) / sizeof(ctx.ring[0]));
  const size_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
  void* p[sizeof(ctx.ring) / sizeof(ctx.ring[0])];
  // Chunk i uses slot i % nslots; up to nslots DMA copies run ahead of the
  // host copies out of the pinned slots.
  auto issue = [&](size_t i) -> int {
    const int k = (int)(i % nslots);
    const size_t off = i * kStageChunk, len = std::min(kStageChunk, bytes - off);
    PA_HIP(hipMemcpyAsync(p[k], (const char*)src + off, len, hipMemcpyDeviceToHost, s));
    PA_HIP(hipEventRecord(ctx.ring[k].done, s));
    return PA_OK;
  };
  for (size_t i = 0; i < nchunks && i < (size_t)nslots; ++i) {
    PA_TRY(ring_slot(ctx, (int)i, kStageChunk, &p[i]));
    PA_TRY(issue(i));
  }
  for (size_t i = 0; i < nchunks; ++i) {
    const int k = (int)(i % nslots);
    const size_t off = i * kStageChunk, len = std::min(kStageChunk, bytes - off);
    PA_HIP(hipEventSynchronize(ctx.ring[k].done));
    par_memcpy((char*)dst + off, p[k], len);
    if (i + nslots < nchunks) PA_TRY(issue(i + nslots));
  }
  return PA_OK;
}

int device_cu_count() {
  constexpr int kMaxDevices = 64;
  static std::atomic<int> cached[kMaxDevices];
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev < 0 || dev >= kMaxDevices) dev = 0;
  int c = cached[dev].load(std::memory_order_relaxed);
  if (c <= 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) {
      (void)hipGetLastError();
      c = 256;
    }
    cached[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

bool is_device_ptr(const void* p) {
  if (p == nullptr) return false;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// ---------------------------------------------------------------------------
// Kernel timing
// ---------------------------------------------------------------------------
struct ProfRec {
  std::string name;
  hipEvent_t start, stop;
};
static std::atomic<bool> g_prof_on{false};
static std::mutex g_prof_mu;
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_prof_pool;  // events of reset records, reused
static std::string g_prof_only;               // time only names with this prefix ("" = all)

// An event from the pool, or a new one; nullptr on failure.
static hipEvent_t prof_event() {
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_pool.empty()) {
      hipEvent_t e = g_prof_pool.back();
      g_prof_pool.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

static thread_local bool g_prof_suppressed = false;
ProfSuppress::ProfSuppress() : prev(g_prof_suppressed) { g_prof_suppressed = true; }
ProfSuppress::~ProfSuppress() { g_prof_suppressed = prev; }

ProfScope::ProfScope(const char* n, hipStream_t s, bool att) : name(n), stream(s), attach(att) {
  if (g_prof_suppressed || !g_prof_on.load(std::memory_order_relaxed)) return;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_only.empty() && std::strncmp(n, g_prof_only.c_str(), g_prof_only.size()) != 0) return;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();  // launches being captured into a graph are not timed
    return;
  }
  start = prof_event();
  stop = start ? prof_event() : nullptr;
  if (!stop) {
    if (start) (void)hipEventDestroy(start);
    start = nullptr;
    return;
  }
  if (!attach) (void)hipEventRecord(start, stream);
}

ProfScope::~ProfScope() {
  if (!start) return;
  if (attach && !attached) {  // the launch did not take the events
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_pool.push_back(start);
    g_prof_pool.push_back(stop);
    return;
  }
  if (!attach) (void)hipEventRecord(stop, stream);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back({name, start, stop});
}

// ---------------------------------------------------------------------------
// Copy kernel (bandwidth calibration)
// ---------------------------------------------------------------------------
// Four 16-byte nontemporal loads per thread in flight, then their stores,
// over a grid of at most kCopyGrid blocks: the fastest form measured for a
// 400 MB copy on MI355X (136 us, 5.86 TB/s; one plain 16-byte load per
// iteration over 4,096 blocks: 157 us, 5.10 TB/s; eight loads in flight or
// larger grids: 141-190 us).  It is the achievable-bandwidth reference the
// kernels' traffic is compared with.
constexpr unsigned kCopyGrid = 1024;
__global__ __launch_bounds__(256) void k_copy_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                  size_t nvec, size_t n) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const v4u* s4 = reinterpret_cast<const v4u*>(src);
  v4u* d4 = reinterpret_cast<v4u*>(dst);
  constexpr int U = 4;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t b = 0; b < nvec; b += stride * U) {
    v4u x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = b + i0 + u * stride;
      if (i < nvec) x[u] = __builtin_nontemporal_load(s4 + i);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = b + i0 + u * stride;
      if (i < nvec) __builtin_nontemporal_store(x[u], d4 + i);
    }
  }
  if (blockIdx.x == 0)
    for (size_t i = nvec * 4 + threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Counter-based generator: x_i = splitmix64(seed + i * golden), one output
// per index, so any slice can be generated independently.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_random(T* __restrict__ dst, size_t n, uint64_t seed) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = (T)splitmix64(seed + i * 0x9E3779B97F4A7C15ull);
}

}  // namespace pa

using namespace pa;

extern "C" {

const char* pa_status_string(int status) {
  switch (status) {
    case PA_OK: return "ok";
    case PA_ERR_ARG: return "invalid argument";
    case PA_ERR_OOM: return "out of memory";
    case PA_ERR_HIP: return "HIP runtime error";
    case PA_ERR_TOO_LARGE: return "input too large for this entry point";
    case PA_ERR_NO_DEVICE: return "no usable device";
    default: return "unknown status";
  }
}

const char* pa_last_error_string(void) { return t_last_error.c_str(); }

int pa_version(void) { return 10000; }

int pa_set_device(int device) {
  PA_HIP(hipSetDevice(device));
  return PA_OK;
}

pa_stream_t pa_thread_stream(void) { return (pa_stream_t)thread_stream(); }

int pa_is_device_pointer(const void* p) { return is_device_ptr(p) ? 1 : 0; }

int pa_synchronize(pa_stream_t stream) {
  PA_HIP(hipStreamSynchronize((hipStream_t)stream));
  return PA_OK;
}

int pa_copy(void* dst, const void* src, size_t bytes, pa_stream_t stream) {
  if (bytes == 0 || dst == src) return PA_OK;
  if (!dst || !src) return arg_error("pa_copy: null pointer");
  hipStream_t s = (hipStream_t)stream;
  PA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
  if (!is_device_ptr(dst) || !is_device_ptr(src)) PA_HIP(hipStreamSynchronize(s));
  return PA_OK;
}

int pa_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, pa_stream_t stream) {
  if (n == 0) return PA_OK;
  if (!dst || !src) return arg_error("pa_copy_u32: null pointer");
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return arg_error("pa_copy_u32: pointers must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  size_t nvec = n / 4;
  unsigned grid = (unsigned)std::min<size_t>((nvec + 255) / 256 + 1, kCopyGrid);
  ProfScope ps("k_copy_u32", s);
  hipLaunchKernelGGL(k_copy_u32, dim3(grid), dim3(256), 0, s, dst, src, nvec, n);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

int pa_profile_enable(int enable) {
  g_prof_on.store(enable != 0);
  return PA_OK;
}

int pa_profile_only(const char* prefix) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_only = prefix ? prefix : "";
  return PA_OK;
}

int pa_profile_query(const char* prefix, double* total_ms, uint64_t* launches) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  double tot = 0.0;
  uint64_t cnt = 0;
  size_t plen = prefix ? std::strlen(prefix) : 0;
  for (auto& r : g_prof) {
    if (plen && r.name.compare(0, plen, prefix) != 0) continue;
    PA_HIP(hipEventSynchronize(r.stop));
    float ms = 0.f;
    PA_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    tot += ms;
    ++cnt;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = cnt;
  return PA_OK;
}

int pa_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (auto& r : g_prof) {
    (void)hipEventSynchronize(r.stop);
    g_prof_pool.push_back(r.start);
    g_prof_pool.push_back(r.stop);
  }
  g_prof.clear();
  return PA_OK;
}

int pa_fill_random_u32(uint32_t* dst, size_t n, uint64_t seed, pa_stream_t stream) {
  if (n == 0) return PA_OK;
  if (!dst || !is_device_ptr(dst)) return arg_error("pa_fill_random_u32: dst must be device memory");
  hipLaunchKernelGGL(k_fill_random<uint32_t>, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0,
                     (hipStream_t)stream, dst, n, seed);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

int pa_fill_random_u64(uint64_t* dst, size_t n, uint64_t seed, pa_stream_t stream) {
  if (n == 0) return PA_OK;
  if (!dst || !is_device_ptr(dst)) return arg_error("pa_fill_random_u64: dst must be device memory");
  hipLaunchKernelGGL(k_fill_random<uint64_t>, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0,
                     (hipStream_t)stream, dst, n, seed);
  PA_HIP(hipGetLastError());
  return PA_OK;
}

void pa_gen_mt19937_u32(uint64_t seed, uint32_t* out, size_t n) {
  std::mt19937_64 g(seed);
  for (size_t i = 0; i < n; ++i) out[i] = static_cast<uint32_t>(g());
}

void pa_gen_mt19937_u64(uint64_t seed, uint64_t* out, size_t n) {
  std::mt19937_64 g(seed);
  for (size_t i = 0; i < n; ++i) out[i] = g();
}

}  // extern "C"
