// pa_merge.hip -- merge-path merge of two sorted runs for gfx950.
//
// Reference: merge_parallel_L5 (ParallelMerge.h:285-315) splits the longer
// run at its midpoint, binary-searches the pivot in the other run
// (my_binary_search, BinarySearch.h:33-47), places it and recurses in parallel
// down to a 32K cut-off, where merge_ptr_1 (:52-75) merges sequentially with
// `a <= b` taking the first run's element.
//
// GPU form: the output is cut into tiles of kMergeTile elements.  A partition
// kernel finds, for every tile boundary (a diagonal of the merge grid), how many
// elements come from run A by a binary search along the diagonal -- the same
// lower-bound search the reference recurses on, done for all splits at once.
// (Measured alternatives, slower at 2 x 100M u32: each tile searching its own
// boundaries at the start of the tile kernel, 64-ary by one wave, 0.77 ms a
// merge against 0.44, as the searches wait behind the streaming traffic; and
// a 16-ary search by 16 lanes per boundary, 136-147 us against 46-60 us for
// the partition, 16 times the scattered reads for 4 times fewer steps.)
// Each workgroup then loads its A and B slices into LDS, every thread finds its
// own sub-diagonal in LDS, merges kMergeItems elements (ties take A), and the
// tile is written out coalesced through LDS.
#include <algorithm>

#include "pa_internal.h"

namespace pa {

// 512 threads x 8 outputs: at 2 x 100M u32 the tile kernel and its partition
// take 440 us against 529 us for 256 x 16 (each thread's merge is a chain of
// dependent LDS reads, so shorter chains and more of them in flight pay) and
// 448-562 us for the other shapes measured (256 x 4/8, 512 x 4, 1024 x 4/8).
constexpr int kMergeBlock = 512;
constexpr int kMergeItems = 8;
constexpr int kMergeTile = kMergeBlock * kMergeItems;

// Number of elements taken from A among the first d outputs of the stable
// merge (ties take A first).
template <typename K>
__device__ __forceinline__ size_t merge_path(const K* a, size_t na, const K* b, size_t nb, size_t d) {
  size_t lo = d > nb ? d - nb : 0;
  size_t hi = d < na ? d : na;
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if (a[mid] <= b[d - 1 - mid])
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The same split as merge_path, for runs in global memory, where every probe
// is a dependent HBM round trip: the probes are placed by interpolation.
// P(i) = a[i] <= b[d - 1 - i] is true below the split and false from it on,
// and f(i) = a[i] - b[d - 1 - i] rises with i, so the split is near the zero
// crossing of the line through the two nearest probed points (one with
// P true, one with P false).  A step that does not halve the interval is
// followed by a bisection step, so there are at most about twice as many
// probes as a binary search makes; on evenly spread keys a handful of probes
// leave an interval of 32 or fewer, which the binary search finishes.  Only
// where the probes go depends on the key values (through f, in double), never
// the result: every decision is the exact comparison P.
template <typename K>
__device__ size_t merge_path_interp(const K* a, size_t na, const K* b, size_t nb, size_t d) {
  size_t lo = d > nb ? d - nb : 0;
  size_t hi = d < na ? d : na;  // invariant: P true below lo, false from hi on
  if (hi - lo > 32) {
    const K a0 = a[lo], b0 = b[d - 1 - lo], a1 = a[hi - 1], b1 = b[d - hi];
    if (!(a0 <= b0)) return lo;
    if (a1 <= b1) return hi;
    // Known points: pl (P true, f <= 0) and ph (P false, f > 0).
    size_t pl = lo, ph = hi - 1;
    double fl = (double)a0 - (double)b0, fh = (double)a1 - (double)b1;
    lo = pl + 1;
    hi = ph;
    bool bisect = false;
    while (hi - lo > 32) {
      size_t mid = lo + (hi - lo) / 2;
      if (!bisect && fh > fl) {
        const double x = (double)pl + (double)(ph - pl) * (-fl / (fh - fl));
        const double m = x < (double)lo ? (double)lo : (x > (double)(hi - 1) ? (double)(hi - 1) : x);
        mid = (size_t)m;
      }
      const K am = a[mid], bm = b[d - 1 - mid];
      const size_t before = hi - lo;
      if (am <= bm) {
        lo = mid + 1;
        pl = mid;
        fl = (double)am - (double)bm;
      } else {
        hi = mid;
        ph = mid;
        fh = (double)am - (double)bm;
      }
      bisect = 2 * (hi - lo) > before;
    }
  }
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if (a[mid] <= b[d - 1 - mid])
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <typename K>
__global__ __launch_bounds__(256) void k_merge_partition(const K* __restrict__ a, size_t na, const K* __restrict__ b,
                                                         size_t nb, size_t ntiles, size_t* __restrict__ split) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const size_t d = std::min(t * (size_t)kMergeTile, na + nb);
  split[t] = merge_path_interp(a, na, b, nb, d);
}

template <typename K>
__global__ __launch_bounds__(kMergeBlock) void k_merge_tile(const K* __restrict__ a, size_t na, const K* __restrict__ b,
                                                            size_t nb, const size_t* __restrict__ split,
                                                            K* __restrict__ out) {
  __shared__ K s_in[kMergeTile];
  __shared__ K s_out[kMergeTile];
  const uint32_t tid = threadIdx.x;
  const size_t tile = blockIdx.x;
  const size_t d0 = tile * kMergeTile;
  const size_t d1 = std::min(d0 + kMergeTile, na + nb);
  const size_t i0 = split[tile], i1 = split[tile + 1];
  const size_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t la = (uint32_t)(i1 - i0), lb = (uint32_t)(j1 - j0);
  // All kMergeItems loads of a thread are issued before the first LDS store,
  // so each thread keeps its whole share of the tile in flight.
  {
    K v[kMergeItems];
#pragma unroll
    for (int x = 0; x < kMergeItems; ++x) {
      const uint32_t e = tid + x * kMergeBlock;
      // Streamed once: nontemporal loads and stores keep the runs and the
      // output from displacing each other in L2 (2 x 100M u32: tile kernel
      // 392 -> 377 us).
      v[x] = e < la ? __builtin_nontemporal_load(a + i0 + e)
                    : (e < la + lb ? __builtin_nontemporal_load(b + j0 + (e - la)) : K(0));
    }
#pragma unroll
    for (int x = 0; x < kMergeItems; ++x) {
      const uint32_t e = tid + x * kMergeBlock;
      if (e < la + lb) s_in[e] = v[x];
    }
  }
  __syncthreads();
  const K* sa = s_in;
  const K* sb = s_in + la;
  const uint32_t total = la + lb;
  const uint32_t dt = std::min<uint32_t>(tid * kMergeItems, total);
  uint32_t ia = (uint32_t)merge_path(sa, la, sb, lb, dt);
  uint32_t ib = dt - ia;
  K r[kMergeItems];
#pragma unroll
  for (int x = 0; x < kMergeItems; ++x) {
    if (dt + x < total) {
      const bool take_a = ib >= lb || (ia < la && sa[ia] <= sb[ib]);
      r[x] = take_a ? sa[ia] : sb[ib];
      ia += take_a;
      ib += !take_a;
    }
  }
#pragma unroll
  for (int x = 0; x < kMergeItems; ++x)
    if (dt + x < total) s_out[dt + x] = r[x];
  __syncthreads();
#pragma unroll
  for (int x = 0; x < kMergeItems; ++x) {
    const uint32_t e = tid + x * kMergeBlock;
    if (e < total) __builtin_nontemporal_store(s_out[e], out + d0 + e);
  }
}

// Key-value form of k_merge_tile (ties take A, so merging an earlier run with
// a later one keeps a stable sort stable): keys and payloads of the tile are
// staged in LDS side by side, each thread records the staging slot of every
// output it merges, and the write-out gathers key and payload through it.
template <typename K>
__global__ __launch_bounds__(kMergeBlock) void k_merge_tile_kv(const K* __restrict__ a, const uint32_t* __restrict__ va,
                                                               size_t na, const K* __restrict__ b,
                                                               const uint32_t* __restrict__ vb, size_t nb,
                                                               const size_t* __restrict__ split, K* __restrict__ out,
                                                               uint32_t* __restrict__ vout) {
  __shared__ K s_in[kMergeTile];
  __shared__ uint32_t s_v[kMergeTile];
  __shared__ uint16_t s_src[kMergeTile];
  static_assert(kMergeTile <= 65536, "16-bit staging slots");
  const uint32_t tid = threadIdx.x;
  const size_t tile = blockIdx.x;
  const size_t d0 = tile * kMergeTile;
  const size_t d1 = std::min(d0 + kMergeTile, na + nb);
  const size_t i0 = split[tile], i1 = split[tile + 1];
  const size_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t la = (uint32_t)(i1 - i0), lb = (uint32_t)(j1 - j0);
  {
    K v[kMergeItems];
    uint32_t w[kMergeItems];
#pragma unroll
    for (int x = 0; x < kMergeItems; ++x) {
      const uint32_t e = tid + x * kMergeBlock;
      if (e < la) {
        v[x] = __builtin_nontemporal_load(a + i0 + e);
        w[x] = __builtin_nontemporal_load(va + i0 + e);
      } else if (e < la + lb) {
        v[x] = __builtin_nontemporal_load(b + j0 + (e - la));
        w[x] = __builtin_nontemporal_load(vb + j0 + (e - la));
      }
    }
#pragma unroll
    for (int x = 0; x < kMergeItems; ++x) {
      const uint32_t e = tid + x * kMergeBlock;
      if (e < la + lb) {
        s_in[e] = v[x];
        s_v[e] = w[x];
      }
    }
  }
  __syncthreads();
  const K* sa = s_in;
  const K* sb = s_in + la;
  const uint32_t total = la + lb;
  const uint32_t dt = std::min<uint32_t>(tid * kMergeItems, total);
  uint32_t ia = (uint32_t)merge_path(sa, la, sb, lb, dt);
  uint32_t ib = dt - ia;
  uint32_t r[kMergeItems];
#pragma unroll
  for (int x = 0; x < kMergeItems; ++x) {
    if (dt + x < total) {
      const bool take_a = ib >= lb || (ia < la && sa[ia] <= sb[ib]);
      r[x] = take_a ? ia : la + ib;
      ia += take_a;
      ib += !take_a;
    }
  }
  __syncthreads();  // every merge read of s_in is done before s_src is written
#pragma unroll
  for (int x = 0; x < kMergeItems; ++x)
    if (dt + x < total) s_src[dt + x] = (uint16_t)r[x];
  __syncthreads();
#pragma unroll
  for (int x = 0; x < kMergeItems; ++x) {
    const uint32_t e = tid + x * kMergeBlock;
    if (e < total) {
      const uint32_t src = s_src[e];
      __builtin_nontemporal_store(s_in[src], out + d0 + e);
      __builtin_nontemporal_store(s_v[src], vout + d0 + e);
    }
  }
}

template <typename K>
int merge_entry(const K* a, size_t na, const K* b, size_t nb, K* out, hipStream_t s, const char* who) {
  const size_t n = na + nb;
  if (n == 0) return PA_OK;
  if (!out || (na && !a) || (nb && !b)) return arg_error(who);
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool da = na == 0 || is_device_ptr(a);
  const bool db = nb == 0 || is_device_ptr(b);
  const bool dout = is_device_ptr(out);
  const size_t ntiles = (n + kMergeTile - 1) / kMergeTile;
  Layout L;
  const size_t o_split = L.take((ntiles + 1) * sizeof(size_t));
  const size_t o_a = da ? 0 : L.take(na * sizeof(K));
  const size_t o_b = db ? 0 : L.take(nb * sizeof(K));
  const size_t o_o = dout ? 0 : L.take(n * sizeof(K));
  void* base;
  PA_TRY(ensure_scratch(ctx, L.off, &base));
  char* p = (char*)base;
  size_t* split = (size_t*)(p + o_split);
  const K* dA = a;
  const K* dB = b;
  K* dO = dout ? out : (K*)(p + o_o);
  if (!da) {
    PA_TRY(stage_to_device(ctx, p + o_a, a, na * sizeof(K)));
    dA = (const K*)(p + o_a);
  }
  if (!db) {
    PA_TRY(stage_to_device(ctx, p + o_b, b, nb * sizeof(K)));
    dB = (const K*)(p + o_b);
  }
  {
    ProfScope ps("k_merge_partition", s);
    hipLaunchKernelGGL(k_merge_partition<K>, dim3((unsigned)((ntiles + 1 + 255) / 256)), dim3(256), 0, s, dA, na, dB,
                       nb, ntiles, split);
  }
  {
    ProfScope ps("k_merge_tile", s);
    hipLaunchKernelGGL(k_merge_tile<K>, dim3((unsigned)ntiles), dim3(kMergeBlock), 0, s, dA, na, dB, nb, split, dO);
  }
  PA_HIP(hipGetLastError());
  if (!dout) {
    PA_TRY(stage_to_host(ctx, out, dO, n * sizeof(K)));
  }
  if (!da || !db || !dout) PA_HIP(hipStreamSynchronize(s));
  return PA_OK;
}

// Device runs into a device output, with the split table in caller scratch of
// merge_split_bytes(na + nb) bytes; stream-ordered, no allocation.
size_t merge_split_bytes(size_t n) { return ((n + kMergeTile - 1) / kMergeTile + 1) * sizeof(size_t); }

template <typename K>
int merge_enqueue(const K* a, size_t na, const K* b, size_t nb, K* out, size_t* split, hipStream_t s) {
  const size_t n = na + nb;
  if (n == 0) return PA_OK;
  const size_t ntiles = (n + kMergeTile - 1) / kMergeTile;
  {
    ProfScope ps("k_merge_partition", s);
    hipLaunchKernelGGL(k_merge_partition<K>, dim3((unsigned)((ntiles + 1 + 255) / 256)), dim3(256), 0, s, a, na, b,
                       nb, ntiles, split);
  }
  {
    ProfScope ps("k_merge_tile", s);
    hipLaunchKernelGGL(k_merge_tile<K>, dim3((unsigned)ntiles), dim3(kMergeBlock), 0, s, a, na, b, nb, split, out);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}
template int merge_enqueue<uint32_t>(const uint32_t*, size_t, const uint32_t*, size_t, uint32_t*, size_t*,
                                     hipStream_t);

// Key-value runs (device memory) into device outputs, split table as above.
template <typename K>
int merge_kv_enqueue(const K* a, const uint32_t* va, size_t na, const K* b, const uint32_t* vb, size_t nb, K* out,
                     uint32_t* vout, size_t* split, hipStream_t s) {
  const size_t n = na + nb;
  if (n == 0) return PA_OK;
  const size_t ntiles = (n + kMergeTile - 1) / kMergeTile;
  {
    ProfScope ps("k_merge_partition", s);
    hipLaunchKernelGGL(k_merge_partition<K>, dim3((unsigned)((ntiles + 1 + 255) / 256)), dim3(256), 0, s, a, na, b,
                       nb, ntiles, split);
  }
  {
    ProfScope ps("k_merge_tile_kv", s);
    hipLaunchKernelGGL(k_merge_tile_kv<K>, dim3((unsigned)ntiles), dim3(kMergeBlock), 0, s, a, va, na, b, vb, nb,
                       split, out, vout);
  }
  PA_HIP(hipGetLastError());
  return PA_OK;
}
template int merge_kv_enqueue<uint32_t>(const uint32_t*, const uint32_t*, size_t, const uint32_t*, const uint32_t*,
                                        size_t, uint32_t*, uint32_t*, size_t*, hipStream_t);
template int merge_kv_enqueue<uint64_t>(const uint64_t*, const uint32_t*, size_t, const uint64_t*, const uint32_t*,
                                        size_t, uint64_t*, uint32_t*, size_t*, hipStream_t);
template int merge_enqueue<uint64_t>(const uint64_t*, size_t, const uint64_t*, size_t, uint64_t*, size_t*,
                                     hipStream_t);

// In-place form: merge t[0..n1) and t[n1..n1+n2) into scratch, copy back.
template <typename K>
int merge_inplace_entry(K* t, size_t n1, size_t n2, hipStream_t s, const char* who) {
  const size_t n = n1 + n2;
  if (n1 == 0 || n2 == 0) return PA_OK;  // one run: already merged
  if (!t) return arg_error(who);
  const bool dt = is_device_ptr(t);
  if (!dt) {
    // Host runs: both are staged to device memory before the merged result is
    // copied back over t (same stream, in order), so t may be the output.
    return merge_entry<K>(t, n1, t + n1, n2, t, s, who);
  }
  StreamCtx& ctx = stream_ctx(s);
  K* out = nullptr;
  {
    std::lock_guard<std::mutex> lk(ctx.mu);
    Layout L;
    const size_t ntiles = (n + kMergeTile - 1) / kMergeTile;
    const size_t o_split = L.take((ntiles + 1) * sizeof(size_t));
    const size_t o_out = L.take(n * sizeof(K));
    void* base;
    PA_TRY(ensure_scratch(ctx, L.off, &base));
    char* p = (char*)base;
    size_t* split = (size_t*)(p + o_split);
    out = (K*)(p + o_out);
    {
      ProfScope ps("k_merge_partition", s);
      hipLaunchKernelGGL(k_merge_partition<K>, dim3((unsigned)((ntiles + 1 + 255) / 256)), dim3(256), 0, s, t, n1,
                         t + n1, n2, ntiles, split);
    }
    {
      ProfScope ps("k_merge_tile", s);
      hipLaunchKernelGGL(k_merge_tile<K>, dim3((unsigned)ntiles), dim3(kMergeBlock), 0, s, t, n1, t + n1, n2, split,
                         out);
    }
    PA_HIP(hipGetLastError());
    PA_HIP(hipMemcpyAsync(t, out, n * sizeof(K), hipMemcpyDeviceToDevice, s));
  }
  return PA_OK;
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_merge_u32(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out, pa_stream_t stream) {
  return merge_entry<uint32_t>(a, na, b, nb, out, (hipStream_t)stream, "pa_merge_u32: null pointer");
}

int pa_merge_u64(const uint64_t* a, size_t na, const uint64_t* b, size_t nb, uint64_t* out, pa_stream_t stream) {
  return merge_entry<uint64_t>(a, na, b, nb, out, (hipStream_t)stream, "pa_merge_u64: null pointer");
}

int pa_merge_inplace_u32(uint32_t* t, size_t n1, size_t n2, pa_stream_t stream) {
  return merge_inplace_entry<uint32_t>(t, n1, n2, (hipStream_t)stream, "pa_merge_inplace_u32: null pointer");
}

int pa_merge_inplace_u64(uint64_t* t, size_t n1, size_t n2, pa_stream_t stream) {
  return merge_inplace_entry<uint64_t>(t, n1, n2, (hipStream_t)stream, "pa_merge_inplace_u64: null pointer");
}

}  // extern "C"
