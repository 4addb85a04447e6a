// pa_partition.hip -- radix partition (nth_element semantics) for gfx950.
//
// Reference: PartitionRadix (RadixPartition.h:94-184) with its MSD in-place
// partition PartitionRadixMsdUIntInner (:30-86): afterwards a[k] holds the
// k-th smallest key, nothing before it is larger and nothing after it is
// smaller, and the array is a permutation of its input.  The reference's exact
// arrangement of the other keys comes from its in-place swap order and is not
// part of that contract; its own benchmark checks a[k] against nth_element
// (RadixPartitionBenchmark.cpp:50-66).
//
// GPU form, one rank k:
//   radix select of v = the k-th smallest (the select kernels, one read of the
//   keys plus passes over a shrinking candidate set), then one stable
//   three-way partition of the keys into (< v | == v | > v):
//     k_part_count    per block: keys < v and == v over its tiles; also
//                     copies the keys into scratch
//     k_part_offsets  one block: each block's start in each of the 3 groups
//     k_part_scatter  per block, same tiles in the same order, reading the
//                     scratch copy: each key goes to its group's next slot
//                     (wave ballots give the stable order inside a tile),
//                     written straight over the keys
//   Since count(< v) <= k < count(<= v), slot k holds v.
// Several ranks: the keys are sorted (a sorted array satisfies every rank's
// partition at once).
#include <algorithm>
#include <vector>

#include "pa_internal.h"

namespace pa {

constexpr int kPartBlock = 256;
constexpr int kPartItems = 16;
constexpr uint32_t kPartTile = kPartBlock * kPartItems;  // 4096 keys
constexpr unsigned kPartGrid = 1024;
// k_part_offsets gives each block of the count and scatter passes one thread
// of its single workgroup, so it serves at most this many blocks.
constexpr unsigned kPartOffsetsThreads = 1024;

// Host check before the launch of k_part_offsets: a grid it cannot serve
// would leave blocks without offsets and the scatter writing through them.
int part_offsets_grid_check(size_t nb) {
  if (nb == 0 || nb > kPartOffsetsThreads)
    return arg_error("partition: the offsets pass serves 1 to 1024 blocks");
  return PA_OK;
}

// Block b owns the contiguous tiles [b * tpb, (b + 1) * tpb), tpb =
// ceil(tiles / blocks), and item j of lane l in wave w of a tile is key
// w*1024 + j*64 + l, so (block, tile, wave, item, lane) order is input order.
__device__ __host__ inline size_t part_tiles_per_block(size_t n, size_t nb) {
  const size_t ntiles = (n + kPartTile - 1) / kPartTile;
  return (ntiles + nb - 1) / nb;
}

__global__ __launch_bounds__(kPartBlock) void k_part_count(const uint32_t* __restrict__ keys, size_t n,
                                                           const uint32_t* __restrict__ pivot,
                                                           uint32_t* __restrict__ counts,
                                                           uint32_t* __restrict__ copy_out) {
  // Counting needs no order, so the block reads its whole key range with
  // 16-byte loads (four in flight per thread) when the keys are 16-byte
  // aligned; the range starts on a tile boundary, a multiple of 4 keys.  The
  // keys are also written to copy_out, the scatter's source, so the scatter
  // can write the result over the keys without a copy back.
  __shared__ uint32_t s_c[2];
  const uint32_t tid = threadIdx.x;
  if (tid < 2) s_c[tid] = 0;
  __syncthreads();
  const uint32_t v = *pivot;
  uint32_t lt = 0, eq = 0;
  const size_t span = part_tiles_per_block(n, gridDim.x) * kPartTile;
  const size_t lo = std::min<size_t>(n, (size_t)blockIdx.x * span), hi = std::min<size_t>(n, lo + span);
  size_t scalar_from = lo;
  if ((((uintptr_t)keys | (uintptr_t)copy_out) & 15) == 0) {
    const uint4* vk = reinterpret_cast<const uint4*>(keys + lo);
    uint4* vo = reinterpret_cast<uint4*>(copy_out + lo);
    const size_t nv = (hi - lo) / 4;
    constexpr int U = 4;
    for (size_t i0 = tid; i0 < nv; i0 += (size_t)kPartBlock * U) {
      uint4 x[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u * kPartBlock < nv) x[u] = vk[i0 + u * kPartBlock];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (i0 + u * kPartBlock >= nv) continue;
        vo[i0 + u * kPartBlock] = x[u];
        const uint32_t w[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          lt += w[j] < v;
          eq += w[j] == v;
        }
      }
    }
    scalar_from = lo + nv * 4;
  }
  for (size_t i = scalar_from + tid; i < hi; i += kPartBlock) {
    const uint32_t k = keys[i];
    copy_out[i] = k;
    lt += k < v;
    eq += k == v;
  }
  // wave reduction, then one atomic per wave into LDS
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lt += __shfl_down(lt, off, 64);
    eq += __shfl_down(eq, off, 64);
  }
  if ((tid & 63) == 0) {
    atomicAdd(&s_c[0], lt);
    atomicAdd(&s_c[1], eq);
  }
  __syncthreads();
  if (tid < 2) counts[blockIdx.x * 2 + tid] = s_c[tid];
}

// offsets[b][g] = first output slot of block b's keys of group g
// (0: < v, 1: == v, 2: > v).  One block of 1024 threads, nb <= 1024
// (part_offsets_grid_check runs on the host before every launch).
__global__ __launch_bounds__(kPartOffsetsThreads) void k_part_offsets(const uint32_t* __restrict__ counts, uint32_t nb, size_t n,
                                                       const uint32_t* __restrict__ items_per_block,
                                                       uint32_t* __restrict__ offsets) {
  __shared__ uint32_t s_sum[3][kPartOffsetsThreads / 64];
  __shared__ uint32_t s_tot[3];
  const uint32_t b = threadIdx.x, lane = b & 63, wave = b >> 6;
  uint32_t c[3] = {0, 0, 0};
  if (b < nb) {
    c[0] = counts[2 * b];
    c[1] = counts[2 * b + 1];
    c[2] = items_per_block[b] - c[0] - c[1];
  }
  uint32_t inc[3];
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    inc[g] = dev::wave_incl_scan(c[g], lane);
    if (lane == 63) s_sum[g][wave] = inc[g];
  }
  __syncthreads();
  if (b < 3) {
    uint32_t t = 0;
    for (unsigned w = 0; w < kPartOffsetsThreads / 64; ++w) t += s_sum[b][w];
    s_tot[b] = t;
  }
  __syncthreads();
  if (b < nb) {
    uint32_t base = 0;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      uint32_t pre = 0;
      for (uint32_t w = 0; w < wave; ++w) pre += s_sum[g][w];
      offsets[3 * b + g] = base + pre + inc[g] - c[g];
      base += s_tot[g];
    }
  }
}

__global__ __launch_bounds__(kPartBlock) void k_part_scatter(const uint32_t* __restrict__ keys, size_t n,
                                                             const uint32_t* __restrict__ pivot,
                                                             const uint32_t* __restrict__ offsets,
                                                             uint32_t* __restrict__ out) {
  __shared__ uint32_t s_wave[4][3];  // per wave, per group: keys in the tile
  __shared__ uint32_t s_run[3];      // the block's next slot per group
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < 3) s_run[tid] = offsets[3 * blockIdx.x + tid];
  const uint32_t v = *pivot;
  const uint64_t below = (1ull << lane) - 1ull;
  const size_t ntiles = (n + kPartTile - 1) / kPartTile;
  const size_t tpb = part_tiles_per_block(n, gridDim.x);
  const size_t t1 = std::min<size_t>(ntiles, (blockIdx.x + 1) * tpb);
  for (size_t t = blockIdx.x * tpb; t < t1; ++t) {
    const size_t base = t * kPartTile + wave * 1024 + lane;
    uint32_t key[kPartItems];
    uint32_t grp[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      const size_t idx = base + j * 64;
      key[j] = idx < n ? keys[idx] : 0u;
      grp[j] = idx < n ? (key[j] < v ? 0u : key[j] == v ? 1u : 2u) : 3u;
    }
    // Rank of every key among its wave's keys of the same group, in order.
    uint32_t seen[3] = {0, 0, 0};
    uint32_t pos[kPartItems];
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) {
      uint32_t r = 0;
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        const uint64_t m = __ballot(grp[j] == (uint32_t)g);
        if (grp[j] == (uint32_t)g) r = seen[g] + (uint32_t)__popcll(m & below);
        seen[g] += (uint32_t)__popcll(m);
      }
      pos[j] = r;
    }
    __syncthreads();  // s_run of the previous tile is final
    if (lane == 0)
      for (int g = 0; g < 3; ++g) s_wave[wave][g] = seen[g];
    __syncthreads();
    uint32_t wbase[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      uint32_t pre = s_run[g];
      for (uint32_t w = 0; w < wave; ++w) pre += s_wave[w][g];
      wbase[g] = pre;
    }
#pragma unroll
    for (int j = 0; j < kPartItems; ++j)
      if (grp[j] < 3) out[wbase[grp[j]] + pos[j]] = key[j];
    __syncthreads();
    if (tid < 3) s_run[tid] += s_wave[0][tid] + s_wave[1][tid] + s_wave[2][tid] + s_wave[3][tid];
  }
}

// Keys per block of the partition grid: [b * tpb * tile, (b + 1) * tpb * tile)
// clipped to n.
__global__ void k_part_block_items(uint32_t nb, size_t n, uint32_t* __restrict__ items) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const size_t span = part_tiles_per_block(n, nb) * kPartTile;
  const size_t lo = std::min<size_t>(n, (size_t)b * span), hi = std::min<size_t>(n, lo + span);
  items[b] = (uint32_t)(hi - lo);
}

}  // namespace pa

using namespace pa;

extern "C" {

// The offsets pass's grid check on its own, for host-side tests (no device
// needed).  Not part of the public header.
int pa_partition_offsets_grid_check(size_t nb) { return part_offsets_grid_check(nb); }

int pa_partition_u32(uint32_t* keys, size_t n, size_t k, uint32_t* result, pa_stream_t stream) {
  if (!keys || !result) return arg_error("pa_partition_u32: null pointer");
  if (n == 0 || k >= n) return arg_error("pa_partition_u32: need 0 <= k < n");
  if (n > 0xFFFFFFFFull) return arg_error("pa_partition_u32: n must be < 2^32");
  hipStream_t s = (hipStream_t)stream;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const bool dk = is_device_ptr(keys);
  const bool dres = is_device_ptr(result);
  const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>((n + kPartTile - 1) / kPartTile, kPartGrid));
  PA_TRY(part_offsets_grid_check(nb));  // before anything is enqueued
  Layout L;
  const size_t o_sel = L.take(select_scratch_bytes_u32(n));
  const size_t o_piv = L.take(4);
  const size_t o_cnt = L.take((size_t)nb * 2 * 4);
  const size_t o_items = L.take((size_t)nb * 4);
  const size_t o_off = L.take((size_t)nb * 3 * 4);
  const size_t o_tmp = L.take(n * 4);
  const size_t o_in = dk ? 0 : L.take(n * 4);
  void* basep;
  PA_TRY(ensure_scratch(ctx, L.off, &basep));
  char* b = (char*)basep;
  uint32_t* pivot = (uint32_t*)(b + o_piv);
  uint32_t* counts = (uint32_t*)(b + o_cnt);
  uint32_t* items = (uint32_t*)(b + o_items);
  uint32_t* offsets = (uint32_t*)(b + o_off);
  uint32_t* tmp = (uint32_t*)(b + o_tmp);
  uint32_t* din = keys;
  if (!dk) {
    din = (uint32_t*)(b + o_in);
    PA_TRY(stage_to_device(ctx, din, keys, n * 4));
  }
  PA_TRY(select_enqueue_u32(din, n, k, b + o_sel, pivot, s));
  {
    ProfScope ps("k_part_count", s);
    hipLaunchKernelGGL(k_part_count, dim3(nb), dim3(kPartBlock), 0, s, din, n, pivot, counts, tmp);
  }
  hipLaunchKernelGGL(k_part_block_items, dim3((nb + 255) / 256), dim3(256), 0, s, nb, n, items);
  PA_TRY(part_offsets_grid_check(nb));
  hipLaunchKernelGGL(k_part_offsets, dim3(1), dim3(kPartOffsetsThreads), 0, s, counts, nb, n, items, offsets);
  {
    ProfScope ps("k_part_scatter", s);
    hipLaunchKernelGGL(k_part_scatter, dim3(nb), dim3(kPartBlock), 0, s, tmp, n, pivot, offsets, din);
  }
  PA_HIP(hipGetLastError());
  if (!dk) PA_TRY(stage_to_host(ctx, keys, din, n * 4));
  PA_HIP(hipMemcpyAsync(result, pivot, 4, dres ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  if (!dk || !dres) PA_HIP(hipStreamSynchronize(s));
  return PA_OK;
}

int pa_partition_multi_u32(uint32_t* keys, size_t n, const size_t* ks, size_t num_ks, pa_stream_t stream) {
  if (n == 0 || num_ks == 0) return PA_OK;
  if (!keys || !ks) return arg_error("pa_partition_multi_u32: null pointer");
  for (size_t i = 0; i < num_ks; ++i) {
    if (ks[i] >= n) return arg_error("pa_partition_multi_u32: every k must be below n");
    if (i && ks[i] < ks[i - 1]) return arg_error("pa_partition_multi_u32: ks must be ascending");
  }
  return pa_sort_u32(keys, nullptr, n, stream);
}

}  // extern "C"
