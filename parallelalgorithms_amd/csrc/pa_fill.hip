// pa_fill.hip -- parallel fill for gfx950.
//
// Reference: parallel_fill (FillParallel.h:22-65), std::fill of
// src[l, r) with one value, split recursively over TBB tasks.  Here one
// launch writes the range with 16-byte stores: the element pattern repeated
// to 16 bytes (16 is a multiple of every element size 1, 2, 4, 8, and the
// range starts on an element boundary, so every 16-byte-aligned address is one
// too), with the unaligned head and tail written byte by byte.  Pure write
// traffic: 1 B per byte filled.
#include "pa_internal.h"

namespace pa {

constexpr int kFillBlock = 256;
constexpr unsigned kFillMaxGrid = 8192;

// a[0, nbytes) = value repeated; byte i of the range gets byte (i % es) of the
// element, so pat is the 16-byte image of the 16-aligned words.
__global__ __launch_bounds__(kFillBlock) void k_fill_pattern(uint8_t* __restrict__ a, size_t nbytes, uint4 pat,
                                                             uint32_t es) {
  const size_t head = std::min<size_t>(nbytes, (16 - ((uintptr_t)a & 15)) & 15);
  const size_t nvec = (nbytes - head) / 16;
  uint4* v = reinterpret_cast<uint4*>(a + head);
  const size_t stride = (size_t)gridDim.x * kFillBlock;
  for (size_t i = (size_t)blockIdx.x * kFillBlock + threadIdx.x; i < nvec; i += stride) v[i] = pat;
  if (blockIdx.x == 0) {
    const uint8_t* pb = reinterpret_cast<const uint8_t*>(&pat);
    const size_t tail = head + nvec * 16;
    for (size_t i = threadIdx.x; i < head; i += kFillBlock) a[i] = pb[i % es];
    for (size_t i = tail + threadIdx.x; i < nbytes; i += kFillBlock) a[i] = pb[i % es];
  }
}

}  // namespace pa

using namespace pa;

extern "C" {

int pa_fill(void* a, size_t elem_bytes, const void* value, size_t n, pa_stream_t stream) {
  if (n == 0) return PA_OK;
  if (!a || !value) return arg_error("pa_fill: null pointer");
  if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8 && elem_bytes != 16)
    return arg_error("pa_fill: elem_bytes must be 1, 2, 4, 8 or 16");
  if ((uintptr_t)a % elem_bytes != 0 && elem_bytes <= 8) return arg_error("pa_fill: a is not aligned to elem_bytes");
  if (elem_bytes == 16 && ((uintptr_t)a & 15) != 0) return arg_error("pa_fill: a is not aligned to 16 bytes");
  uint8_t bytes[16];
  const uint8_t* vb = static_cast<const uint8_t*>(value);
  for (int i = 0; i < 16; ++i) bytes[i] = vb[i % elem_bytes];
  uint4 pat;
  __builtin_memcpy(&pat, bytes, 16);
  hipStream_t s = (hipStream_t)stream;
  StreamCtx& ctx = stream_ctx(s);
  std::lock_guard<std::mutex> lk(ctx.mu);
  const size_t nbytes = n * elem_bytes;
  const bool dev = is_device_ptr(a);
  uint8_t* dst = static_cast<uint8_t*>(a);
  if (!dev) {
    // Host memory: fill a device copy, then bring it back (blocking), the
    // same staging as every other entry point.
    void* base;
    PA_TRY(ensure_scratch(ctx, nbytes + 16, &base));
    dst = static_cast<uint8_t*>(base) + ((uintptr_t)a & 15);  // same alignment phase as the host range
  }
  const size_t nvec = nbytes / 16 + 1;
  const unsigned grid =
      (unsigned)std::max<size_t>(1, std::min<size_t>((nvec + kFillBlock - 1) / kFillBlock, kFillMaxGrid));
  {
    ProfScope ps("k_fill_pattern", s);
    hipLaunchKernelGGL(k_fill_pattern, dim3(grid), dim3(kFillBlock), 0, s, dst, nbytes, pat, (uint32_t)elem_bytes);
  }
  PA_HIP(hipGetLastError());
  if (!dev) {
    PA_TRY(stage_to_host(ctx, a, dst, nbytes));
  }
  return PA_OK;
}

}  // extern "C"
