// pa_host_small.h -- the host cut-offs of the drop-in headers: sorts of fewer
// than 100 keys and merges of at most 32,768 keys whose arrays live in host
// memory stay on the calling thread, as in the reference, instead of paying
// staging, a launch and a stream synchronisation.
//
//   - Sorts: RadixSortLsdParallel.h:368,383-386 switches to an insertion sort
//     below 100 keys.
//   - Merges: merge_parallel_L5 merges sequentially at or below its
//     parallel_threshold of 32,768 (ParallelMerge.h:286,296-299).
//
// The results are those of the GPU path: the same ascending order (for
// float and double the IEEE total order the GPU sorts use,
// -NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN) and, in merges, ties
// taken from the first run.  Arrays in device memory always go to the GPU.
#ifndef PA_DROPIN_HOST_SMALL_H
#define PA_DROPIN_HOST_SMALL_H

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "pa_hip.h"

namespace pa_host_small {

constexpr size_t kSortBelow = 100;     // RadixSortLsdParallel.h:368
constexpr size_t kMergeMax = 32768;    // ParallelMerge.h:286

inline bool on_host(const void* p) { return p != nullptr && !pa_is_device_pointer(p); }

// Unsigned key whose order is the GPU sorts' order of T.
template <class T>
inline auto order_key(T x) {
  if constexpr (std::is_same_v<T, float>) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  } else if constexpr (std::is_same_v<T, double>) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
  } else if constexpr (std::is_signed_v<T>) {
    using U = std::make_unsigned_t<T>;
    return (U)((U)x ^ ((U)1 << (8 * sizeof(T) - 1)));
  } else {
    return x;
  }
}

template <class T>
inline bool before(const T& a, const T& b) {
  return order_key(a) < order_key(b);
}

// Stable insertion sort of a[0, n).
template <class T>
inline void insertion_sort(T* a, size_t n) {
  for (size_t i = 1; i < n; ++i) {
    const T v = a[i];
    size_t j = i;
    for (; j > 0 && before(v, a[j - 1]); --j) a[j] = a[j - 1];
    a[j] = v;
  }
}

// dst[0, na + nb) = a[0, na) merged with b[0, nb); ties take a first.
template <class T>
inline void merge(const T* a, size_t na, const T* b, size_t nb, T* dst) {
  size_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) dst[k++] = before(b[j], a[i]) ? b[j++] : a[i++];
  while (i < na) dst[k++] = a[i++];
  while (j < nb) dst[k++] = b[j++];
}

// t[0, n1) and t[n1, n1 + n2) merged in place (through a copy of the first).
template <class T>
inline void merge_in_place(T* t, size_t n1, size_t n2) {
  std::vector<T> first(t, t + n1);
  size_t i = 0, j = n1, k = 0;
  while (i < n1 && j < n1 + n2) t[k++] = before(t[j], first[i]) ? t[j++] : first[i++];
  while (i < n1) t[k++] = first[i++];
}

}  // namespace pa_host_small

#endif  // PA_DROPIN_HOST_SMALL_H
