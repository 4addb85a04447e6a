// pa_sort_keys.h -- shared by the drop-in comparison-sort headers
// (SortParallel.h, ParallelMergeSort.h, RadixSortMsdParallel.h): sorts a range
// of scalar keys on an MI355X through libpa_hip.so (include/pa_hip.h).
//
// The reference's merge sorts, adaptive sorts and MSD sort are templates over
// any comparable element type.  For scalar keys the ascending order is unique,
// so every one of them has the same result as one radix sort of the range:
// 32- and 64-bit unsigned and signed integers, float and double (IEEE total
// order, -NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN, where the reference's
// operator< leaves NaN order undefined).  Other element types do not compile.
#ifndef PA_DROPIN_SORT_KEYS_H
#define PA_DROPIN_SORT_KEYS_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "pa_hip.h"
#include "pa_host_small.h"

namespace pa_sort_keys_detail {

template <class T>
struct unsupported : std::false_type {};

inline void check(int status, const char* what) {
  if (status != PA_OK)
    throw std::runtime_error(std::string(what) + ": " + pa_status_string(status) + ": " + pa_last_error_string());
}

// Sorts a[0, n) ascending in place (host or device memory) and waits; fewer
// than 100 keys in host memory are sorted on the calling thread.
template <class T>
inline void sort_keys(T* a, size_t n, const char* who) {
  if (n == 0) return;
  if (n < pa_host_small::kSortBelow && pa_host_small::on_host(a)) {
    // Same order as the GPU sorts; the reference's sorts also finish small
    // ranges with an insertion sort.
    pa_host_small::insertion_sort(a, n);
    return;
  }
  int st;
  if constexpr (std::is_same_v<T, float>) {
    st = pa_sort_f32(a, nullptr, n, pa_thread_stream());
  } else if constexpr (std::is_same_v<T, double>) {
    st = pa_sort_f64(a, nullptr, n, pa_thread_stream());
  } else if constexpr (std::is_integral_v<T> && !std::is_same_v<T, bool> && sizeof(T) == 4) {
    if constexpr (std::is_signed_v<T>)
      st = pa_sort_i32(reinterpret_cast<int32_t*>(a), nullptr, n, pa_thread_stream());
    else
      st = pa_sort_u32(reinterpret_cast<uint32_t*>(a), nullptr, n, pa_thread_stream());
  } else if constexpr (std::is_integral_v<T> && !std::is_same_v<T, bool> && sizeof(T) == 8) {
    if constexpr (std::is_signed_v<T>)
      st = pa_sort_i64(reinterpret_cast<int64_t*>(a), nullptr, n, pa_thread_stream());
    else
      st = pa_sort_u64(reinterpret_cast<uint64_t*>(a), nullptr, n, pa_thread_stream());
  } else {
    static_assert(unsupported<T>::value,
                  "GPU sorts take 32/64-bit integers, float or double elements");
  }
  check(st, who);
  check(pa_synchronize(pa_thread_stream()), who);
}

// Sorts src[l, r] (inclusive bounds, the reference merge sorts' convention).
// With src_to_dst the result goes to dst[l, r] and src is left as it was (the
// reference leaves it in an unspecified intermediate state); otherwise the
// result is in src[l, r] and dst is not touched.
template <class T>
inline void sort_inclusive(T* src, size_t l, size_t r, T* dst, bool src_to_dst, const char* who) {
  if (r < l) return;
  const size_t n = r - l + 1;
  T* target = src;
  if (src_to_dst) {
    if (!dst) throw std::invalid_argument(std::string(who) + ": dst is null");
    check(pa_copy(dst + l, src + l, n * sizeof(T), pa_thread_stream()), who);
    target = dst;
  }
  sort_keys(target + l, n, who);
}

}  // namespace pa_sort_keys_detail

#endif  // PA_DROPIN_SORT_KEYS_H
