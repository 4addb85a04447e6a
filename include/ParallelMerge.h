// ParallelMerge.h -- drop-in replacement for the reference header of the same
// name (DragonSpit/ParallelAlgorithms), merging on an MI355X through
// libpa_hip.so (include/pa_hip.h).
//
// Same namespace, names, bounds conventions and default arguments as the
// reference.  Element types: 32- and 64-bit unsigned integers (the key types
// of the GPU path); other types fail to compile with a static_assert rather
// than silently running somewhere else.  Every merge here produces the merged
// run the reference produces; its parallel thresholds and in-place strategies
// (block rotations, memory-pressure checks) only chose how the CPU got there.
// As in the reference, merges of at most 32,768 keys (merge_parallel_L5: its
// parallel_threshold) whose arrays are in host memory run on the calling
// thread; larger ones and any in device memory run on the GPU, on the calling
// thread's stream.  A failure throws std::runtime_error.
#ifndef PA_DROPIN_PARALLELMERGE_H
#define PA_DROPIN_PARALLELMERGE_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "pa_hip.h"
#include "pa_host_small.h"

namespace ParallelAlgorithms {

namespace pa_merge_detail {
template <class T>
constexpr bool supported = std::is_integral<T>::value && std::is_unsigned<T>::value &&
                           (sizeof(T) == 4 || sizeof(T) == 8);

inline void check(int status, const char* what) {
  if (status != PA_OK)
    throw std::runtime_error(std::string(what) + ": " + pa_status_string(status) + ": " + pa_last_error_string());
}

// dst[0 .. na+nb) = merge of a[0..na) and b[0..nb); ties take a first.
// Up to host_max keys with every array in host memory are merged on the
// calling thread (merge_parallel_L5's sequential case, ParallelMerge.h:296-299).
template <class T>
inline void merge(const T* a, size_t na, const T* b, size_t nb, T* dst, const char* who,
                  size_t host_max = pa_host_small::kMergeMax) {
  static_assert(supported<T>, "the GPU merge handles 32/64-bit unsigned keys");
  if (na + nb == 0) return;
  if (na + nb <= host_max && pa_host_small::on_host(dst) && (na == 0 || pa_host_small::on_host(a)) &&
      (nb == 0 || pa_host_small::on_host(b))) {
    pa_host_small::merge(a, na, b, nb, dst);
    return;
  }
  int st;
  if (sizeof(T) == 4)
    st = pa_merge_u32(reinterpret_cast<const uint32_t*>(a), na, reinterpret_cast<const uint32_t*>(b), nb,
                      reinterpret_cast<uint32_t*>(dst), pa_thread_stream());
  else
    st = pa_merge_u64(reinterpret_cast<const uint64_t*>(a), na, reinterpret_cast<const uint64_t*>(b), nb,
                      reinterpret_cast<uint64_t*>(dst), pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  check(st, who);
}

// t[l..m] and t[m+1..r] (inclusive) merged in place.
template <class T>
inline void merge_in_place(T* t, size_t l, size_t m, size_t r, const char* who) {
  static_assert(supported<T>, "the GPU merge handles 32/64-bit unsigned keys");
  if (r + 1 <= l) return;  // empty range (size_t wrap allowed, as in the reference)
  const size_t n1 = m + 1 - l, n2 = r - m;
  if (n1 + n2 <= pa_host_small::kMergeMax && pa_host_small::on_host(t + l)) {
    pa_host_small::merge_in_place(t + l, n1, n2);
    return;
  }
  int st;
  if (sizeof(T) == 4)
    st = pa_merge_inplace_u32(reinterpret_cast<uint32_t*>(t) + l, n1, n2, pa_thread_stream());
  else
    st = pa_merge_inplace_u64(reinterpret_cast<uint64_t*>(t) + l, n1, n2, pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  check(st, who);
}

// Length of the inclusive range [p, r]; an empty range is r = p - 1 (size_t
// wrap, ParallelMerge.h:288-289).
inline size_t incl_len(size_t p, size_t r) { return r - p + 1; }
}  // namespace pa_merge_detail

// Pointer-range merges: [a_start, a_end) and [b_start, b_end) into dst
// (ParallelMerge.h:40-206: merge_ptr, _1, _1_unrolled, _2, _3, _adaptive_2).
template <class _Type>
inline void merge_ptr(const _Type* a_start, const _Type* a_end, const _Type* b_start, const _Type* b_end, _Type* dst) {
  pa_merge_detail::merge(a_start, (size_t)(a_end - a_start), b_start, (size_t)(b_end - b_start), dst, "merge_ptr");
}
template <class _Type>
inline void merge_ptr_1(const _Type* a_start, const _Type* a_end, const _Type* b_start, const _Type* b_end,
                        _Type* dst) {
  pa_merge_detail::merge(a_start, (size_t)(a_end - a_start), b_start, (size_t)(b_end - b_start), dst, "merge_ptr_1");
}
template <class _Type>
inline void merge_ptr_1_unrolled(const _Type* a_start, const _Type* a_end, const _Type* b_start, const _Type* b_end,
                                 _Type* dst) {
  pa_merge_detail::merge(a_start, (size_t)(a_end - a_start), b_start, (size_t)(b_end - b_start), dst,
                         "merge_ptr_1_unrolled");
}
template <class _Type>
inline void merge_ptr_2(const _Type* a_start, const _Type* a_end, const _Type* b_start, const _Type* b_end,
                        _Type* dst) {
  pa_merge_detail::merge(a_start, (size_t)(a_end - a_start), b_start, (size_t)(b_end - b_start), dst, "merge_ptr_2");
}
template <class _Type>
inline void merge_ptr_3(const _Type* a_start, const _Type* a_end, const _Type* b_start, const _Type* b_end, _Type* dst,
                        long threshold = 10 * 1024) {
  (void)threshold;
  pa_merge_detail::merge(a_start, (size_t)(a_end - a_start), b_start, (size_t)(b_end - b_start), dst, "merge_ptr_3");
}
// The reference's merge_ptr_adaptive_2 emits some elements twice when a run
// has fewer than 128 left (ParallelMerge.h:192-193); this returns the merge.
template <class _Type>
inline void merge_ptr_adaptive_2(const _Type* a_start, const _Type* a_end, const _Type* b_start, const _Type* b_end,
                                 _Type* dst) {
  pa_merge_detail::merge(a_start, (size_t)(a_end - a_start), b_start, (size_t)(b_end - b_start), dst,
                         "merge_ptr_adaptive_2");
}

// Divide-and-conquer merges of t[p1..r1] and t[p2..r2] into a[p3..]
// (ParallelMerge.h:212-347).
template <class _Type>
inline void merge_dac(const _Type* t, int p1, int r1, int p2, int r2, _Type* a, int p3) {
  pa_merge_detail::merge(t + p1, (size_t)(r1 - p1 + 1), t + p2, (size_t)(r2 - p2 + 1), a + p3, "merge_dac");
}
template <class _Type>
inline void merge_parallel_L3(_Type* t, int p1, int r1, int p2, int r2, _Type* a, int p3) {
  pa_merge_detail::merge(t + p1, (size_t)(r1 - p1 + 1), t + p2, (size_t)(r2 - p2 + 1), a + p3, "merge_parallel_L3");
}
template <class _Type>
inline void merge_dac_hybrid(const _Type* t, size_t p1, size_t r1, size_t p2, size_t r2, _Type* a, size_t p3) {
  pa_merge_detail::merge(t + p1, pa_merge_detail::incl_len(p1, r1), t + p2, pa_merge_detail::incl_len(p2, r2), a + p3,
                         "merge_dac_hybrid");
}
// ParallelMerge.h:285-315 -- the merge every parallel merge sort calls.
template <class _Type>
inline void merge_parallel_L5(_Type* t, size_t p1, size_t r1, size_t p2, size_t r2, _Type* a, size_t p3,
                              size_t parallel_threshold = 32768) {
  pa_merge_detail::merge(t + p1, pa_merge_detail::incl_len(p1, r1), t + p2, pa_merge_detail::incl_len(p2, r2), a + p3,
                         "merge_parallel_L5", parallel_threshold);
}
template <class _Type>
inline void merge_parallel_quad(_Type* t, size_t p1, size_t r1, size_t p2, size_t r2, _Type* a, size_t p3) {
  pa_merge_detail::merge(t + p1, pa_merge_detail::incl_len(p1, r1), t + p2, pa_merge_detail::incl_len(p2, r2), a + p3,
                         "merge_parallel_quad");
}

// In-place merges of t[l..m] and t[m+1..r] (ParallelMerge.h:400-726).
template <class _Type>
inline void merge_truly_in_place(_Type* t, size_t l, size_t m, size_t r) {
  pa_merge_detail::merge_in_place(t, l, m, r, "merge_truly_in_place");
}
template <class _Type>
inline void merge_in_place(_Type* t, int l, int m, int r) {
  if (r < l) return;
  pa_merge_detail::merge_in_place(t, (size_t)l, (size_t)m, (size_t)r, "merge_in_place");
}
template <class _Type>
inline void p_merge_in_place_2(_Type* t, size_t l, size_t m, size_t r) {
  pa_merge_detail::merge_in_place(t, l, m, r, "p_merge_in_place_2");
}
// mergeInPlace(a, a_size): a[0..a_size) and a[a_size..2*a_size) (:561-573).
template <class _Type>
inline void mergeInPlace(_Type* a, unsigned long a_size) {
  if (a_size == 0) return;
  pa_merge_detail::merge_in_place(a, 0, a_size - 1, 2 * a_size - 1, "mergeInPlace");
}
// mergeInPlace(a, l, m, r): the reference starts at index 0 instead of l
// (:578); this merges [l..m] and [m+1..r] as the signature says.
template <class _Type>
inline void mergeInPlace(_Type* a, int l, int m, int r) {
  if (r < l) return;
  pa_merge_detail::merge_in_place(a, (size_t)l, (size_t)m, (size_t)r, "mergeInPlace");
}
template <class _Type>
inline void p_merge_truly_in_place(_Type* t, size_t l, size_t m, size_t r) {
  pa_merge_detail::merge_in_place(t, l, m, r, "p_merge_truly_in_place");
}
template <class _Type>
inline void p_merge_in_place_adaptive(_Type* src, size_t l, size_t m, size_t r) {
  pa_merge_detail::merge_in_place(src, l, m, r, "p_merge_in_place_adaptive");
}
template <class _Type>
inline void merge_inplace_preventative_adaptive(_Type* src, size_t l, size_t m, size_t r,
                                                double physical_memory_threshold = 0.75) {
  (void)physical_memory_threshold;
  pa_merge_detail::merge_in_place(src, l, m, r, "merge_inplace_preventative_adaptive");
}
template <class _Type>
inline void p_merge_in_place_preventative_adaptive(_Type* src, size_t l, size_t m, size_t r,
                                                   double physical_memory_threshold_post = 0.75) {
  (void)physical_memory_threshold_post;
  pa_merge_detail::merge_in_place(src, l, m, r, "p_merge_in_place_preventative_adaptive");
}

}  // namespace ParallelAlgorithms

#endif  // PA_DROPIN_PARALLELMERGE_H
