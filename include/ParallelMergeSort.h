// ParallelMergeSort.h -- drop-in replacement for the reference header of the
// same name (DragonSpit/ParallelAlgorithms): the merge-sort family on an
// MI355X through libpa_hip.so (include/pa_hip.h).
//
// Same namespace, names, signatures and default arguments as the reference
// (ParallelMergeSort.h:53-638).  The reference's sorts recurse over TBB tasks
// and merge with merge_parallel_L5; for scalar keys every one of them yields
// the unique ascending order of its range, which is what one GPU radix sort of
// the range produces (pa_sort_keys.h).  The radix-leaf family
// (parallel_merge_sort_hybrid_radix and its _inner and _single_buffer forms)
// keeps the reference's shape instead: leaves radix-sorted, merged upward
// (pa_merge_sort_u32).  Element types: 32/64-bit integers,
// float, double.  Thresholds, stability flags and memory-fraction arguments
// steer the reference's CPU recursion and have no effect on that result.
//
// Bounds: l and r are inclusive ([l, r]), as in the reference; start/length
// forms cover [start, start + length).  srcToDst = true puts the result in
// dst[l, r] (src keeps its input; the reference leaves it scrambled), false in
// src[l, r] (dst untouched; the reference uses it as scratch).
//
// The merge primitives (merge_parallel_L5 and the rest) come with this header
// as they do in the reference, from the drop-in ParallelMerge.h.
#ifndef PA_DROPIN_PARALLELMERGESORT_H
#define PA_DROPIN_PARALLELMERGESORT_H

#include <cstddef>
#include <stdexcept>
#include <string>
#include <thread>

#include "ParallelMerge.h"
#include "RadixSortLsdParallel.h"
#include "RadixSortMsdParallel.h"
#include "pa_sort_keys.h"

namespace ParallelAlgorithms {

namespace pa_msort_detail {
using pa_sort_keys_detail::sort_inclusive;
// The radix-leaf merge sort on the GPU (pa_merge_sort_u32): leaves of `leaf`
// keys radix-sorted, then merged upward.  src[l, r] inclusive; the result in
// dst[l, r] (src kept) for src_to_dst, else in src[l, r] with dst[l, r] as
// the ping-pong buffer, as the reference uses it.
inline void merge_sort_radix_leaves(unsigned* src, size_t l, size_t r, unsigned* dst, bool src_to_dst, size_t leaf,
                                    const char* who) {
  if (r < l) return;
  const size_t n = r - l + 1;
  if (n < pa_host_small::kSortBelow && pa_host_small::on_host(src) && (!src_to_dst || pa_host_small::on_host(dst))) {
    unsigned* t = src;
    if (src_to_dst) {
      for (size_t i = l; i <= r; ++i) dst[i] = src[i];
      t = dst;
    }
    pa_host_small::insertion_sort(t + l, n);
    return;
  }
  unsigned* target = src;
  unsigned* work = dst ? dst + l : nullptr;
  if (src_to_dst) {
    if (!dst) throw std::invalid_argument(std::string(who) + ": dst is null");
    pa_sort_keys_detail::check(pa_copy(dst + l, src + l, n * sizeof(unsigned), pa_thread_stream()), who);
    target = dst;
    work = nullptr;  // src keeps its input
  }
  pa_sort_keys_detail::check(pa_merge_sort_u32(target + l, work, n, leaf == 0 ? 1 : leaf, pa_thread_stream()), who);
  pa_sort_keys_detail::check(pa_synchronize(pa_thread_stream()), who);
}
// The reference's leaf size for n keys: parallelThreshold, raised to
// n / hardware threads when that is larger (ParallelMergeSort.h:288-292).
inline size_t leaf_keys(size_t n, size_t threshold) {
  const size_t p = std::thread::hardware_concurrency();
  return (p > 0 && threshold * p < n) ? n / p : threshold;
}
// The int-bounded forms allow r = -1 for an empty range (ParallelMergeSort.h:74-77).
template <class T>
inline void sort_int_bounds(T* src, int l, int r, T* dst, bool src_to_dst, const char* who) {
  if (r < l) return;
  sort_inclusive(src, (size_t)l, (size_t)r, dst, src_to_dst, who);
}
}  // namespace pa_msort_detail

// ---- not-in-place merge sorts (src, l, r, dst, srcToDst) -------------------

// ParallelMergeSort.h:53-71 (Listing 1).
template <class _Type>
inline void parallel_merge_sort_simplest_r(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true) {
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "parallel_merge_sort_simplest_r");
}

// ParallelMergeSort.h:73-78.
template <class _Type>
inline void parallel_merge_sort_simplest(_Type* src, int l, int r, _Type* dst, bool srcToDst = true) {
  pa_msort_detail::sort_int_bounds(src, l, r, dst, srcToDst, "parallel_merge_sort_simplest");
}

// ParallelMergeSort.h:81-85 (Listing 2): result in dst.
template <class _Type>
inline void parallel_merge_sort(_Type* src, int l, int r, _Type* dst) {
  pa_msort_detail::sort_int_bounds(src, l, r, dst, true, "parallel_merge_sort");
}

// ParallelMergeSort.h:87-91: result in srcDst, aux is scratch.
template <class _Type>
inline void parallel_merge_sort_pseudo_inplace(_Type* srcDst, int l, int r, _Type* aux) {
  pa_msort_detail::sort_int_bounds(srcDst, l, r, aux, false, "parallel_merge_sort_pseudo_inplace");
}

// ParallelMergeSort.h:94-119 (Listing 3).
template <class _Type>
inline void parallel_merge_sort_hybrid_rh(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true) {
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "parallel_merge_sort_hybrid_rh");
}

// ParallelMergeSort.h:122-145 (Listing 4).
template <class _Type>
inline void parallel_merge_sort_hybrid_rh_1(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true) {
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "parallel_merge_sort_hybrid_rh_1");
}

// ParallelMergeSort.h:147-176.
template <class _Type>
inline void parallel_merge_sort_hybrid_rh_2(_Type* src, size_t l, size_t r, _Type* dst, bool stable = true,
                                            bool srcToDst = true, size_t parallelThreshold = 32 * 1024) {
  (void)stable;
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "parallel_merge_sort_hybrid_rh_2");
}

// ParallelMergeSort.h:180-200 (serial in the reference).
template <class _Type>
inline void merge_sort_hybrid(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true) {
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "merge_sort_hybrid");
}

// ParallelMergeSort.h:203-226.
template <class _Type>
inline void parallel_merge_merge_sort_hybrid_inner(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true,
                                                   size_t parallelThreshold = 32 * 1024) {
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "parallel_merge_merge_sort_hybrid_inner");
}

// ParallelMergeSort.h:228-239.
template <class _Type>
inline void parallel_merge_merge_sort_hybrid(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true,
                                             size_t parallelThreshold = 32 * 1024) {
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "parallel_merge_merge_sort_hybrid");
}

// ParallelMergeSort.h:241-253.
template <class _Type>
inline void parallel_merge_sort_hybrid(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true,
                                       size_t parallelThreshold = 16 * 1024) {
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "parallel_merge_sort_hybrid");
}

// ParallelMergeSort.h:255-283: radix-sorted leaves of parallelThreshold keys,
// merged upward -- on the GPU, the leaves as one segmented batch and the
// merges as merge-path rounds (pa_merge_sort_u32).
inline void parallel_merge_sort_hybrid_radix_inner(unsigned* src, size_t l, size_t r, unsigned* dst,
                                                   bool srcToDst = true, size_t parallelThreshold = 32 * 1024) {
  pa_msort_detail::merge_sort_radix_leaves(src, l, r, dst, srcToDst, parallelThreshold,
                                           "parallel_merge_sort_hybrid_radix_inner");
}

// ParallelMergeSort.h:285-295: the same, with the leaf size raised to
// n / hardware threads.
inline void parallel_merge_sort_hybrid_radix(unsigned* src, size_t l, size_t r, unsigned* dst, bool srcToDst = true,
                                             size_t parallelThreshold = 24 * 1024) {
  const size_t n = r >= l ? r - l + 1 : 0;
  pa_msort_detail::merge_sort_radix_leaves(src, l, r, dst, srcToDst, pa_msort_detail::leaf_keys(n, parallelThreshold),
                                           "parallel_merge_sort_hybrid_radix");
}

// ParallelMergeSort.h:297-307.
inline void parallel_merge_sort_hybrid_radix_single_buffer(unsigned* src, size_t l, size_t r, unsigned* dst,
                                                           bool srcToDst = true, size_t parallelThreshold = 24 * 1024) {
  const size_t n = r >= l ? r - l + 1 : 0;
  pa_msort_detail::merge_sort_radix_leaves(src, l, r, dst, srcToDst, pa_msort_detail::leaf_keys(n, parallelThreshold),
                                           "parallel_merge_sort_hybrid_radix_single_buffer");
}

// ParallelMergeSort.h:310-325 (serial in the reference).
template <class _Type>
inline void merge_sort(_Type* src, size_t l, size_t r, _Type* dst, bool srcToDst = true) {
  pa_sort_keys_detail::sort_inclusive(src, l, r, dst, srcToDst, "merge_sort");
}

// ---- in-place merge sorts (src, l, r): result in src[l, r] -----------------

// ParallelMergeSort.h:327-346.
template <class _Type>
inline void merge_sort_inplace_hybrid_with_sort(_Type* src, size_t l, size_t r, bool stable = false,
                                                size_t threshold = 1024) {
  (void)stable;
  (void)threshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "merge_sort_inplace_hybrid_with_sort");
}

// ParallelMergeSort.h:348-390.
template <class _Type>
inline void parallel_inplace_merge_sort_hybrid_inner(_Type* src, size_t l, size_t r, bool stable = false,
                                                     size_t parallelThreshold = 1024) {
  (void)stable;
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "parallel_inplace_merge_sort_hybrid_inner");
}

// ParallelMergeSort.h:392-403.
template <class _Type>
inline void parallel_inplace_merge_sort_hybrid(_Type* src, size_t l, size_t r, bool stable = false,
                                               size_t parallelThreshold = 24 * 1024) {
  (void)stable;
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "parallel_inplace_merge_sort_hybrid");
}

// ParallelMergeSort.h:405-430 (MSD-radix-sorted leaves, merged in place).
template <class _Type>
inline void parallel_inplace_merge_sort_radix_hybrid_inner(_Type* src, size_t l, size_t r,
                                                           size_t parallelThreshold = 1024) {
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "parallel_inplace_merge_sort_radix_hybrid_inner");
}

// ParallelMergeSort.h:432-448 (serial in the reference).
template <class _Type>
inline void preventative_adaptive_inplace_merge_sort(_Type* src, size_t l, size_t r,
                                                     double physical_memory_threshold = 0.75, size_t threshold = 48) {
  (void)physical_memory_threshold;
  (void)threshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "preventative_adaptive_inplace_merge_sort");
}

// ParallelMergeSort.h:450-470.
template <class _Type>
inline void parallel_preventative_adaptive_inplace_merge_sort(_Type* src, size_t l, size_t r,
                                                              double physical_memory_threshold = 0.75,
                                                              size_t parallelThreshold = 48) {
  (void)physical_memory_threshold;
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "parallel_preventative_adaptive_inplace_merge_sort");
}

// ParallelMergeSort.h:472-495.
template <class _Type>
inline void parallel_preventative_adaptive_inplace_merge_sort(_Type* src, size_t l, size_t r, bool stable = false,
                                                              double physical_memory_threshold = 0.75,
                                                              size_t parallelThreshold = 48) {
  (void)stable;
  (void)physical_memory_threshold;
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "parallel_preventative_adaptive_inplace_merge_sort");
}

// ParallelMergeSort.h:501-534.
template <class _Type>
inline void parallel_preventative_adaptive_inplace_merge_sort_2(_Type* src, size_t l, size_t r,
                                                                double physical_memory_threshold_post = 0.75,
                                                                size_t parallelThreshold = 48) {
  (void)physical_memory_threshold_post;
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "parallel_preventative_adaptive_inplace_merge_sort_2");
}

// ParallelMergeSort.h:536-547.
template <class _Type>
inline void parallel_inplace_merge_sort_radix_hybrid(_Type* src, size_t l, size_t r,
                                                     size_t parallelThreshold = 24 * 1024) {
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "parallel_inplace_merge_sort_radix_hybrid");
}

// ParallelMergeSort.h:549-573: all src_size elements.
inline void parallel_linear_in_place_preventative_adaptive_sort(unsigned* src, size_t src_size, bool stable = true,
                                                                double physical_memory_threshold_post = 0.75,
                                                                size_t parallelThreshold = 24 * 1024) {
  (void)stable;
  (void)physical_memory_threshold_post;
  (void)parallelThreshold;
  pa_sort_keys_detail::sort_keys(src, src_size, "parallel_linear_in_place_preventative_adaptive_sort");
}

// ParallelMergeSort.h:575-586 (serial in the reference).
template <class _Type>
inline void merge_sort_inplace(_Type* src, size_t l, size_t r) {
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "merge_sort_inplace");
}

// ParallelMergeSort.h:588-598: [start, start + length).
template <class _Type>
inline void merge_sort_bottom_up_inplace(_Type* src, size_t start, size_t length) {
  pa_sort_keys_detail::sort_keys(src + start, length, "merge_sort_bottom_up_inplace");
}

// ParallelMergeSort.h:600-615.
template <class _Type>
inline void merge_sort_inplace_hybrid_with_insertion(_Type* src, size_t l, size_t r) {
  pa_sort_keys_detail::sort_inclusive(src, l, r, src, false, "merge_sort_inplace_hybrid_with_insertion");
}

// ParallelMergeSort.h:618-638: [start, start + length).
template <class _Type>
inline void merge_sort_bottom_up_inplace_hybrid(_Type* src, size_t start, size_t length) {
  pa_sort_keys_detail::sort_keys(src + start, length, "merge_sort_bottom_up_inplace_hybrid");
}

}  // namespace ParallelAlgorithms

#endif  // PA_DROPIN_PARALLELMERGESORT_H
