// SortParallel.h -- drop-in replacement for the reference header of the same
// name (DragonSpit/ParallelAlgorithms): sort_par on an MI355X through
// libpa_hip.so (include/pa_hip.h).
//
// Same namespace, names, signatures and default arguments as the reference
// (SortParallel.h:23-93).  The reference sorts with a parallel merge sort into
// a temporary buffer (or std::sort when that allocation fails); for scalar keys
// the result is the unique ascending order, which is what the GPU radix sort
// produces.  Element types: 32/64-bit integers, float, double.  Arrays may be
// in host or device memory; calls block; failures throw.
//
// Divergences from the reference code, each where it contradicts its own
// comments and benchmark (ParallelMergeSortBenchmark.cpp:61-71):
//  - sort_par(src, l, r, dst, dst_size, srcToDst) passes srcToDst into
//    parallel_merge_sort_hybrid_rh_2's `stable` slot (SortParallel.h:82), so
//    the reference always ends with the result in dst.  Here the result is in
//    dst[l, r) and, when srcToDst is false (the in-place interface), also in
//    src[l, r), which is what the reference's benchmark checks.
//  - sort_par(src, src_size, dst, dst_size, srcToDst) passes src_size - 1 as
//    the exclusive bound (SortParallel.h:92), leaving the last element out of
//    the sort; here all src_size elements are sorted.
#ifndef PA_DROPIN_SORTPARALLEL_H
#define PA_DROPIN_SORTPARALLEL_H

#include <cstddef>
#include <stdexcept>
#include <vector>

#include "ParallelMergeSort.h"
#include "pa_sort_keys.h"

namespace ParallelAlgorithms {

// SortParallel.h:36-50: src[l, r) sorted in place.
template <class _Type>
inline void sort_par(_Type* src, size_t l, size_t r) {
  if (r <= l) return;
  pa_sort_keys_detail::sort_keys(src + l, r - l, "sort_par");
}

// SortParallel.h:23-27.
template <class _Type>
inline void sort_par(_Type* src, size_t src_size) {
  ParallelAlgorithms::sort_par(src, (size_t)0, src_size);
}

// SortParallel.h:53-66.
template <class _Type>
inline void sort_par(std::vector<_Type>& src, size_t l, size_t r) {
  if (r <= l) return;
  pa_sort_keys_detail::sort_keys(src.data() + l, r - l, "sort_par");
}

// SortParallel.h:29-33.
template <class _Type>
inline void sort_par(std::vector<_Type>& src) {
  ParallelAlgorithms::sort_par(src, (size_t)0, src.size());
}

// SortParallel.h:73-83: [l, r), dst must hold at least r elements.
template <class _Type>
inline void sort_par(_Type* src, size_t l, size_t r, _Type* dst, size_t dst_size, bool srcToDst = false) {
  if (!dst) throw std::invalid_argument("dst is null, which is not supported");
  if (dst_size < r) throw std::invalid_argument("dst_size must be larger or equal to r, to be able to return dst[l to r-1]");
  if (r <= l) return;
  pa_sort_keys_detail::sort_inclusive(src, l, r - 1, dst, true, "sort_par");
  if (!srcToDst) pa_sort_keys_detail::check(pa_copy(src + l, dst + l, (r - l) * sizeof(_Type), pa_thread_stream()), "sort_par");
}

// SortParallel.h:89-93.
template <class _Type>
inline void sort_par(_Type* src, size_t src_size, _Type* dst, size_t dst_size, bool srcToDst = false) {
  ParallelAlgorithms::sort_par(src, (size_t)0, src_size, dst, dst_size, srcToDst);
}

}  // namespace ParallelAlgorithms

#endif  // PA_DROPIN_SORTPARALLEL_H
