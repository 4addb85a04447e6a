// CountingSortParallel.h -- drop-in replacement for the reference header of
// the same name (DragonSpit/ParallelAlgorithms): the byte counting sort on an
// MI355X through libpa_hip.so (include/pa_hip.h, pa_counting_sort_u8).
//
// Same namespace, names, signatures and default arguments as the reference
// (CountingSortParallel.h:29-90): one histogram read of the bytes, then every
// bin's value written over its range.  The thresholds steer the reference's
// CPU task split and have no effect on the result.
//
// Divergence: counting_sort_parallel_inner(a, l, r) in the reference counts
// [l, r) but fills from index 0 (its bin starts begin at 0,
// CountingSortParallel.h:56-59), so for l > 0 it writes [0, r - l) and leaves
// [l, r) unsorted.  Here [l, r) is sorted in place; for l = 0 (the only form
// counting_sort_parallel uses) the two agree.  Arrays of up to 2^32 - 1 bytes.
#ifndef PA_DROPIN_COUNTINGSORTPARALLEL_H
#define PA_DROPIN_COUNTINGSORTPARALLEL_H

#include <cstddef>
#include <stdexcept>
#include <string>

#include "FillParallel.h"
#include "HistogramParallel.h"
#include "RadixSortMsdParallel.h"
#include "pa_hip.h"

namespace ParallelAlgorithms {

// CountingSortParallel.h:29-76.  NumberOfBins is 256 (one byte).
template <unsigned long NumberOfBins>
inline void counting_sort_parallel_inner(unsigned char* array_to_sort, size_t l, size_t r,
                                         size_t threshold_count = 64 * 1024, size_t threshold_fill = 64 * 1024) {
  static_assert(NumberOfBins == 256, "the byte counting sort has 256 bins");
  (void)threshold_count;
  (void)threshold_fill;
  if (r <= l) return;
  int st = pa_counting_sort_u8(array_to_sort + l, r - l, pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  if (st != PA_OK)
    throw std::runtime_error(std::string("counting_sort_parallel: ") + pa_status_string(st) + ": " +
                             pa_last_error_string());
}

// CountingSortParallel.h:78-90.
inline void counting_sort_parallel(unsigned char* a, size_t a_size) {
  if (a_size == 0) return;
  counting_sort_parallel_inner<256>(a, 0, a_size);
}

}  // namespace ParallelAlgorithms

#endif  // PA_DROPIN_COUNTINGSORTPARALLEL_H
