// RadixSortMsdParallel.h -- drop-in replacement for the reference header of the
// same name (DragonSpit/ParallelAlgorithms): the in-place MSD radix sort on an
// MI355X through libpa_hip.so (include/pa_hip.h).
//
// Reference: parallel_hybrid_inplace_msd_radix_sort (RadixSortMsdParallel.h:
// 291-317), an unstable in-place MSD sort of u32 keys with TBB tasks per bin.
// Sorting scalar keys ascending has one result, so the GPU's LSD radix sort of
// the array is that result.  The reference's internal recursion helpers are
// not part of the drop-in.  Host or device memory; blocks; throws on failure.
#ifndef PA_DROPIN_RADIXSORTMSDPARALLEL_H
#define PA_DROPIN_RADIXSORTMSDPARALLEL_H

#include <cstddef>

#include "pa_sort_keys.h"

namespace ParallelAlgorithms {

// RadixSortMsdParallel.h:291-317.
inline void parallel_hybrid_inplace_msd_radix_sort(unsigned* a, size_t a_size) {
  if (a_size < 2) return;
  pa_sort_keys_detail::sort_keys(a, a_size, "parallel_hybrid_inplace_msd_radix_sort");
}

}  // namespace ParallelAlgorithms

#endif  // PA_DROPIN_RADIXSORTMSDPARALLEL_H
