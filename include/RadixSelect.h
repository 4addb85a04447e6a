// RadixSelect.h -- drop-in replacement for the reference header of the same
// name (DragonSpit/ParallelAlgorithms), selecting on an MI355X through
// libpa_hip.so (include/pa_hip.h).
//
// Same global-namespace names and signatures as the reference (RadixSelect.h:
// 139-168).  Returns the k-th smallest element (0-based; k is an absolute
// index into the array), and a[k] holds that value afterwards, as the
// reference's callers expect (RadixSelectBenchmark.cpp:57-58).  The reference
// leaves the rest of [start, start+length) reordered and partly overwritten
// (RadixSelect.h:52); here the rest of the array is left as it was.  A failure
// throws; there is no CPU fallback.
#ifndef PA_DROPIN_RADIXSELECT_H
#define PA_DROPIN_RADIXSELECT_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "pa_hip.h"

namespace pa_select_detail {
inline unsigned select(unsigned a[], size_t start, size_t length, size_t k) {
  if (k < start || k >= start + length) throw std::out_of_range("SelectRadix: k must lie in [start, start + length)");
  uint32_t value = 0;
  int st = pa_select_radix_u32(a + start, length, k - start, &value, pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  if (st != PA_OK)
    throw std::runtime_error(std::string("SelectRadix: ") + pa_status_string(st) + ": " + pa_last_error_string());
  return value;
}
}  // namespace pa_select_detail

// RadixSelect.h:139-150.
inline unsigned SelectRadix(unsigned arrayToBeSelected[], size_t start, size_t length, size_t k) {
  return pa_select_detail::select(arrayToBeSelected, start, length, k);
}

// RadixSelect.h:157-168.
inline unsigned SelectRadix(unsigned arrayToBeSelected[], size_t length, size_t k) {
  return pa_select_detail::select(arrayToBeSelected, 0, length, k);
}

#endif  // PA_DROPIN_RADIXSELECT_H
