// RadixPartition.h -- drop-in replacement for the reference header of the
// same name (DragonSpit/ParallelAlgorithms), partitioning on an MI355X through
// libpa_hip.so (include/pa_hip.h).
//
// Same global-namespace names and signatures as the reference
// (RadixPartition.h:94-184).  Afterwards a[k] holds the k-th smallest key of
// [aStart, aStart + aLength) (k is an absolute index, as in the reference),
// nothing before it is larger and nothing after it is smaller, and the range
// is a permutation of its input; for several ranks this holds for every one
// of them.  The order of the other keys differs from the reference's, whose
// in-place swap order is not part of that contract.  A failure throws; there
// is no CPU fallback.
#ifndef PA_DROPIN_RADIXPARTITION_H
#define PA_DROPIN_RADIXPARTITION_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pa_hip.h"

namespace pa_partition_detail {
inline void check(int st, const char* who) {
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  if (st != PA_OK) throw std::runtime_error(std::string(who) + ": " + pa_status_string(st) + ": " + pa_last_error_string());
}
inline unsigned one(unsigned a[], size_t start, size_t length, size_t k) {
  if (k < start || k >= start + length) throw std::out_of_range("PartitionRadix: k must lie in [aStart, aStart + aLength)");
  uint32_t value = 0;
  check(pa_partition_u32(a + start, length, k - start, &value, pa_thread_stream()), "PartitionRadix");
  return value;
}
inline void many(unsigned a[], size_t start, size_t length, const size_t k[], size_t k_start, size_t k_length) {
  std::vector<size_t> ks(k_length);
  for (size_t i = 0; i < k_length; ++i) {
    const size_t kk = k[k_start + i];
    if (kk < start || kk >= start + length) throw std::out_of_range("PartitionRadix: k must lie in [aStart, aStart + aLength)");
    ks[i] = kk - start;
  }
  check(pa_partition_multi_u32(a + start, length, ks.data(), ks.size(), pa_thread_stream()), "PartitionRadix");
}
}  // namespace pa_partition_detail

// RadixPartition.h:94-113.
inline unsigned PartitionRadix(unsigned arrayToBePartitioned[], size_t aLength, size_t k) {
  return pa_partition_detail::one(arrayToBePartitioned, 0, aLength, k);
}

// RadixPartition.h:122-140.
inline unsigned PartitionRadix(unsigned arrayToBePartitioned[], size_t aStart, size_t aLength, size_t k) {
  return pa_partition_detail::one(arrayToBePartitioned, aStart, aLength, k);
}

// RadixPartition.h:147-163 (k ascending).
inline void PartitionRadix(unsigned arrayToBePartitioned[], size_t aLength, size_t k[], size_t kLength) {
  pa_partition_detail::many(arrayToBePartitioned, 0, aLength, k, 0, kLength);
}

// RadixPartition.h:172-184 (k ascending).
inline void PartitionRadix(unsigned arrayToBePartitioned[], size_t aStart, size_t aLength, size_t k[], size_t kStart,
                           size_t kLength) {
  pa_partition_detail::many(arrayToBePartitioned, aStart, aLength, k, kStart, kLength);
}

#endif  // PA_DROPIN_RADIXPARTITION_H
