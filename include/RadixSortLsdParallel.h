// RadixSortLsdParallel.h -- drop-in replacement for the reference header of the
// same name (DragonSpit/ParallelAlgorithms), running the LSD radix sort on an
// MI355X through libpa_hip.so (include/pa_hip.h).
//
// Same namespace, names, template parameters and default arguments as the
// reference.  Arrays may live in host memory (staged to the GPU and back, as
// the reference's callers pass them) or in device memory.  Calls block until
// the result is in place; each host thread uses its own stream
// (pa_thread_stream).  Host arrays of fewer than 100 keys are insertion-sorted
// on the calling thread, as the reference does.  A failure throws
// std::runtime_error; nothing larger falls back to the CPU.
#ifndef PA_DROPIN_RADIXSORTLSDPARALLEL_H
#define PA_DROPIN_RADIXSORTLSDPARALLEL_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "pa_hip.h"
#include "pa_host_small.h"

namespace ParallelAlgorithms {

namespace pa_detail {
inline void check(int status, const char* what) {
  if (status != PA_OK)
    throw std::runtime_error(std::string(what) + ": " + pa_status_string(status) + ": " + pa_last_error_string());
}
// Host-staged calls already return complete; device-pointer calls are
// enqueued on the calling thread's stream, so wait for it to keep the
// blocking contract.
inline void finish(const char* what) { check(pa_synchronize(pa_thread_stream()), what); }
// Fewer than 100 keys in host memory: insertion sort on this thread, as the
// reference does (RadixSortLsdParallel.h:383-386).
template <class T>
inline bool small_on_host(T* a, size_t n) {
  if (n == 0) return true;
  if (n >= pa_host_small::kSortBelow || !pa_host_small::on_host(a)) return false;
  pa_host_small::insertion_sort(a, n);
  return true;
}
}  // namespace pa_detail

// RadixSortLsdParallel.h:340-364.  The reference allocates its own scratch and
// sizes work quanta from the thread count; the GPU uses device scratch and its
// own tiles, so parallelThreshold has no effect on the (identical) result.
inline void SortRadixPar(unsigned* a, size_t a_size, size_t parallelThreshold = 64 * 1024) {
  (void)parallelThreshold;
  if (pa_detail::small_on_host(a, a_size)) return;
  pa_detail::check(pa_sort_u32(a, nullptr, a_size, pa_thread_stream()), "SortRadixPar");
  pa_detail::finish("SortRadixPar");
}

// RadixSortLsdParallel.h:367-387.  Result in a; tmp_work_buff is used as the
// ping-pong buffer when it is device memory (contents unspecified afterwards).
inline void SortRadixPar(unsigned* a, unsigned* tmp_work_buff, size_t a_size, size_t parallelThreshold = 512 * 1024) {
  (void)parallelThreshold;
  if (pa_detail::small_on_host(a, a_size)) return;
  pa_detail::check(pa_sort_u32(a, tmp_work_buff, a_size, pa_thread_stream()), "SortRadixPar");
  pa_detail::finish("SortRadixPar");
}

// Extensions with no reference counterpart: 64-bit keys, float keys (order
// transform, -NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN), and stable
// key-value sorts.
inline void SortRadixPar(uint64_t* a, uint64_t* tmp_work_buff, size_t a_size, size_t parallelThreshold = 512 * 1024) {
  (void)parallelThreshold;
  if (pa_detail::small_on_host(a, a_size)) return;
  pa_detail::check(pa_sort_u64(a, tmp_work_buff, a_size, pa_thread_stream()), "SortRadixPar(u64)");
  pa_detail::finish("SortRadixPar(u64)");
}
inline void SortRadixPar(float* a, float* tmp_work_buff, size_t a_size, size_t parallelThreshold = 512 * 1024) {
  (void)parallelThreshold;
  if (pa_detail::small_on_host(a, a_size)) return;
  pa_detail::check(pa_sort_f32(a, tmp_work_buff, a_size, pa_thread_stream()), "SortRadixPar(float)");
  pa_detail::finish("SortRadixPar(float)");
}
inline void SortRadixParKeyValue(unsigned* keys, unsigned* values, size_t n) {
  pa_detail::check(pa_sort_kv_u32_u32(keys, values, nullptr, nullptr, n, pa_thread_stream()), "SortRadixParKeyValue");
  pa_detail::finish("SortRadixParKeyValue");
}
inline void SortRadixParKeyValue(uint64_t* keys, unsigned* values, size_t n) {
  pa_detail::check(pa_sort_kv_u64_u32(keys, values, nullptr, nullptr, n, pa_thread_stream()), "SortRadixParKeyValue");
  pa_detail::finish("SortRadixParKeyValue");
}

// RadixSortLsdParallel.h:86-106: sorted result in BOTH a and b (:80-82).
inline void RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase(unsigned* a, unsigned* b, unsigned long a_size) {
  pa_detail::check(pa_sort_two_phase_u32(a, b, a_size, pa_thread_stream()), "RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase");
  pa_detail::finish("RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase");
}

// RadixSortLsdParallel.h:491-511.  In the reference this is a no-op for
// a_size >= 100 (its body is commented out, :503); here it sorts like the
// TwoPhase form (result in a and b) -- a documented divergence.
inline void RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase_DeRandomize(unsigned* a, unsigned* b, size_t a_size) {
  pa_detail::check(pa_sort_two_phase_u32(a, b, a_size, pa_thread_stream()),
                   "RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase_DeRandomize");
  pa_detail::finish("RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase_DeRandomize");
}

// RadixSortLsdParallel.h:215-336 (8-bit digits only, as instantiated by the
// reference's callers).  Result in inputArray (4 passes: even).
template <unsigned long PowerOfTwoRadix, unsigned long Log2ofPowerOfTwoRadix>
inline void SortRadixInnerPar(unsigned* inputArray, unsigned* workArray, size_t inputSize,
                              size_t ParallelWorkQuantum = 64 * 1024) {
  static_assert(PowerOfTwoRadix == 256 && Log2ofPowerOfTwoRadix == 8, "the GPU sort uses 8-bit digits");
  (void)ParallelWorkQuantum;
  pa_detail::check(pa_sort_u32(inputArray, workArray, inputSize, pa_thread_stream()), "SortRadixInnerPar");
  pa_detail::finish("SortRadixInnerPar");
}

// RadixSortLsdParallel.h:108-165.  Returns startOfBin[numberOfQuantas][256]
// allocated with new[] (caller deletes each row, then the outer array).
template <unsigned PowerOfTwoRadix, unsigned Log2ofPowerOfTwoRadix>
inline size_t** ComputeStartOfBinsPar(unsigned* inArray, size_t size, size_t workQuanta, size_t numberOfQuantas,
                                      unsigned digit, size_t parallelThreshold = 16 * 1024) {
  static_assert(PowerOfTwoRadix == 256 && Log2ofPowerOfTwoRadix == 8, "8-bit digits");
  (void)parallelThreshold;
  uint64_t* flat = new uint64_t[numberOfQuantas * 256];
  int st = pa_start_of_bins_u32(inArray, size, workQuanta, numberOfQuantas, digit, flat, pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  if (st != PA_OK) {
    delete[] flat;
    pa_detail::check(st, "ComputeStartOfBinsPar");
  }
  size_t** startOfBin = new size_t*[numberOfQuantas];
  for (size_t q = 0; q < numberOfQuantas; q++) {
    startOfBin[q] = new size_t[PowerOfTwoRadix];
    for (unsigned b = 0; b < PowerOfTwoRadix; b++) startOfBin[q][b] = (size_t)flat[q * 256 + b];
  }
  delete[] flat;
  return startOfBin;
}

}  // namespace ParallelAlgorithms

#include "HistogramParallel.h"

#endif  // PA_DROPIN_RADIXSORTLSDPARALLEL_H
