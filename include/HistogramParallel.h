// HistogramParallel.h -- drop-in replacement for the reference header of the
// same name (DragonSpit/ParallelAlgorithms), computing the histograms on an
// MI355X through libpa_hip.so (include/pa_hip.h).
//
// Same namespace, names, template parameters, bounds conventions, default
// arguments and new[]-allocated return shapes as the reference; the caller
// frees the result exactly as before.  parallelThreshold only steered the
// reference's CPU bisection and has no effect on the counts.  A failure throws
// std::runtime_error; there is no CPU fallback.
#ifndef PA_DROPIN_HISTOGRAMPARALLEL_H
#define PA_DROPIN_HISTOGRAMPARALLEL_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "pa_hip.h"

namespace ParallelAlgorithms {

namespace pa_hist_detail {
inline void check(int status, const char* what) {
  if (status != PA_OK)
    throw std::runtime_error(std::string(what) + ": " + pa_status_string(status) + ": " + pa_last_error_string());
}
constexpr unsigned log2_exact(unsigned long v) { return v <= 1 ? 0 : 1 + log2_exact(v >> 1); }

// Counts of bytes over [l, r) into a new size_t[NumberOfBins].
template <unsigned NumberOfBins>
inline size_t* byte_histogram(const unsigned char* in, size_t l, size_t r, const char* who) {
  static_assert(NumberOfBins == 256, "byte histograms have 256 bins");
  size_t* out = new size_t[NumberOfBins]{};
  if (l >= r) return out;
  uint64_t tmp[256];
  int st = pa_hist_u8(in + l, r - l, tmp, pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  if (st != PA_OK) {
    delete[] out;
    check(st, who);
  }
  for (unsigned b = 0; b < 256; ++b) out[b] = (size_t)tmp[b];
  return out;
}

// Counts of ((in[i] >> shift) & (R-1)) over the inclusive range [l, r].
template <unsigned long R>
inline size_t* digit_histogram(unsigned* in, size_t l, size_t r, unsigned long shift, const char* who) {
  static_assert(R >= 2 && R <= 4096 && (R & (R - 1)) == 0, "R must be a power of two in [2, 4096]");
  size_t* out = new size_t[R]{};
  if (l > r) return out;
  uint64_t* tmp = new uint64_t[R];
  int st = pa_hist_u32(in + l, r - l + 1, (unsigned)shift, log2_exact(R), tmp, pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  if (st != PA_OK) {
    delete[] tmp;
    delete[] out;
    check(st, who);
  }
  for (unsigned long b = 0; b < R; ++b) out[b] = (size_t)tmp[b];
  delete[] tmp;
  return out;
}

// count[numberOfQuantas][256] of byte whichByte over [l, r] (q = index / quantum).
inline size_t** quanta_histogram(unsigned* in, size_t l, size_t r, size_t workQuanta, size_t numberOfQuantas,
                                 unsigned whichByte, const char* who) {
  uint64_t* flat = new uint64_t[numberOfQuantas * 256]();
  int st = PA_OK;
  if (l <= r) {
    st = pa_hist_quanta_u32(in, l, r + 1, workQuanta, numberOfQuantas, whichByte, flat, pa_thread_stream());
    if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  }
  if (st != PA_OK) {
    delete[] flat;
    check(st, who);
  }
  size_t** count = new size_t*[numberOfQuantas];
  for (size_t q = 0; q < numberOfQuantas; ++q) {
    count[q] = new size_t[256];
    for (unsigned b = 0; b < 256; ++b) count[q][b] = (size_t)flat[q * 256 + b];
  }
  delete[] flat;
  return count;
}
}  // namespace pa_hist_detail

// Byte-array histograms over [l, r) (HistogramParallel.h:16-247).
template <unsigned NumberOfBins>
inline size_t* HistogramOneByteComponentParallel(unsigned char inArray[], size_t l, size_t r,
                                                 size_t parallelThreshold = 64 * 1024) {
  (void)parallelThreshold;
  return pa_hist_detail::byte_histogram<NumberOfBins>(inArray, l, r, "HistogramOneByteComponentParallel");
}
template <unsigned NumberOfBins>
inline size_t* HistogramOneByteComponentParallel_4(unsigned char inArray[], size_t l, size_t r,
                                                   size_t parallelThreshold = 64 * 1024) {
  (void)parallelThreshold;
  return pa_hist_detail::byte_histogram<NumberOfBins>(inArray, l, r, "HistogramOneByteComponentParallel_4");
}
template <unsigned NumberOfBins>
inline size_t* HistogramOneByteComponentParallel_2(unsigned char inArray[], size_t l, size_t r,
                                                   size_t parallelThreshold = 64 * 1024) {
  (void)parallelThreshold;
  return pa_hist_detail::byte_histogram<NumberOfBins>(inArray, l, r, "HistogramOneByteComponentParallel_2");
}
template <unsigned NumberOfBins>
inline size_t* HistogramOneByteComponentParallel_3(unsigned char inArray[], size_t l, size_t r,
                                                   size_t parallelThreshold = 64 * 1024) {
  (void)parallelThreshold;
  return pa_hist_detail::byte_histogram<NumberOfBins>(inArray, l, r, "HistogramOneByteComponentParallel_3");
}

// HistogramParallel.h:249-329: counts of the 4 byte digits over [l, r].
// Returns Log2ofPowerOfTwoRadix rows (8 for <256,8>, as the reference
// allocates) of PowerOfTwoRadix counts; rows 4..7 are zero.
template <unsigned PowerOfTwoRadix, unsigned Log2ofPowerOfTwoRadix>
inline size_t** HistogramByteComponentsParallel(unsigned inArray[], size_t l, size_t r,
                                                size_t parallelThreshold = 64 * 1024) {
  static_assert(PowerOfTwoRadix == 256 && Log2ofPowerOfTwoRadix >= 4, "byte digits of 32-bit keys");
  (void)parallelThreshold;
  uint64_t tmp[4 * 256] = {};
  if (l <= r) {
    int st = pa_hist_digits_u32(inArray + l, r - l + 1, tmp, pa_thread_stream());
    if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
    pa_hist_detail::check(st, "HistogramByteComponentsParallel");
  }
  size_t** count = new size_t*[Log2ofPowerOfTwoRadix];
  for (unsigned d = 0; d < Log2ofPowerOfTwoRadix; ++d) {
    count[d] = new size_t[PowerOfTwoRadix];
    for (unsigned b = 0; b < PowerOfTwoRadix; ++b) count[d][b] = d < 4 ? (size_t)tmp[d * 256 + b] : 0;
  }
  return count;
}

// HistogramParallel.h:331-397 (one quantum pass), :399-442 and :444-464.
template <unsigned PowerOfTwoRadix, unsigned Log2ofPowerOfTwoRadix>
inline size_t** HistogramByteComponentsAcrossWorkQuantasQC(unsigned inArray[], size_t l, size_t r, size_t workQuanta,
                                                           size_t numberOfQuantas, unsigned whichByte) {
  static_assert(PowerOfTwoRadix == 256 && Log2ofPowerOfTwoRadix == 8, "8-bit digits");
  return pa_hist_detail::quanta_histogram(inArray, l, r, workQuanta, numberOfQuantas, whichByte,
                                          "HistogramByteComponentsAcrossWorkQuantasQC");
}
template <unsigned PowerOfTwoRadix, unsigned Log2ofPowerOfTwoRadix>
inline size_t** HistogramByteComponentsQCParInner(unsigned inArray[], size_t l, size_t r, size_t workQuanta,
                                                  size_t numberOfQuantas, unsigned whichByte,
                                                  size_t parallelThreshold = 16 * 1024) {
  static_assert(PowerOfTwoRadix == 256 && Log2ofPowerOfTwoRadix == 8, "8-bit digits");
  (void)parallelThreshold;
  return pa_hist_detail::quanta_histogram(inArray, l, r, workQuanta, numberOfQuantas, whichByte,
                                          "HistogramByteComponentsQCParInner");
}
template <unsigned PowerOfTwoRadix, unsigned Log2ofPowerOfTwoRadix>
inline size_t** HistogramByteComponentsQCPar(unsigned* inArray, size_t l, size_t r, size_t workQuanta,
                                             size_t numberOfQuantas, unsigned whichByte,
                                             size_t parallelThreshold = 16 * 1024) {
  static_assert(PowerOfTwoRadix == 256 && Log2ofPowerOfTwoRadix == 8, "8-bit digits");
  (void)parallelThreshold;
  return pa_hist_detail::quanta_histogram(inArray, l, r, workQuanta, numberOfQuantas, whichByte,
                                          "HistogramByteComponentsQCPar");
}

// HistogramParallel.h:535-576: counts of (a >> shiftRight) & (R-1) over [l, r].
template <unsigned long PowerOfTwoRadix, unsigned long Log2ofPowerOfTwoRadix>
inline size_t* HistogramOneByteComponentParallel(unsigned inArray[], size_t l, size_t r, unsigned long shiftRight,
                                                 size_t parallelThreshold = 64 * 1024) {
  (void)parallelThreshold;
  return pa_hist_detail::digit_histogram<PowerOfTwoRadix>(inArray, l, r, shiftRight,
                                                          "HistogramOneByteComponentParallel");
}

// HistogramParallel.h:466-533.  The reference's leaf loop drops one element
// per leaf (:499); this returns the full counts (a documented divergence).
template <unsigned long PowerOfTwoRadix, unsigned long Log2ofPowerOfTwoRadix>
inline size_t* HistogramOneByteComponentParallel_2(unsigned inArray[], size_t l, size_t r, unsigned long shiftRight,
                                                   size_t parallelThreshold = 64 * 1024) {
  (void)parallelThreshold;
  return pa_hist_detail::digit_histogram<PowerOfTwoRadix>(inArray, l, r, shiftRight,
                                                          "HistogramOneByteComponentParallel_2");
}

}  // namespace ParallelAlgorithms

#endif  // PA_DROPIN_HISTOGRAMPARALLEL_H
