// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin extern "C" wrappers around the few reference headers on the hot path
// that compile as shipped, without TBB or the missing Copy.h:
//   Histogram.h     (HistogramOneComponentOpt, radix select's histogram, :103-133;
//                    HistogramByteComponents, serial 4-digit counts, :5-27)
//   BinarySearch.h  (my_binary_search, :33-47)
//   InsertionSort.h (insertionSortSimilarToSTLnoSelfAssignment, :6-24)
//   RadixSortCommon.h (extractDigit_1, :57-61)
// Built by oracle/Makefile into oracle/_ref/libpa_ref.so with -I pointing at the
// reference tree.  Used only by tests/ to cross-check the oracle's helpers.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include "RadixSortCommon.h"
#include "BinarySearch.h"
#include "InsertionSort.h"
#include "Histogram.h"

extern "C" {
size_t ref_binary_search_u32(unsigned value, const unsigned* a, size_t left, size_t right) {
  return my_binary_search<unsigned>(value, a, left, right);
}
void ref_insertion_sort_u32(unsigned* a, size_t n) { insertionSortSimilarToSTLnoSelfAssignment<unsigned>(a, n); }
void ref_hist_one_component_opt_u32(unsigned* a, size_t l, size_t r, unsigned shift, unsigned bits, uint64_t* out) {
  size_t nb = (size_t)1 << bits;
  size_t* c = new size_t[nb]{};
  HistogramOneComponentOpt(a, l, r, shift, bits, c);
  for (size_t i = 0; i < nb; ++i) out[i] = c[i];
  delete[] c;
}
void ref_hist_byte_components_serial_u32(unsigned* a, size_t l, size_t r, uint64_t* out) {
  size_t* c = HistogramByteComponents(a, l, r);
  for (size_t i = 0; i < 4 * 256; ++i) out[i] = c[i];
  delete[] c;
}
unsigned ref_extract_digit_1_u32(unsigned a, unsigned mask, unsigned shift) { return extractDigit_1<unsigned>(a, mask, shift); }
}
