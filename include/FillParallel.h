// FillParallel.h -- drop-in replacement for the reference header of the same
// name (DragonSpit/ParallelAlgorithms): parallel_fill on an MI355X through
// libpa_hip.so (include/pa_hip.h, pa_fill).
//
// Same namespace, names, signatures and default arguments as the reference
// (FillParallel.h:22-66): src[l, r) = value, inclusive-left and exclusive-right
// bounds.  Elements of 1, 2, 4, 8 or 16 bytes that are trivially copyable;
// the value's bytes are written as they are.  Host or device memory; blocks;
// throws on failure.
#ifndef PA_DROPIN_FILLPARALLEL_H
#define PA_DROPIN_FILLPARALLEL_H

#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "pa_hip.h"

namespace ParallelAlgorithms {

namespace pa_fill_detail {
template <class T>
inline void fill(T* src, const T& value, size_t l, size_t r, const char* who) {
  static_assert(std::is_trivially_copyable_v<T>, "parallel_fill needs trivially copyable elements");
  static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16,
                "parallel_fill takes elements of 1, 2, 4, 8 or 16 bytes");
  if (r <= l) return;
  int st = pa_fill(src + l, sizeof(T), &value, r - l, pa_thread_stream());
  if (st == PA_OK) st = pa_synchronize(pa_thread_stream());
  if (st != PA_OK)
    throw std::runtime_error(std::string(who) + ": " + pa_status_string(st) + ": " + pa_last_error_string());
}
}  // namespace pa_fill_detail

// FillParallel.h:22-43.
template <class _Type>
inline void parallel_fill(_Type* src, _Type value, size_t l, size_t r, size_t parallel_threshold = 16 * 1024) {
  (void)parallel_threshold;
  pa_fill_detail::fill(src, value, l, r, "parallel_fill");
}

// FillParallel.h:45-66.
inline void parallel_fill(unsigned char* src, unsigned char value, size_t l, size_t r,
                          size_t parallel_threshold = 16 * 1024) {
  (void)parallel_threshold;
  pa_fill_detail::fill(src, value, l, r, "parallel_fill");
}

}  // namespace ParallelAlgorithms

#endif  // PA_DROPIN_FILLPARALLEL_H
