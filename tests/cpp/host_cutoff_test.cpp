// host_cutoff_test.cpp -- the drop-in headers' host cut-offs need no GPU:
// host arrays of fewer than 100 keys are sorted, and host merges of at most
// 32,768 keys are merged, on the calling thread (RadixSortLsdParallel.h:383-386,
// ParallelMerge.h:296-299).  Runs on a machine without a GPU, where any call
// that reached the device would fail and throw.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "ParallelMerge.h"
#include "RadixSortLsdParallel.h"
#include "SortParallel.h"

static int failures = 0;
#define EXPECT(cond, what)               \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: %s\n", what);   \
      ++failures;                        \
    }                                    \
  } while (0)

int main() {
  std::mt19937_64 g(7);
  for (size_t len = 0; len < 100; ++len) {
    std::vector<unsigned> a(len);
    for (auto& x : a) x = (unsigned)(g() % 50);  // many ties
    std::vector<unsigned> e(a), tmp(len);
    std::stable_sort(e.begin(), e.end());
    ParallelAlgorithms::SortRadixPar(a.data(), tmp.data(), a.size());
    EXPECT(a == e, "SortRadixPar below 100 keys");
  }
  {
    std::vector<int> v = {5, -3, 0, -2147483647 - 1, 2147483647, -1, 7};
    std::vector<int> e(v);
    std::sort(e.begin(), e.end());
    ParallelAlgorithms::sort_par(v);
    EXPECT(v == e, "sort_par(vector<int>) signed order");
    const float nan = std::numeric_limits<float>::quiet_NaN();
    std::vector<float> f = {1.0f, -nan, 0.0f, -0.0f, nan, -std::numeric_limits<float>::infinity(), -1.0f};
    ParallelAlgorithms::sort_par(f);
    // -NaN < -inf < -1 < -0 < +0 < 1 < +NaN
    bool ok = std::signbit(f[0]) && std::isnan(f[0]) && std::isinf(f[1]) && f[2] == -1.0f && std::signbit(f[3]) &&
              f[3] == 0.0f && !std::signbit(f[4]) && f[4] == 0.0f && f[5] == 1.0f && std::isnan(f[6]) &&
              !std::signbit(f[6]);
    EXPECT(ok, "sort_par(vector<float>) total order");
  }
  for (size_t n1 : {size_t(0), size_t(1), size_t(1000), size_t(16384)}) {
    for (size_t n2 : {size_t(0), size_t(3), size_t(16384)}) {
      std::vector<unsigned> t(n1 + n2);
      for (auto& x : t) x = (unsigned)(g() % 1000);
      std::sort(t.begin(), t.begin() + n1);
      std::sort(t.begin() + n1, t.end());
      std::vector<unsigned> e(n1 + n2), out(n1 + n2 + 2, 0);
      std::merge(t.begin(), t.begin() + n1, t.begin() + n1, t.end(), e.begin());
      if (n1 + n2 == 0) continue;
      ParallelAlgorithms::merge_parallel_L5(t.data(), 0, n1 - 1, n1, n1 + n2 - 1, out.data(), 2);
      EXPECT(std::equal(e.begin(), e.end(), out.begin() + 2), "merge_parallel_L5 on the host");
      std::vector<unsigned> ip(t);
      ParallelAlgorithms::p_merge_in_place_adaptive(ip.data(), 0, n1 - 1, n1 + n2 - 1);
      EXPECT(ip == e, "p_merge_in_place_adaptive on the host");
    }
  }
  std::printf(failures ? "FAILED (%d)\n" : "ALL OK\n", failures);
  return failures ? 1 : 0;
}
