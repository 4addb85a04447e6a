// dropin_test.cpp -- the reference's self-checking benchmarks, re-run against
// the drop-in headers (host std::vector inputs, as the reference's callers
// pass them).  Mirrors RadixSortLsdBenchmark.cpp:78-114 (sort, compare with
// std::stable_sort, exit(1) on mismatch), RadixSelectBenchmark.cpp:33-67
// (k = n - 21 vs std::nth_element, then read a[k]) and checks the histogram
// and merge headers against brute force / std::merge.
// Usage: dropin_test [n]   (default 10,000,000; needs a GPU)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "HistogramParallel.h"
#include "ParallelMerge.h"
#include "RadixSelect.h"
#include "RadixSortLsdParallel.h"

static int failures = 0;
#define EXPECT(cond, ...)                  \
  do {                                     \
    if (!(cond)) {                         \
      std::printf("FAIL: " __VA_ARGS__);   \
      std::printf("\n");                   \
      ++failures;                          \
    }                                      \
  } while (0)

int main(int argc, char** argv) {
  const size_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 10000000;
  std::mt19937_64 dist(1234);  // ParallelAlgorithms.cpp:91
  std::vector<unsigned> uints(n);
  for (auto& d : uints) d = static_cast<unsigned>(dist());

  // ---- LSD sort: 3-arg and 2-arg SortRadixPar, TwoPhase (both buffers)
  std::vector<unsigned> sorted_reference(uints);
  std::stable_sort(sorted_reference.begin(), sorted_reference.end());
  for (size_t len : {size_t(0), size_t(1), size_t(99), size_t(100), size_t(101), n}) {
    std::vector<unsigned> a(uints.begin(), uints.begin() + len), tmp(len);
    std::vector<unsigned> expect(a);
    std::stable_sort(expect.begin(), expect.end());
    ParallelAlgorithms::SortRadixPar(a.data(), tmp.data(), a.size());
    EXPECT(a == expect, "SortRadixPar(a, tmp, %zu)", len);
    std::vector<unsigned> b(uints.begin(), uints.begin() + len);
    ParallelAlgorithms::SortRadixPar(b.data(), b.size());
    EXPECT(b == expect, "SortRadixPar(a, %zu)", len);
    std::vector<unsigned> c(uints.begin(), uints.begin() + len), d(len);
    ParallelAlgorithms::RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase(c.data(), d.data(), len);
    EXPECT(c == expect && d == expect, "TwoPhase(%zu): result in both buffers", len);
  }

  // ---- Radix select: k = n - 21 (RadixSelectBenchmark.cpp:37), n/2, n/100
  for (size_t k : {n - 21, n / 2, n / 100}) {
    std::vector<unsigned> a(uints);
    unsigned v = SelectRadix(a.data(), a.size(), k);
    EXPECT(v == sorted_reference[k] && a[k] == v, "SelectRadix(k=%zu)", k);
  }
  {
    std::vector<unsigned> a(uints);
    size_t start = 1000, length = std::min<size_t>(n - start, 50000), k = start + length / 3;
    std::vector<unsigned> part(uints.begin() + start, uints.begin() + start + length);
    std::nth_element(part.begin(), part.begin() + (k - start), part.end());
    EXPECT(SelectRadix(a.data(), start, length, k) == part[k - start], "SelectRadix(start, length, k)");
  }

  // ---- Histograms vs brute force
  {
    size_t** h = ParallelAlgorithms::HistogramByteComponentsParallel<256, 8>(uints.data(), 0, n - 1);
    std::vector<size_t> bf(4 * 256, 0);
    for (unsigned v : uints)
      for (int d = 0; d < 4; ++d) bf[d * 256 + ((v >> (8 * d)) & 0xff)]++;
    bool ok = true;
    for (int d = 0; d < 8; ++d)
      for (int b = 0; b < 256; ++b) ok = ok && h[d][b] == (d < 4 ? bf[d * 256 + b] : 0);
    EXPECT(ok, "HistogramByteComponentsParallel<256,8>");
    for (int d = 0; d < 8; ++d) delete[] h[d];
    delete[] h;
    size_t* top = ParallelAlgorithms::HistogramOneByteComponentParallel<256, 8>(uints.data(), 0, n - 1, 24);
    ok = true;
    for (int b = 0; b < 256; ++b) ok = ok && top[b] == bf[3 * 256 + b];
    EXPECT(ok, "HistogramOneByteComponentParallel<256,8>(shift 24)");
    delete[] top;
    std::vector<unsigned char> bytes(n);
    for (size_t i = 0; i < n; ++i) bytes[i] = (unsigned char)(uints[i] >> 5);
    size_t* hb = ParallelAlgorithms::HistogramOneByteComponentParallel<256>(bytes.data(), 0, n);
    std::vector<size_t> bfb(256, 0);
    for (unsigned char c : bytes) bfb[c]++;
    ok = true;
    for (int b = 0; b < 256; ++b) ok = ok && hb[b] == bfb[b];
    EXPECT(ok, "HistogramOneByteComponentParallel<256>(bytes)");
    delete[] hb;
  }

  // ---- Merge: two sorted halves of one array (merge_parallel_L5 convention)
  {
    size_t n1 = n / 2 + 7;
    std::vector<unsigned> t(uints);
    std::sort(t.begin(), t.begin() + n1);
    std::sort(t.begin() + n1, t.end());
    std::vector<unsigned> expect(n), out(n + 5, 0);
    std::merge(t.begin(), t.begin() + n1, t.begin() + n1, t.end(), expect.begin());
    ParallelAlgorithms::merge_parallel_L5(t.data(), 0, n1 - 1, n1, n - 1, out.data(), 5);
    EXPECT(std::equal(expect.begin(), expect.end(), out.begin() + 5), "merge_parallel_L5");
    std::vector<unsigned> ip(t);
    ParallelAlgorithms::p_merge_in_place_adaptive(ip.data(), 0, n1 - 1, n - 1);
    EXPECT(ip == expect, "p_merge_in_place_adaptive");
  }

  std::printf(failures ? "FAILED (%d)\n" : "ALL OK\n", failures);
  return failures ? 1 : 0;
}
