// dropin_test.cpp -- the reference's self-checking benchmarks, re-run against
// the drop-in headers (host std::vector inputs, as the reference's callers
// pass them).  Mirrors RadixSortLsdBenchmark.cpp:78-114 (sort, compare with
// std::stable_sort, exit(1) on mismatch), RadixSelectBenchmark.cpp:33-67
// (k = n - 21 vs std::nth_element, then read a[k]),
// ParallelMergeSortBenchmark.cpp:43-225 (sort_par on doubles and the merge
// sorts on uints vs std::sort), RadixSortMsdBenchmark.cpp:52-67 and
// CountingSortParallelBenchmark.cpp:70-92 (vs std::sort), and checks the
// histogram, merge and fill headers against brute force / std::merge.
// Usage: dropin_test [n]   (default 10,000,000; needs a GPU)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "CountingSortParallel.h"
#include "FillParallel.h"
#include "HistogramParallel.h"
#include "ParallelMerge.h"
#include "ParallelMergeSort.h"
#include "RadixPartition.h"
#include "RadixSelect.h"
#include "RadixSortLsdParallel.h"
#include "RadixSortMsdParallel.h"
#include "SortParallel.h"

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

  // ---- Radix partition: the reference's self-check (RadixPartitionBenchmark.cpp:50-66)
  // compares a[k] with nth_element; also the partition property and the
  // permutation, for one rank, a sub-range and several ranks.
  {
    std::vector<unsigned> a(uints);
    const size_t k = n - 21;
    const unsigned v = PartitionRadix(a.data(), a.size(), k);
    bool side = true;
    for (size_t i = 0; i < n; ++i) side = side && (i < k ? a[i] <= v : a[i] >= v);
    std::vector<unsigned> sa(a);
    std::sort(sa.begin(), sa.end());
    EXPECT(v == sorted_reference[k] && a[k] == v && side && sa == sorted_reference, "PartitionRadix(k=%zu)", k);
  }
  {
    std::vector<unsigned> a(uints);
    size_t start = 777, length = std::min<size_t>(n - start, 100000), k = start + length / 2;
    std::vector<unsigned> part(uints.begin() + start, uints.begin() + start + length);
    std::nth_element(part.begin(), part.begin() + (k - start), part.end());
    const unsigned v = PartitionRadix(a.data(), start, length, k);
    bool outside = std::equal(a.begin(), a.begin() + start, uints.begin()) &&
                   std::equal(a.begin() + start + length, a.end(), uints.begin() + start + length);
    EXPECT(v == part[k - start] && a[k] == v && outside, "PartitionRadix(start, length, k)");
  }
  {
    std::vector<unsigned> a(uints);
    size_t ks[3] = {n / 100, n / 2, n - 1};
    PartitionRadix(a.data(), a.size(), ks, 3);
    EXPECT(a[ks[0]] == sorted_reference[ks[0]] && a[ks[1]] == sorted_reference[ks[1]] &&
               a[ks[2]] == sorted_reference[ks[2]],
           "PartitionRadix(k[3])");
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

  // ---- Merge sorts and sort_par (ParallelMergeSortBenchmark.cpp:43-225)
  {
    std::vector<double> doubles(n);
    std::mt19937_64 gen(5678);
    std::uniform_real_distribution<double> ud(-1e6, 1e6);
    for (auto& x : doubles) x = ud(gen);
    std::vector<double> expect(doubles);
    std::sort(expect.begin(), expect.end());
    std::vector<double> src(doubles), dst(n);
    ParallelAlgorithms::sort_par(src.data(), src.size(), dst.data(), dst.size(), false);  // :61
    EXPECT(src == expect && dst == expect, "sort_par(doubles, dst, srcToDst=false)");
    std::vector<double> src2(doubles), dst2(n, 0.0);
    ParallelAlgorithms::sort_par(src2.data(), src2.size(), dst2.data(), dst2.size(), true);
    EXPECT(dst2 == expect && src2 == doubles, "sort_par(doubles, dst, srcToDst=true)");
    std::vector<double> v(doubles);
    ParallelAlgorithms::sort_par(v);
    EXPECT(v == expect, "sort_par(vector<double>)");
    std::vector<double> part(doubles);
    ParallelAlgorithms::sort_par(part.data(), (size_t)10, n - 10);
    std::vector<double> pe(doubles);
    std::sort(pe.begin() + 10, pe.end() - 10);
    EXPECT(part == pe, "sort_par(src, l, r)");

    std::vector<unsigned> u(uints), sorted(n);
    ParallelAlgorithms::parallel_merge_sort_hybrid_radix(u.data(), 0, n - 1, sorted.data(), false);  // :124
    EXPECT(u == sorted_reference, "parallel_merge_sort_hybrid_radix(srcToDst=false)");
    std::vector<unsigned> u2(uints), sorted2(n);
    ParallelAlgorithms::parallel_merge_sort_hybrid(u2.data(), (size_t)0, n - 1, sorted2.data(), true);
    EXPECT(sorted2 == sorted_reference, "parallel_merge_sort_hybrid(srcToDst=true)");
    std::vector<unsigned> u3(uints);
    ParallelAlgorithms::merge_sort_inplace_hybrid_with_insertion(u3.data(), 0, n - 1);  // :171
    EXPECT(u3 == sorted_reference, "merge_sort_inplace_hybrid_with_insertion");
    std::vector<unsigned> u4(uints);
    ParallelAlgorithms::parallel_linear_in_place_preventative_adaptive_sort(u4.data(), u4.size());
    EXPECT(u4 == sorted_reference, "parallel_linear_in_place_preventative_adaptive_sort");
    std::vector<int> ints(n);
    for (size_t i = 0; i < n; ++i) ints[i] = static_cast<int>(uints[i]);
    std::vector<int> ie(ints), is(ints), aux(n);
    std::sort(ie.begin(), ie.end());
    ParallelAlgorithms::parallel_merge_sort_pseudo_inplace(is.data(), 0, (int)n - 1, aux.data());
    EXPECT(is == ie, "parallel_merge_sort_pseudo_inplace(int)");
    std::vector<long long> ll(n);
    for (size_t i = 0; i < n; ++i) ll[i] = static_cast<long long>(uints[i]) * 4099 - (1ll << 40);
    std::vector<long long> lle(ll), lls(ll);
    std::sort(lle.begin(), lle.end());
    ParallelAlgorithms::merge_sort_bottom_up_inplace(lls.data(), 0, n);
    EXPECT(lls == lle, "merge_sort_bottom_up_inplace(long long)");
  }

  // ---- MSD sort (RadixSortMsdBenchmark.cpp:52-67)
  {
    std::vector<unsigned> a(uints);
    ParallelAlgorithms::parallel_hybrid_inplace_msd_radix_sort(a.data(), a.size());
    EXPECT(a == sorted_reference, "parallel_hybrid_inplace_msd_radix_sort");
  }

  // ---- Byte counting sort (CountingSortParallelBenchmark.cpp:70-92) and fill
  {
    std::vector<unsigned char> bytes(n);
    for (size_t i = 0; i < n; ++i) bytes[i] = static_cast<unsigned char>(uints[i]);
    std::vector<unsigned char> expect(bytes);
    std::sort(expect.begin(), expect.end());
    std::vector<unsigned char> a(bytes);
    ParallelAlgorithms::counting_sort_parallel(a.data(), a.size());
    EXPECT(a == expect, "counting_sort_parallel");
    std::vector<unsigned char> b(bytes);
    ParallelAlgorithms::counting_sort_parallel_inner<256>(b.data(), 3, n - 5);
    std::vector<unsigned char> be(bytes);
    std::sort(be.begin() + 3, be.end() - 5);
    EXPECT(b == be, "counting_sort_parallel_inner(l = 3)");
    std::vector<unsigned> f(uints);
    ParallelAlgorithms::parallel_fill(f.data(), 0xDEADBEEFu, 7, n - 3);
    bool ok = std::equal(f.begin(), f.begin() + 7, uints.begin()) && std::equal(f.end() - 3, f.end(), uints.end() - 3);
    for (size_t i = 7; i < n - 3 && ok; ++i) ok = f[i] == 0xDEADBEEFu;
    EXPECT(ok, "parallel_fill(unsigned)");
    std::vector<unsigned char> fb(bytes);
    ParallelAlgorithms::parallel_fill(fb.data(), (unsigned char)0x5A, 1, n);
    ok = fb[0] == bytes[0];
    for (size_t i = 1; i < n && ok; ++i) ok = fb[i] == 0x5A;
    EXPECT(ok, "parallel_fill(unsigned char)");
  }

  std::printf(failures ? "FAILED (%d)\n" : "ALL OK\n", failures);
  return failures ? 1 : 0;
}
