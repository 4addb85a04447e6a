// forwards_test.cpp -- the drop-in forwards the benchmarks do not reach, the
// host cut-offs, and concurrent callers.
//   1. Each forward against a brute-force or std:: result on host vectors:
//      SortRadixInnerPar, ComputeStartOfBinsPar (RadixSortLsdParallel.h:108-165),
//      the TwoPhase_DeRandomize sort, HistogramOneByteComponentParallel_2 on
//      unsigned keys, HistogramByteComponentsQCPar over a range that starts
//      and ends inside a quantum (the l, r + 1 conversion), merge_in_place with
//      int bounds, both mergeInPlace forms, merge_dac and merge_parallel_L3
//      (int bounds), and merge_truly_in_place on 64-bit keys.
//   2. Host cut-offs: sorts below 100 keys and merges up to 32,768 keys of
//      host arrays stay on the calling thread, in the GPU path's order (float
//      total order with NaNs and signed zeros included).
//   3. The radix-leaf merge sorts (parallel_merge_sort_hybrid_radix and its
//      _inner and _single_buffer forms) in both result directions.
//   4. merge_parallel_L5 and SortRadixPar called from 8 host threads at once
//      (the reference calls merge_parallel_L5 from many TBB tasks,
//      ParallelMergeSort.h:117-118), GPU-sized and host-sized, every result
//      checked against std::merge / std::sort.
// Usage: forwards_test   (needs a GPU)
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

#include "HistogramParallel.h"
#include "ParallelMerge.h"
#include "ParallelMergeSort.h"
#include "RadixSortLsdParallel.h"
#include "SortParallel.h"

static std::atomic<int> failures{0};
#define EXPECT(cond, ...)                  \
  do {                                     \
    if (!(cond)) {                         \
      std::printf("FAIL: " __VA_ARGS__);   \
      std::printf("\n");                   \
      ++failures;                          \
    }                                      \
  } while (0)

static std::vector<unsigned> rand_u32(size_t n, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::vector<unsigned> v(n);
  for (auto& x : v) x = (unsigned)g();
  return v;
}

static void forwards() {
  const size_t n = 3000017;
  const std::vector<unsigned> u = rand_u32(n, 4242);
  std::vector<unsigned> sorted(u);
  std::sort(sorted.begin(), sorted.end());

  {
    std::vector<unsigned> a(u), w(n);
    ParallelAlgorithms::SortRadixInnerPar<256, 8>(a.data(), w.data(), n);
    EXPECT(a == sorted, "SortRadixInnerPar<256,8>");
  }
  {
    std::vector<unsigned> a(u), b(n, 0);
    ParallelAlgorithms::RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase_DeRandomize(a.data(), b.data(), n);
    EXPECT(a == sorted && b == sorted, "TwoPhase_DeRandomize: result in both buffers");
  }
  {
    // startOfBin[q][b] = (keys in bins below b) + (keys of bin b in quanta below q).
    const size_t quantum = 100000, nq = (n + quantum - 1) / quantum;
    for (unsigned digit : {0u, 3u}) {
      std::vector<size_t> cnt(nq * 256, 0), total(256, 0);
      for (size_t i = 0; i < n; ++i) {
        const unsigned b = (u[i] >> (8 * digit)) & 255;
        cnt[(i / quantum) * 256 + b]++;
        total[b]++;
      }
      std::vector<unsigned> in(u);
      size_t** sob = ParallelAlgorithms::ComputeStartOfBinsPar<256, 8>(in.data(), n, quantum, nq, digit);
      bool ok = true;
      size_t below = 0;
      for (unsigned b = 0; b < 256; ++b) {
        size_t run = below;
        for (size_t q = 0; q < nq; ++q) {
          ok = ok && sob[q][b] == run;
          run += cnt[q * 256 + b];
        }
        below += total[b];
      }
      EXPECT(ok, "ComputeStartOfBinsPar<256,8>(digit %u)", digit);
      for (size_t q = 0; q < nq; ++q) delete[] sob[q];
      delete[] sob;
    }
  }
  {
    std::vector<unsigned> in(u);
    const size_t l = 12345, r = n - 777;
    for (unsigned long shift : {0ul, 13ul, 20ul}) {
      size_t* h = ParallelAlgorithms::HistogramOneByteComponentParallel_2<4096, 12>(in.data(), l, r, shift);
      std::vector<size_t> bf(4096, 0);
      for (size_t i = l; i <= r; ++i) bf[(u[i] >> shift) & 4095]++;
      EXPECT(std::equal(bf.begin(), bf.end(), h), "HistogramOneByteComponentParallel_2<4096,12>(shift %lu)", shift);
      delete[] h;
    }
  }
  {
    // A range that starts and ends inside a quantum: counts land in quanta
    // l / quantum .. r / quantum by absolute index (HistogramParallel.h:350-351).
    std::vector<unsigned> in(u);
    const size_t quantum = 65536, nq = n / quantum + 1, l = 70001, r = n - 5001;
    size_t** c = ParallelAlgorithms::HistogramByteComponentsQCPar<256, 8>(in.data(), l, r, quantum, nq, 1);
    std::vector<size_t> bf(nq * 256, 0);
    for (size_t i = l; i <= r; ++i) bf[(i / quantum) * 256 + ((u[i] >> 8) & 255)]++;
    bool ok = true;
    for (size_t q = 0; q < nq; ++q)
      for (unsigned b = 0; b < 256; ++b) ok = ok && c[q][b] == bf[q * 256 + b];
    EXPECT(ok, "HistogramByteComponentsQCPar<256,8>(l=%zu, r=%zu)", l, r);
    for (size_t q = 0; q < nq; ++q) delete[] c[q];
    delete[] c;
  }

  // Merges with int bounds and the in-place forms, GPU-sized and host-sized.
  for (size_t half : {size_t(1000000), size_t(5000)}) {
    const size_t n1 = half + 3, n2 = half;
    std::vector<unsigned> t(u.begin(), u.begin() + n1 + n2);
    std::sort(t.begin(), t.begin() + n1);
    std::sort(t.begin() + n1, t.end());
    std::vector<unsigned> expect(n1 + n2);
    std::merge(t.begin(), t.begin() + n1, t.begin() + n1, t.end(), expect.begin());
    std::vector<unsigned> out(n1 + n2 + 9, 0);
    ParallelAlgorithms::merge_dac(t.data(), 0, (int)n1 - 1, (int)n1, (int)(n1 + n2) - 1, out.data(), 9);
    EXPECT(std::equal(expect.begin(), expect.end(), out.begin() + 9), "merge_dac(int bounds, %zu)", n1 + n2);
    std::vector<unsigned> out3(n1 + n2, 0);
    ParallelAlgorithms::merge_parallel_L3(t.data(), 0, (int)n1 - 1, (int)n1, (int)(n1 + n2) - 1, out3.data(), 0);
    EXPECT(out3 == expect, "merge_parallel_L3(int bounds, %zu)", n1 + n2);
    std::vector<unsigned> ip(t);
    ParallelAlgorithms::merge_in_place(ip.data(), 0, (int)n1 - 1, (int)(n1 + n2) - 1);
    EXPECT(ip == expect, "merge_in_place(int bounds, %zu)", n1 + n2);
    // mergeInPlace(a, a_size): a[0..a_size) and a[a_size..2*a_size).
    std::vector<unsigned> eq(u.begin(), u.begin() + 2 * half);
    std::sort(eq.begin(), eq.begin() + half);
    std::sort(eq.begin() + half, eq.end());
    std::vector<unsigned> eqx(2 * half);
    std::merge(eq.begin(), eq.begin() + half, eq.begin() + half, eq.end(), eqx.begin());
    ParallelAlgorithms::mergeInPlace(eq.data(), (unsigned long)half);
    EXPECT(eq == eqx, "mergeInPlace(a, a_size=%zu)", half);
    // mergeInPlace(a, l, m, r) on an interior range; outside stays untouched.
    std::vector<unsigned> mid(t.size() + 20, 7u);
    std::copy(t.begin(), t.end(), mid.begin() + 10);
    ParallelAlgorithms::mergeInPlace(mid.data(), 10, (int)(10 + n1 - 1), (int)(10 + n1 + n2 - 1));
    EXPECT(std::equal(expect.begin(), expect.end(), mid.begin() + 10) && mid[9] == 7u && mid[10 + n1 + n2] == 7u,
           "mergeInPlace(a, l, m, r) (%zu)", n1 + n2);
    // 64-bit keys in place (pa_merge_inplace_u64 for the GPU-sized case).
    std::mt19937_64 g(99);
    std::vector<uint64_t> w(n1 + n2);
    for (auto& x : w) x = g();
    std::sort(w.begin(), w.begin() + n1);
    std::sort(w.begin() + n1, w.end());
    std::vector<uint64_t> wx(n1 + n2);
    std::merge(w.begin(), w.begin() + n1, w.begin() + n1, w.end(), wx.begin());
    ParallelAlgorithms::merge_truly_in_place(w.data(), 0, n1 - 1, n1 + n2 - 1);
    EXPECT(w == wx, "merge_truly_in_place(uint64_t, %zu)", n1 + n2);
  }
}

// The radix-leaf merge sorts (ParallelMergeSort.h:255-307) with both result
// directions and leaf sizes from every segmented class.
static void merge_sorts() {
  const size_t n = 2000003;
  const std::vector<unsigned> u = rand_u32(n, 5151);
  std::vector<unsigned> sorted(u);
  std::sort(sorted.begin(), sorted.end());
  for (size_t thr : {size_t(4096), size_t(70000), size_t(24 * 1024), size_t(1) << 20}) {
    std::vector<unsigned> a(u), d(n, 0);
    ParallelAlgorithms::parallel_merge_sort_hybrid_radix_inner(a.data(), 0, n - 1, d.data(), false, thr);
    EXPECT(a == sorted, "parallel_merge_sort_hybrid_radix_inner(srcToDst=false, thr=%zu)", thr);
    std::vector<unsigned> b(u), e(n, 0);
    ParallelAlgorithms::parallel_merge_sort_hybrid_radix_inner(b.data(), 0, n - 1, e.data(), true, thr);
    EXPECT(e == sorted && b == u, "parallel_merge_sort_hybrid_radix_inner(srcToDst=true, thr=%zu)", thr);
  }
  std::vector<unsigned> c(u), w(n);
  ParallelAlgorithms::parallel_merge_sort_hybrid_radix_single_buffer(c.data(), 0, n - 1, w.data(), false);
  EXPECT(c == sorted, "parallel_merge_sort_hybrid_radix_single_buffer");
  std::vector<unsigned> part(u), pw(n);
  ParallelAlgorithms::parallel_merge_sort_hybrid_radix(part.data(), 100, n - 101, pw.data(), false, 5000);
  std::vector<unsigned> pe(u);
  std::sort(pe.begin() + 100, pe.end() - 100);
  EXPECT(part == pe, "parallel_merge_sort_hybrid_radix(l=100, r=n-101)");
}

static uint32_t order_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  return (x & 0x80000000u) ? ~x : (x | 0x80000000u);
}

static void host_cutoffs() {
  // Below 100 keys: the insertion sort on the host, same order as the GPU.
  for (size_t len : {size_t(1), size_t(2), size_t(50), size_t(99)}) {
    std::vector<unsigned> a = rand_u32(len, 17 + len);
    std::vector<unsigned> e(a);
    std::sort(e.begin(), e.end());
    ParallelAlgorithms::SortRadixPar(a.data(), a.size());
    EXPECT(a == e, "SortRadixPar host cut-off (%zu)", len);
    std::vector<uint64_t> b(len);
    std::mt19937_64 g(len);
    for (auto& x : b) x = g();
    std::vector<uint64_t> be(b), bt(len);
    std::sort(be.begin(), be.end());
    ParallelAlgorithms::SortRadixPar(b.data(), bt.data(), b.size());
    EXPECT(b == be, "SortRadixPar(u64) host cut-off (%zu)", len);
  }
  {
    const float nan = std::numeric_limits<float>::quiet_NaN(), inf = std::numeric_limits<float>::infinity();
    std::vector<float> f = {3.5f, -0.0f, 0.0f, nan, -inf, 1e-40f, -nan, inf, -2.0f, 0.0f, -0.0f, 7.0f};
    for (size_t i = 0; i < 60; ++i) f.push_back((float)i * 0.37f - 11.0f);
    auto bits = [](const std::vector<float>& v) {
      std::vector<uint32_t> o;
      for (float x : v) o.push_back(order_bits(x));
      return o;
    };
    std::vector<uint32_t> want = bits(f);
    std::sort(want.begin(), want.end());
    std::vector<float> small(f), tmp(f.size());
    ParallelAlgorithms::SortRadixPar(small.data(), tmp.data(), small.size());  // 72 keys: host
    EXPECT(bits(small) == want, "SortRadixPar(float) host cut-off in the GPU's total order");
    std::vector<float> big(f);
    while (big.size() < 5000) big.insert(big.end(), f.begin(), f.end());
    std::vector<uint32_t> wb = bits(big);
    std::sort(wb.begin(), wb.end());
    std::vector<float> tb(big.size());
    ParallelAlgorithms::SortRadixPar(big.data(), tb.data(), big.size());  // GPU
    EXPECT(bits(big) == wb, "SortRadixPar(float) on the GPU in the same order");
    std::vector<double> d = {2.0, -0.0, 0.0, -std::numeric_limits<double>::infinity(), 1.5, -3.25};
    std::vector<double> de = {-std::numeric_limits<double>::infinity(), -3.25, -0.0, 0.0, 1.5, 2.0};
    ParallelAlgorithms::sort_par(d);
    bool same = d.size() == de.size();
    for (size_t i = 0; same && i < d.size(); ++i) same = std::memcmp(&d[i], &de[i], 8) == 0;
    EXPECT(same, "sort_par(vector<double>) host cut-off, -0 before +0");
  }
}

static void concurrent_callers() {
  std::vector<std::thread> th;
  for (int t = 0; t < 8; ++t) {
    th.emplace_back([t] {
      for (int rep = 0; rep < 6; ++rep) {
        const size_t n1 = (rep % 2 ? 20000 : 1500000) + 31 * t, n2 = (rep % 2 ? 9000 : 1200000) + 17 * rep;
        std::vector<unsigned> src = rand_u32(n1 + n2, 1000 * t + rep);
        std::sort(src.begin(), src.begin() + n1);
        std::sort(src.begin() + n1, src.end());
        std::vector<unsigned> expect(n1 + n2), out(n1 + n2 + 3, 0);
        std::merge(src.begin(), src.begin() + n1, src.begin() + n1, src.end(), expect.begin());
        ParallelAlgorithms::merge_parallel_L5(src.data(), 0, n1 - 1, n1, n1 + n2 - 1, out.data(), 3);
        EXPECT(std::equal(expect.begin(), expect.end(), out.begin() + 3), "thread %d rep %d: merge_parallel_L5", t,
               rep);
        std::vector<unsigned> s = rand_u32(700000 + 1001 * t, 77 * t + rep), se(s);
        std::sort(se.begin(), se.end());
        ParallelAlgorithms::SortRadixPar(s.data(), s.size());
        EXPECT(s == se, "thread %d rep %d: SortRadixPar", t, rep);
      }
    });
  }
  for (auto& x : th) x.join();
}

int main() {
  forwards();
  merge_sorts();
  host_cutoffs();
  concurrent_callers();
  std::printf(failures ? "FAILED (%d)\n" : "ALL OK\n", failures.load());
  return failures ? 1 : 0;
}
