// pa_oracle.cpp -- CPU restatement of the reference's parallel radix-sort path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the timed CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load liboracle.so; the product library (libpa_hip.so) never links,
// calls or falls back to it.
//
// What it restates (reference file:line, DragonSpit/ParallelAlgorithms):
//   SortRadixPar (3-arg)            RadixSortLsdParallel.h:367-387
//   SortRadixPar (2-arg)            RadixSortLsdParallel.h:340-364
//   SortRadixInnerPar               RadixSortLsdParallel.h:215-336
//   ComputeStartOfBinsPar           RadixSortLsdParallel.h:108-165
//   PermuteDerandomizedNew          RadixSortLsdParallel.h:169-212
//   ..._unsigned_TwoPhase           RadixSortLsdParallel.h:43-106
//   HistogramByteComponentsParallel HistogramParallel.h:249-329
//   HistogramByteComponentsQCPar    HistogramParallel.h:331-464
//   HistogramOneByteComponentParallel (unsigned)  HistogramParallel.h:535-576
//   HistogramOneByteComponentParallel_2 (unsigned) HistogramParallel.h:466-533
//   byte-array histograms           HistogramParallel.h:16-247
//   SelectRadix                     RadixSelect.h:36-168, Histogram.h:103-133
//   merge_parallel_L5 / merge_ptr_1 ParallelMerge.h:52-75, 285-315
//   my_binary_search                BinarySearch.h:33-47
//   insertionSortSimilarToSTLnoSelfAssignment  InsertionSort.h:6-24
//
// TBB's parallel_invoke / task_group are replaced by a small std::thread pool.
// Recursive bisections keep the reference's split points (the leaves are
// enumerated with the same midpoint formulas, then run on the pool); every
// output is schedule independent (disjoint writes, exact integer sums), so
// the results equal the reference's.
//
// New components that have no parallel reference (u64 keys, key-value, float
// keys) are generalisations of the same quantum/derandomised structure and are
// labelled as such below.  How each function is pinned (closed forms, brute
// force, std:: routines) is in tests/test_oracle.py; no function here is
// "parity unpinned".

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Thread pool: a persistent set of workers running parallel_for(count, fn).
// Stands in for TBB's scheduler (Configuration.h:7-22).
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(unsigned nthreads) { resize(nthreads); }
  ~Pool() { stop(); }

  unsigned size() const { return nthreads_; }

  void resize(unsigned nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads == nthreads_) return;
    stop();
    nthreads_ = nthreads;
    quit_ = false;
    for (unsigned i = 1; i < nthreads_; ++i) workers_.emplace_back([this] { loop(); });
  }

  // Runs fn(0..count-1) on all pool threads plus the caller; blocks until done.
  void parallel_for(size_t count, const std::function<void(size_t)>& fn) {
    if (count == 0) return;
    if (nthreads_ == 1 || count == 1) {
      for (size_t i = 0; i < count; ++i) fn(i);
      return;
    }
    std::unique_lock<std::mutex> job_lock(job_mu_);  // one job at a time
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      count_ = count;
      next_.store(0);
      active_ = nthreads_ - 1;
      ++epoch_;
    }
    cv_.notify_all();
    run_items();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return active_ == 0; });
    fn_ = nullptr;
  }

 private:
  void run_items() {
    for (;;) {
      size_t i = next_.fetch_add(1);
      if (i >= count_) break;
      (*fn_)(i);
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return quit_ || epoch_ != seen; });
        if (quit_) return;
        seen = epoch_;
      }
      run_items();
      std::lock_guard<std::mutex> lk(mu_);
      if (--active_ == 0) done_cv_.notify_one();
    }
  }
  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      quit_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
    workers_.clear();
  }

  unsigned nthreads_ = 0;
  std::vector<std::thread> workers_;
  std::mutex mu_, job_mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t)>* fn_ = nullptr;
  size_t count_ = 0;
  std::atomic<size_t> next_{0};
  unsigned active_ = 0;
  uint64_t epoch_ = 0;
  bool quit_ = false;
};

unsigned g_hw = 0;  // stands in for std::thread::hardware_concurrency()

unsigned hw_threads() {
  if (g_hw == 0) {
    g_hw = std::thread::hardware_concurrency();
    if (g_hw == 0) g_hw = 1;
  }
  return g_hw;
}

Pool& pool() {
  static Pool p(hw_threads());
  if (p.size() != hw_threads()) p.resize(hw_threads());
  return p;
}

// ---------------------------------------------------------------------------
// Helpers: InsertionSort.h:6-24, BinarySearch.h:33-47, RadixSortCommon.h:57-61
// ---------------------------------------------------------------------------
template <class T>
void insertion_sort(T* a, size_t n) {
  for (size_t i = 1; i < n; i++) {
    if (a[i] < a[i - 1]) {
      T cur = a[i];
      a[i] = a[i - 1];
      size_t j = i - 1;
      for (; j > 0 && cur < a[j - 1]; j--) a[j] = a[j - 1];
      a[j] = cur;
    }
  }
}

// Lower bound over the inclusive range [left, right]; returns right+1 when all
// elements are < value (BinarySearch.h:33-47).
template <class T>
size_t lower_bound_incl(T value, const T* a, size_t left, size_t right) {
  size_t lo = left;
  size_t hi = std::max(left, right + 1);
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (value <= a[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  return hi;
}

template <class T>
inline unsigned digit_of(T v, unsigned shift, unsigned mask) {
  return (unsigned)((v >> shift) & mask);
}

// Midpoint used by the reference's unsigned bisections (HistogramParallel.h:309,
// 516, 559): average without overflow.
inline size_t mid_no_overflow(size_t l, size_t r) { return r / 2 + l / 2 + (r % 2 + l % 2) / 2; }

// Collects the leaves [l, r] (inclusive) of the reference's recursive bisection:
// split while r-l+1 > thr at m, left = [l, m], right = [m+1, r].
template <class MidFn>
void bisect_leaves_incl(size_t l, size_t r, size_t thr, MidFn mid, std::vector<std::pair<size_t, size_t>>& out) {
  if (l > r) return;
  if (r - l + 1 <= thr) {
    out.emplace_back(l, r);
    return;
  }
  size_t m = mid(l, r);
  bisect_leaves_incl(l, m, thr, mid, out);
  bisect_leaves_incl(m + 1, r, thr, mid, out);
}

// ---------------------------------------------------------------------------
// HistogramByteComponentsParallel<256,8> (HistogramParallel.h:249-329).
// out: 8 rows x 256 (the reference allocates Log2ofPowerOfTwoRadix = 8 rows;
// rows 4..7 stay zero).  Leaves: [l, m], [m+1, r] with m = mid_no_overflow.
// ---------------------------------------------------------------------------
void hist_byte_components_u32(const uint32_t* a, size_t l, size_t r, size_t thr, uint64_t* out) {
  std::memset(out, 0, sizeof(uint64_t) * 8 * 256);
  if (l > r) return;
  if (thr == 0) thr = 1;
  std::vector<std::pair<size_t, size_t>> leaves;
  bisect_leaves_incl(l, r, thr, mid_no_overflow, leaves);
  std::vector<uint64_t> part(leaves.size() * 4 * 256, 0);
  pool().parallel_for(leaves.size(), [&](size_t li) {
    uint64_t* c = &part[li * 4 * 256];
    for (size_t i = leaves[li].first; i <= leaves[li].second; i++) {
      uint32_t v = a[i];
      c[0 * 256 + (v & 0xff)]++;
      c[1 * 256 + ((v >> 8) & 0xff)]++;
      c[2 * 256 + ((v >> 16) & 0xff)]++;
      c[3 * 256 + ((v >> 24) & 0xff)]++;
    }
  });
  for (size_t li = 0; li < leaves.size(); ++li)
    for (size_t j = 0; j < 4 * 256; ++j) out[j] += part[li * 4 * 256 + j];
}

// HistogramByteComponentsQCPar -> QCParInner (HistogramParallel.h:399-464);
// each leaf is HistogramByteComponentsAcrossWorkQuantasQC (:331-397), i.e.
// count[q][byte] over [l, r] with q = index / quantum.
// out: numberOfQuantas x 256.  Inner bisection uses m = (l + r) / 2 (:420).
void hist_qc_par_u32(const uint32_t* a, size_t l, size_t r, size_t quantum, size_t nq, unsigned which_byte, size_t thr,
                     uint64_t* out) {
  std::memset(out, 0, sizeof(uint64_t) * nq * 256);
  if (l > r) return;
  size_t length = r - l + 1;
  size_t hw = hw_threads();
  if (thr * hw < length) thr = length / hw;  // :457-458
  if (thr == 0) thr = 1;
  std::vector<std::pair<size_t, size_t>> leaves;
  bisect_leaves_incl(l, r, thr, [](size_t lo, size_t hi) { return (hi + lo) / 2; }, leaves);
  // Leaves touch disjoint index ranges but may share a quantum row, so each
  // leaf counts privately and rows are summed afterwards (the reference sums
  // left and right after each parallel_invoke, :431-435).
  std::vector<std::vector<uint64_t>> part(leaves.size());
  pool().parallel_for(leaves.size(), [&](size_t li) {
    size_t ql = leaves[li].first / quantum, qr = leaves[li].second / quantum;
    part[li].assign((qr - ql + 1) * 256, 0);
    const unsigned shift = 8 * which_byte;
    for (size_t i = leaves[li].first; i <= leaves[li].second; ++i)
      part[li][(i / quantum - ql) * 256 + ((a[i] >> shift) & 0xff)]++;
  });
  for (size_t li = 0; li < leaves.size(); ++li) {
    size_t ql = leaves[li].first / quantum;
    for (size_t j = 0; j < part[li].size(); ++j) out[ql * 256 + j] += part[li][j];
  }
}

// ComputeStartOfBinsPar<256,8> (RadixSortLsdParallel.h:108-165):
// startOfBin[q][b] = sum_{b'<b} size[b'] + sum_{q'<q} count[q'][b].
void compute_start_of_bins_u32(const uint32_t* a, size_t n, size_t quantum, size_t nq, unsigned digit, size_t thr,
                               uint64_t* start_of_bin) {
  std::vector<uint64_t> count(nq * 256);
  hist_qc_par_u32(a, 0, n - 1, quantum, nq, digit, thr, count.data());
  uint64_t size_of_bin[256];
  for (unsigned b = 0; b < 256; b++) {
    size_of_bin[b] = 0;
    for (size_t q = 0; q < nq; q++) size_of_bin[b] += count[q * 256 + b];
  }
  start_of_bin[0] = 0;
  for (unsigned b = 1; b < 256; b++) start_of_bin[b] = start_of_bin[b - 1] + size_of_bin[b - 1];
  for (size_t q = 1; q < nq; q++)
    for (unsigned b = 0; b < 256; b++) start_of_bin[q * 256 + b] = start_of_bin[(q - 1) * 256 + b] + count[(q - 1) * 256 + b];
}

// ---------------------------------------------------------------------------
// Generic per-quantum stable LSD pass with 64-deep per-bin write buffers:
// PermuteDerandomizedNew (RadixSortLsdParallel.h:169-212).  K = key type,
// V = payload type (void payloads are not moved).
// ---------------------------------------------------------------------------
constexpr size_t kBufferDepth = 64;  // RadixSortLsdParallel.h:224

template <class K>
void permute_derandomized(const K* in, K* out, uint64_t* start_of_bin_q, size_t start, size_t end, unsigned shift,
                          K* buf /*256*64*/, size_t* buf_idx /*256*/) {
  for (size_t b = 0; b < 256; ++b) buf_idx[b] = b * kBufferDepth;
  for (size_t i = start; i < end; ++i) {
    K v = in[i];
    unsigned d = digit_of<K>(v, shift, 0xff);
    if (buf_idx[d] < (d + 1) * kBufferDepth) {
      buf[buf_idx[d]++] = v;
    } else {
      std::memcpy(out + start_of_bin_q[d], buf + d * kBufferDepth, kBufferDepth * sizeof(K));
      start_of_bin_q[d] += kBufferDepth;
      buf[d * kBufferDepth] = v;
      buf_idx[d] = d * kBufferDepth + 1;
    }
  }
  for (size_t b = 0; b < 256; ++b) {
    size_t items = buf_idx[b] - b * kBufferDepth;
    std::memcpy(out + start_of_bin_q[b], buf + b * kBufferDepth, items * sizeof(K));
    start_of_bin_q[b] += items;
  }
}

// Key-value variant (new: the reference has no key-value LSD).  Same buffered
// structure; the payload travels in a parallel buffer.
template <class K, class V>
void permute_derandomized_kv(const K* kin, const V* vin, K* kout, V* vout, uint64_t* sob, size_t start, size_t end,
                             unsigned shift, K* kbuf, V* vbuf, size_t* buf_idx) {
  for (size_t b = 0; b < 256; ++b) buf_idx[b] = b * kBufferDepth;
  for (size_t i = start; i < end; ++i) {
    K k = kin[i];
    V v = vin[i];
    unsigned d = digit_of<K>(k, shift, 0xff);
    if (buf_idx[d] < (d + 1) * kBufferDepth) {
      kbuf[buf_idx[d]] = k;
      vbuf[buf_idx[d]++] = v;
    } else {
      std::memcpy(kout + sob[d], kbuf + d * kBufferDepth, kBufferDepth * sizeof(K));
      std::memcpy(vout + sob[d], vbuf + d * kBufferDepth, kBufferDepth * sizeof(V));
      sob[d] += kBufferDepth;
      kbuf[d * kBufferDepth] = k;
      vbuf[d * kBufferDepth] = v;
      buf_idx[d] = d * kBufferDepth + 1;
    }
  }
  for (size_t b = 0; b < 256; ++b) {
    size_t items = buf_idx[b] - b * kBufferDepth;
    std::memcpy(kout + sob[b], kbuf + b * kBufferDepth, items * sizeof(K));
    std::memcpy(vout + sob[b], vbuf + b * kBufferDepth, items * sizeof(V));
    sob[b] += items;
  }
}

// Per-quantum counts of one digit for key type K (generalises
// hist_qc_par_u32 to 64-bit keys; identical leaf structure).
template <class K>
void quanta_counts(const K* a, size_t n, size_t quantum, size_t nq, unsigned shift, std::vector<uint64_t>& count) {
  count.assign(nq * 256, 0);
  pool().parallel_for(nq, [&](size_t q) {
    size_t s = q * quantum, e = std::min(n, s + quantum);
    uint64_t* c = &count[q * 256];
    for (size_t i = s; i < e; ++i) c[digit_of<K>(a[i], shift, 0xff)]++;
  });
}

// SortRadixInnerPar<256,8> (RadixSortLsdParallel.h:215-336), generalised over
// key width.  For u32 keys the start-of-bin table comes from the restated
// ComputeStartOfBinsPar (with the QCPar bisection); for wider keys (new) the
// same table is formed from per-quantum counts.
template <class K>
void sort_inner_par(K* in, K* work, size_t n, size_t quantum) {
  if (n == 0) return;
  const size_t nq = (n % quantum) == 0 ? n / quantum : n / quantum + 1;
  const unsigned digits = sizeof(K);
  std::vector<uint64_t> sob(nq * 256);
  for (unsigned d = 0; d < digits; ++d) {
    const unsigned shift = 8 * d;
    if (sizeof(K) == 4) {
      compute_start_of_bins_u32(reinterpret_cast<const uint32_t*>(in), n, quantum, nq, d, 16 * 1024, sob.data());
    } else {
      std::vector<uint64_t> count;
      quanta_counts<K>(in, n, quantum, nq, shift, count);
      uint64_t run = 0;
      for (unsigned b = 0; b < 256; ++b) {
        uint64_t acc = run;
        for (size_t q = 0; q < nq; ++q) {
          sob[q * 256 + b] = acc;
          acc += count[q * 256 + b];
        }
        run = acc;
      }
    }
    pool().parallel_for(nq, [&](size_t q) {
      thread_local std::vector<K> buf;
      thread_local std::vector<size_t> idx;
      buf.resize(256 * kBufferDepth);
      idx.resize(256);
      size_t s = q * quantum, e = std::min(n, s + quantum);
      permute_derandomized<K>(in, work, &sob[q * 256], s, e, shift, buf.data(), idx.data());
    });
    std::swap(in, work);
  }
}

// SortRadixInnerPar<PowerOfTwoRadix, Log2ofPowerOfTwoRadix> at any radix of
// 2^4 .. 2^8 (RadixSortLsdParallel.h:215-336): the digit loop runs while the
// shifted mask has bits left (:250-253, :309-311), i.e. ceil(key bits / bits)
// passes, each a per-quantum count, a (bin x quantum) start-of-bin prefix and
// the buffered stable permutation (PermuteDerandomizedNew, :169-212).  Two
// departures, both where the reference only works for 8-bit digits: its
// ComputeStartOfBinsPar always counts bytes (HistogramByteComponentsQCPar's
// whichByte, HistogramParallel.h:337-338), so the counts here are of the
// pass's own digit; and its copy of an odd pass count's result back into the
// input is commented out (:327-329), so the result here always ends in `in`.
// max_passes < passes stops early (no copy back): the state after that many
// passes, for tests of the digit order.
template <class K>
void sort_inner_par_bits(K* in, K* work, size_t n, size_t quantum, unsigned bits, unsigned max_passes) {
  if (n == 0) return;
  const unsigned R = 1u << bits;
  const K mask = (K)(R - 1);
  const unsigned passes = (8 * sizeof(K) + bits - 1) / bits;
  const size_t nq = (n % quantum) == 0 ? n / quantum : n / quantum + 1;
  std::vector<uint64_t> count(nq * R), sob(nq * R);
  K* const orig = in;
  unsigned done = 0;
  for (unsigned p = 0; p < passes && p < max_passes; ++p, ++done) {
    const unsigned shift = bits * p;
    pool().parallel_for(nq, [&](size_t q) {
      uint64_t* c = &count[q * R];
      std::fill(c, c + R, 0);
      const size_t s0 = q * quantum, e0 = std::min(n, s0 + quantum);
      for (size_t i = s0; i < e0; ++i) c[(in[i] >> shift) & mask]++;
    });
    uint64_t run = 0;
    for (unsigned b = 0; b < R; ++b)
      for (size_t q = 0; q < nq; ++q) {
        sob[q * R + b] = run;
        run += count[q * R + b];
      }
    pool().parallel_for(nq, [&](size_t q) {
      thread_local std::vector<K> buf;
      thread_local std::vector<size_t> idx;
      buf.resize((size_t)R * kBufferDepth);
      idx.resize(R);
      uint64_t* start = &sob[q * R];
      for (unsigned b = 0; b < R; ++b) idx[b] = b * kBufferDepth;
      const size_t s0 = q * quantum, e0 = std::min(n, s0 + quantum);
      for (size_t i = s0; i < e0; ++i) {
        const K v = in[i];
        const unsigned d = (unsigned)((v >> shift) & mask);
        if (idx[d] == (d + 1) * kBufferDepth) {
          std::memcpy(work + start[d], buf.data() + d * kBufferDepth, kBufferDepth * sizeof(K));
          start[d] += kBufferDepth;
          idx[d] = d * kBufferDepth;
        }
        buf[idx[d]++] = v;
      }
      for (unsigned b = 0; b < R; ++b) {
        const size_t items = idx[b] - b * kBufferDepth;
        std::memcpy(work + start[b], buf.data() + b * kBufferDepth, items * sizeof(K));
      }
    });
    std::swap(in, work);
  }
  if (in != orig) std::memcpy(orig, in, n * sizeof(K));
}

// Key-value generalisation (new).  Stable: ties keep input order.
template <class K, class V>
void sort_inner_par_kv(K* kin, V* vin, K* kwork, V* vwork, size_t n, size_t quantum) {
  if (n == 0) return;
  const size_t nq = (n % quantum) == 0 ? n / quantum : n / quantum + 1;
  std::vector<uint64_t> sob(nq * 256);
  for (unsigned d = 0; d < sizeof(K); ++d) {
    const unsigned shift = 8 * d;
    std::vector<uint64_t> count;
    quanta_counts<K>(kin, n, quantum, nq, shift, count);
    uint64_t run = 0;
    for (unsigned b = 0; b < 256; ++b) {
      uint64_t acc = run;
      for (size_t q = 0; q < nq; ++q) {
        sob[q * 256 + b] = acc;
        acc += count[q * 256 + b];
      }
      run = acc;
    }
    pool().parallel_for(nq, [&](size_t q) {
      thread_local std::vector<K> kb;
      thread_local std::vector<V> vb;
      thread_local std::vector<size_t> idx;
      kb.resize(256 * kBufferDepth);
      vb.resize(256 * kBufferDepth);
      idx.resize(256);
      size_t s = q * quantum, e = std::min(n, s + quantum);
      permute_derandomized_kv<K, V>(kin, vin, kwork, vwork, &sob[q * 256], s, e, shift, kb.data(), vb.data(),
                                    idx.data());
    });
    std::swap(kin, kwork);
    std::swap(vin, vwork);
  }
}

// Quantum sizing of the 3-arg SortRadixPar (RadixSortLsdParallel.h:374-379).
size_t quantum_3arg(size_t n, size_t thr) {
  size_t hw = hw_threads();
  if (hw > 0 && thr * hw < n) thr = n / hw;
  return thr == 0 ? 1 : thr;
}

// ---------------------------------------------------------------------------
// RadixSelect (RadixSelect.h:36-168) with HistogramOneComponentOpt
// (Histogram.h:103-133).  The reference's second MoveOutsideOfKthBinIn call
// scans one element past `last` (RadixSelect.h:112 -> :45); that element can
// never be moved (the k-th bin is full by then), so the restatement stops at
// `last` and returns the same value without the out-of-range read.
// ---------------------------------------------------------------------------
size_t move_outside_kth_bin_in(uint32_t* a, size_t start_ob, size_t len_ob, size_t start_kb, size_t len_kb, int shift,
                               unsigned mask, size_t kbin) {
  size_t end_kb = start_kb + len_kb;
  size_t end_ob = start_ob + len_ob;
  size_t ob = start_ob, ib = start_kb;
  for (;;) {
    for (; ob < end_ob; ob++)
      if (((a[ob] >> shift) & mask) == kbin) break;
    if (ob < end_ob)
      for (; ib < end_kb; ib++)
        if (((a[ib] >> shift) & mask) != kbin) break;
    if (ob >= end_ob || ib >= end_kb) break;
    a[ib++] = a[ob++];
  }
  return ib;
}

void hist_one_component(const uint32_t* a, size_t l, size_t r, unsigned shift, unsigned bits, uint64_t* count) {
  const unsigned nb = 1u << bits, mask = nb - 1;
  // Histogram.h:113 leaves `count` untouched when l >= r; every caller in the
  // select path passes l < r, and zeroing here keeps the restatement defined.
  for (unsigned i = 0; i < nb; ++i) count[i] = 0;
  for (size_t i = l; i < r; ++i) count[(a[i] >> shift) & mask]++;
}

uint32_t select_radix_inner(uint32_t* a, size_t first, size_t length, int shift, unsigned bits, size_t k,
                            size_t* pass_counts /*optional: candidates per pass, 4 entries*/) {
  const unsigned mask = (1u << bits) - 1;
  const size_t nb = (size_t)1 << bits;
  std::vector<uint64_t> sob(nb + 1), count(nb);
  int pass = 0;
  while (shift >= 0) {
    size_t last = first + length;
    if (pass_counts && pass < 4) pass_counts[pass] = length;
    hist_one_component(a, first, last, (unsigned)shift, bits, count.data());
    sob[0] = first;
    sob[nb] = last;
    for (size_t i = 1; i < nb; i++) sob[i] = sob[i - 1] + count[i - 1];
    size_t kbin = 0;
    for (; kbin < nb; kbin++)
      if (k >= sob[kbin] && k < sob[kbin + 1]) break;
    size_t ib = move_outside_kth_bin_in(a, first, sob[kbin] - first, sob[kbin], sob[kbin + 1] - sob[kbin], shift, mask,
                                        kbin);
    move_outside_kth_bin_in(a, sob[kbin + 1], last - sob[kbin + 1], ib, sob[kbin + 1] - ib, shift, mask, kbin);
    ++pass;
    if (shift == 0) break;
    shift = shift >= 8 ? shift - 8 : 0;
    if (sob[kbin + 1] - sob[kbin] > 1) {
      first = sob[kbin];
      length = sob[kbin + 1] - sob[kbin];
    } else if (sob[kbin + 1] - sob[kbin] == 1) {
      break;
    }
  }
  return a[k];
}

// ---------------------------------------------------------------------------
// merge_parallel_L5 (ParallelMerge.h:285-315) with merge_ptr_1 (:52-75).
// The recursion is unrolled on the calling thread (pivots written, leaves
// collected with the same split formula), then the leaves run on the pool.
// ---------------------------------------------------------------------------
template <class T>
void merge_ptr_1(const T* a, const T* a_end, const T* b, const T* b_end, T* dst) {
  if (a < a_end && b < b_end) {
    for (;;) {
      if (*a <= *b) {
        *dst++ = *a++;
        if (a >= a_end) break;
      } else {
        *dst++ = *b++;
        if (b >= b_end) break;
      }
    }
  }
  while (a < a_end) *dst++ = *a++;
  while (b < b_end) *dst++ = *b++;
}

struct MergeLeaf {
  size_t p1, n1, p2, n2, p3;
};

template <class T>
void merge_l5_split(const T* t, size_t p1, size_t r1, size_t p2, size_t r2, T* a, size_t p3, size_t thr,
                    std::vector<MergeLeaf>& leaves) {
  size_t len1 = r1 - p1 + 1;
  size_t len2 = r2 - p2 + 1;
  if (len1 < len2) {
    std::swap(p1, p2);
    std::swap(r1, r2);
    std::swap(len1, len2);
  }
  if (len1 == 0) return;
  if (len1 + len2 <= thr) {
    leaves.push_back({p1, len1, p2, len2, p3});
    return;
  }
  size_t q1 = mid_no_overflow(p1, r1);
  size_t q2 = lower_bound_incl(t[q1], t, p2, r2);
  size_t q3 = p3 + (q1 - p1) + (q2 - p2);
  a[q3] = t[q1];
  // Recursive calls use the default threshold (ParallelMerge.h:311-312).
  merge_l5_split(t, p1, q1 - 1, p2, q2 - 1, a, p3, 32768, leaves);
  merge_l5_split(t, q1 + 1, r1, q2, r2, a, q3 + 1, 32768, leaves);
}

template <class T>
void merge_l5(const T* t, size_t p1, size_t r1, size_t p2, size_t r2, T* a, size_t p3, size_t thr) {
  std::vector<MergeLeaf> leaves;
  merge_l5_split(t, p1, r1, p2, r2, a, p3, thr, leaves);
  pool().parallel_for(leaves.size(), [&](size_t i) {
    const MergeLeaf& L = leaves[i];
    merge_ptr_1(t + L.p1, t + L.p1 + L.n1, t + L.p2, t + L.p2 + L.n2, a + L.p3);
  });
}

// ---------------------------------------------------------------------------
// PartitionRadixMsdUIntInner (RadixPartition.h:30-86): MSD in-place radix
// partition.  One 256-bin count of the current digit over [start, start+size),
// an in-place cycle permutation into the bins (each element is swapped into
// the next free slot of its bin until the slot it came from receives an
// element of the bin being filled), then recursion into only the bins that
// hold one of the requested ranks k[kStart .. kStart+kLength) (k sorted
// ascending); bins below Threshold (48) are finished by insertion sort.
// ---------------------------------------------------------------------------
template <class T>
void partition_radix_inner(T* a, size_t start, size_t size, int shift, const size_t* k, size_t k0, size_t kn) {
  constexpr size_t R = 256;
  constexpr long Threshold = 48;
  const size_t last = start + size;
  size_t count[R] = {0};
  for (size_t i = start; i < last; ++i) count[(a[i] >> shift) & 0xFF]++;
  size_t bin_start[R + 1], bin_end[R];
  bin_start[0] = bin_end[0] = start;
  bin_start[R] = start;  // sentinel, as in the reference (:44)
  for (size_t i = 1; i < R; ++i) bin_start[i] = bin_end[i] = bin_start[i - 1] + count[i - 1];
  size_t next_bin = 1;
  for (size_t cur = start; cur < last;) {
    unsigned digit;
    T e = a[cur];
    while (bin_end[digit = (unsigned)((e >> shift) & 0xFF)] != cur) std::swap(e, a[bin_end[digit]++]);
    a[cur] = e;
    bin_end[digit]++;
    while (bin_end[next_bin - 1] == bin_start[next_bin]) next_bin++;
    cur = bin_end[next_bin - 1];
  }
  if (shift <= 0) return;
  shift = shift >= 8 ? shift - 8 : 0;
  size_t ki = k0;
  for (size_t bin = 0; bin < R && ki < k0 + kn; ++bin) {
    if (bin_start[bin] < bin_end[bin] && k[ki] >= bin_start[bin] && k[ki] < bin_end[bin]) {
      const size_t first_k = ki++;
      size_t nk = 1;
      for (; ki < k0 + kn && k[ki] >= bin_start[bin] && k[ki] < bin_end[bin]; ++ki) ++nk;
      const size_t m = bin_end[bin] - bin_start[bin];
      if ((long)m >= Threshold)
        partition_radix_inner(a, bin_start[bin], m, shift, k, first_k, nk);
      else
        insertion_sort(&a[bin_start[bin]], m);
    }
  }
}

inline uint32_t f32_to_ordered(uint32_t u) { return u ^ ((u >> 31) ? 0xFFFFFFFFu : 0x80000000u); }
inline uint32_t ordered_to_f32(uint32_t v) { return v ^ ((v >> 31) ? 0x80000000u : 0xFFFFFFFFu); }
inline uint64_t f64_to_ordered(uint64_t u) { return u ^ ((u >> 63) ? ~0ull : 1ull << 63); }
inline uint64_t ordered_to_f64(uint64_t v) { return v ^ ((v >> 63) ? 1ull << 63 : ~0ull); }

// Signed keys: flipping the sign bit maps two's-complement order onto unsigned
// order (the top-digit flip of extractDigitNegate, RadixSortCommon.h:62-68).
template <typename U>
void sort_signed(U* a, U* tmp, size_t n, size_t thr) {
  const U top = U(1) << (sizeof(U) * 8 - 1);
  for (size_t i = 0; i < n; ++i) a[i] ^= top;
  if (n >= 100)
    sort_inner_par<U>(a, tmp, n, quantum_3arg(n, thr));
  else
    insertion_sort(a, n);
  for (size_t i = 0; i < n; ++i) a[i] ^= top;
}

}  // namespace

extern "C" {

// Sets the thread count used in place of hardware_concurrency(); 0 = detect.
void orc_set_threads(unsigned n) {
  g_hw = n;
  hw_threads();
}
unsigned orc_get_threads(void) { return hw_threads(); }

void orc_insertion_sort_u32(uint32_t* a, size_t n) { insertion_sort(a, n); }

size_t orc_binary_search_u32(uint32_t value, const uint32_t* a, size_t left, size_t right) {
  return lower_bound_incl(value, a, left, right);
}

// SortRadixPar(unsigned* a, unsigned* tmp, size_t n, size_t thr)
// RadixSortLsdParallel.h:367-387.
void orc_sort_radix_par_u32(uint32_t* a, uint32_t* tmp, size_t n, size_t thr) {
  if (n >= 100)
    sort_inner_par<uint32_t>(a, tmp, n, quantum_3arg(n, thr));
  else
    insertion_sort(a, n);
}

// SortRadixPar(unsigned* a, size_t n, size_t thr): RadixSortLsdParallel.h:340-364.
void orc_sort_radix_par_alloc_u32(uint32_t* a, size_t n, size_t thr) {
  std::vector<uint32_t> b(n);
  size_t pc = hw_threads() * 4;  // :349-354
  if (pc > 0 && thr * pc < n) thr = n / pc;
  if (thr == 0) thr = 1;
  if (n >= 100)
    sort_inner_par<uint32_t>(a, b.data(), n, thr);
  else
    insertion_sort(a, n);
}

// RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase (RadixSortLsdParallel.h:43-106):
// one parallel 4-digit histogram, then serial permutations; the result ends in
// both a and b.
void orc_sort_two_phase_u32(uint32_t* a, uint32_t* b, size_t n) {
  if (n < 100) {
    insertion_sort(a, n);
    for (size_t j = 0; j < n; ++j) b[j] = a[j];
    return;
  }
  std::vector<uint64_t> count2d(8 * 256);
  hist_byte_components_u32(a, 0, n - 1, 64 * 1024, count2d.data());
  uint32_t* in = a;
  uint32_t* out = b;
  for (unsigned d = 0; d < 4; ++d) {
    uint64_t end_of_bin[256];
    end_of_bin[0] = 0;
    for (unsigned i = 1; i < 256; ++i) end_of_bin[i] = end_of_bin[i - 1] + count2d[d * 256 + i - 1];
    for (size_t i = 0; i < n; ++i) out[end_of_bin[(in[i] >> (8 * d)) & 0xff]++] = in[i];
    std::swap(in, out);
  }
  // 4 passes: the result is in `a`; copy it to `b` as well (:80-82).
  std::memcpy(b, a, n * sizeof(uint32_t));
}

// SortRadixInnerPar at digit widths of 4 to 8 bits (the GPU sort's tuned
// widths); result in a whatever the pass count (see sort_inner_par_bits).
void orc_sort_radix_bits_u32(uint32_t* a, uint32_t* tmp, size_t n, unsigned bits, unsigned max_passes, size_t thr) {
  if (bits < 4 || bits > 8) return;
  sort_inner_par_bits<uint32_t>(a, tmp, n, quantum_3arg(n, thr), bits, max_passes);
}
void orc_sort_radix_bits_u64(uint64_t* a, uint64_t* tmp, size_t n, unsigned bits, unsigned max_passes, size_t thr) {
  if (bits < 4 || bits > 8) return;
  sort_inner_par_bits<uint64_t>(a, tmp, n, quantum_3arg(n, thr), bits, max_passes);
}

// New: 64-bit keys through the same quantum structure (8 passes).
void orc_sort_radix_par_u64(uint64_t* a, uint64_t* tmp, size_t n, size_t thr) {
  if (n >= 100)
    sort_inner_par<uint64_t>(a, tmp, n, quantum_3arg(n, thr));
  else
    insertion_sort(a, n);
}

// New: stable key-value sorts (keys ascending, payload moved with its key).
void orc_sort_kv_u64_u32(uint64_t* k, uint32_t* v, uint64_t* ktmp, uint32_t* vtmp, size_t n, size_t thr) {
  sort_inner_par_kv<uint64_t, uint32_t>(k, v, ktmp, vtmp, n, quantum_3arg(n, thr));
}
void orc_sort_kv_u32_u32(uint32_t* k, uint32_t* v, uint32_t* ktmp, uint32_t* vtmp, size_t n, size_t thr) {
  sort_inner_par_kv<uint32_t, uint32_t>(k, v, ktmp, vtmp, n, quantum_3arg(n, thr));
}

// New: float32 keys (as raw bits) through the order-preserving transform.
uint32_t orc_f32_to_ordered(uint32_t u) { return f32_to_ordered(u); }
uint32_t orc_ordered_to_f32(uint32_t v) { return ordered_to_f32(v); }

// Exhaustive round trip g(f(u)) == u over all 2^32 patterns, plus monotonicity
// of f against the IEEE total order on the non-NaN patterns checked in order.
// Returns the number of failures.
uint64_t orc_f32_transform_exhaustive_check(void) {
  std::atomic<uint64_t> bad{0};
  pool().parallel_for(256, [&](size_t hi) {
    uint64_t local = 0;
    for (uint64_t lo = 0; lo < (1ull << 24); ++lo) {
      uint32_t u = (uint32_t)((hi << 24) | lo);
      if (ordered_to_f32(f32_to_ordered(u)) != u) ++local;
    }
    bad += local;
  });
  return bad.load();
}

void orc_sort_f32(uint32_t* bits, uint32_t* tmp, size_t n, size_t thr) {
  for (size_t i = 0; i < n; ++i) bits[i] = f32_to_ordered(bits[i]);
  if (n >= 100)
    sort_inner_par<uint32_t>(bits, tmp, n, quantum_3arg(n, thr));
  else
    insertion_sort(bits, n);
  for (size_t i = 0; i < n; ++i) bits[i] = ordered_to_f32(bits[i]);
}

// float64 bit patterns: the 64-bit form of the float transform around the
// u64 LSD restatement.
void orc_sort_f64(uint64_t* bits, uint64_t* tmp, size_t n, size_t thr) {
  for (size_t i = 0; i < n; ++i) bits[i] = f64_to_ordered(bits[i]);
  if (n >= 100)
    sort_inner_par<uint64_t>(bits, tmp, n, quantum_3arg(n, thr));
  else
    insertion_sort(bits, n);
  for (size_t i = 0; i < n; ++i) bits[i] = ordered_to_f64(bits[i]);
}

void orc_sort_i32(uint32_t* a, uint32_t* tmp, size_t n, size_t thr) { sort_signed<uint32_t>(a, tmp, n, thr); }
void orc_sort_i64(uint64_t* a, uint64_t* tmp, size_t n, size_t thr) { sort_signed<uint64_t>(a, tmp, n, thr); }

// HistogramByteComponentsParallel<256,8>(a, l, r, thr): out is 8 x 256.
void orc_hist_byte_components_u32(const uint32_t* a, size_t l, size_t r, size_t thr, uint64_t* out) {
  hist_byte_components_u32(a, l, r, thr, out);
}

// HistogramByteComponentsQCPar<256,8>: out is nq x 256.
void orc_hist_qc_par_u32(const uint32_t* a, size_t l, size_t r, size_t quantum, size_t nq, unsigned which_byte,
                         size_t thr, uint64_t* out) {
  hist_qc_par_u32(a, l, r, quantum, nq, which_byte, thr, out);
}

// ComputeStartOfBinsPar<256,8>: out is nq x 256.
void orc_compute_start_of_bins_u32(const uint32_t* a, size_t n, size_t quantum, size_t nq, unsigned digit, size_t thr,
                                   uint64_t* out) {
  compute_start_of_bins_u32(a, n, quantum, nq, digit, thr, out);
}

// HistogramOneByteComponentParallel<R,L>(unsigned[], l, r, shift, thr)
// (HistogramParallel.h:535-576): bins = (a >> shift) & (R-1) over [l, r].
void orc_hist_one_byte_u32(const uint32_t* a, size_t l, size_t r, unsigned shift, unsigned radix, size_t thr,
                           uint64_t* out) {
  const unsigned mask = radix - 1;
  std::memset(out, 0, sizeof(uint64_t) * radix);
  if (l > r) return;
  if (thr == 0) thr = 1;
  std::vector<std::pair<size_t, size_t>> leaves;
  bisect_leaves_incl(l, r, thr, mid_no_overflow, leaves);
  std::vector<uint64_t> part(leaves.size() * radix, 0);
  pool().parallel_for(leaves.size(), [&](size_t li) {
    uint64_t* c = &part[li * radix];
    for (size_t i = leaves[li].first; i <= leaves[li].second; ++i) c[(a[i] >> shift) & mask]++;
  });
  for (size_t li = 0; li < leaves.size(); ++li)
    for (unsigned j = 0; j < radix; ++j) out[j] += part[li * radix + j];
}

// HistogramOneByteComponentParallel_2<R,L>(unsigned[], ...) restated WITH its
// leaf tail bug (HistogramParallel.h:499: `current < r` with inclusive r drops
// one element per leaf).  Kept only to document the divergence: the product's
// _2 returns the corrected counts.
void orc_hist_one_byte_2_bugcompat_u32(const uint32_t* a, size_t l, size_t r, unsigned shift, size_t thr,
                                       uint64_t* out) {
  std::memset(out, 0, sizeof(uint64_t) * 256);
  if (l > r) return;
  if (thr == 0) thr = 1;
  std::vector<std::pair<size_t, size_t>> leaves;
  bisect_leaves_incl(l, r, thr, mid_no_overflow, leaves);
  for (auto& lf : leaves) {
    size_t ll = lf.first, rr = lf.second;
    size_t last_by_four = ll + ((rr - ll) / 4) * 4;
    size_t cur = ll;
    for (; cur < last_by_four; ++cur) out[(a[cur] >> shift) & 0xff]++;
    for (; cur < rr; ++cur) out[(a[cur] >> shift) & 0xff]++;
  }
}

// Byte-array histograms HistogramOneByteComponentParallel{,_2,_3,_4}<256>
// (HistogramParallel.h:16-247): counts of bytes over [l, r).
void orc_hist_bytes(const uint8_t* a, size_t l, size_t r, size_t thr, uint64_t* out) {
  std::memset(out, 0, sizeof(uint64_t) * 256);
  if (l >= r) return;
  if (thr == 0) thr = 1;
  // Exclusive-bound bisection (:27,37): split while r-l > thr at m.
  std::vector<std::pair<size_t, size_t>> leaves;
  std::function<void(size_t, size_t)> rec = [&](size_t lo, size_t hi) {
    if (lo >= hi) return;
    if (hi - lo <= thr) {
      leaves.emplace_back(lo, hi);
      return;
    }
    size_t m = mid_no_overflow(lo, hi);
    rec(lo, m);
    rec(m, hi);
  };
  rec(l, r);
  std::vector<uint64_t> part(leaves.size() * 256, 0);
  pool().parallel_for(leaves.size(), [&](size_t li) {
    uint64_t* c = &part[li * 256];
    for (size_t i = leaves[li].first; i < leaves[li].second; ++i) c[a[i]]++;
  });
  for (size_t li = 0; li < leaves.size(); ++li)
    for (unsigned j = 0; j < 256; ++j) out[j] += part[li * 256 + j];
}

// counting_sort_parallel_inner<256> (CountingSortParallel.h:29-76): the byte
// histogram of [l, r), bin starts by an exclusive scan, then every bin's value
// written over its run (parallel_fill, FillParallel.h:45-66).  The reference's
// bin starts begin at index 0 (:56-59), so for l > 0 it writes [0, r - l); this
// restatement writes [l, r), the sorted range the function's name promises.
void orc_counting_sort_u8(uint8_t* a, size_t l, size_t r, size_t thr) {
  if (r <= l) return;
  uint64_t counts[256];
  orc_hist_bytes(a, l, r, thr, counts);
  size_t start = l;
  for (int b = 0; b < 256; ++b) {
    std::memset(a + start, b, counts[b]);
    start += counts[b];
  }
}

// SelectRadix(a, start, len, k) / SelectRadix(a, len, k) (RadixSelect.h:139-168).
// Mutates `a` like the reference.  pass_counts (4 entries, may be NULL)
// receives the candidate-range length at the start of each digit pass.
uint32_t orc_select_radix_u32(uint32_t* a, size_t start, size_t len, size_t k, size_t* pass_counts) {
  if (pass_counts)
    for (int i = 0; i < 4; ++i) pass_counts[i] = 0;
  return select_radix_inner(a, start, len, 24, 8, k, pass_counts);
}

// PartitionRadix(a, start, len, k[], kStart, kLength) and the single-k forms
// (RadixPartition.h:94-184): a[start, start+len) partitioned in place so that
// each a[k[i]] holds its rank's value; k ascending, absolute indices.
void orc_partition_radix_u32(uint32_t* a, size_t start, size_t len, const size_t* k, size_t k_len) {
  partition_radix_inner(a, start, len, 24, k, 0, k_len);
}

// merge_parallel_L5<T>(t, p1, r1, p2, r2, a, p3, thr) (ParallelMerge.h:285-315).
void orc_merge_parallel_l5_u32(const uint32_t* t, size_t p1, size_t r1, size_t p2, size_t r2, uint32_t* a, size_t p3,
                               size_t thr) {
  merge_l5<uint32_t>(t, p1, r1, p2, r2, a, p3, thr);
}
void orc_merge_parallel_l5_u64(const uint64_t* t, size_t p1, size_t r1, size_t p2, size_t r2, uint64_t* a, size_t p3,
                               size_t thr) {
  merge_l5<uint64_t>(t, p1, r1, p2, r2, a, p3, thr);
}

// One sequential stable counting-sort pass on digit `shift` (textbook form):
// the per-pass pin for the GPU's onesweep passes.
void orc_counting_pass_u32(const uint32_t* in, uint32_t* out, size_t n, unsigned shift) {
  uint64_t c[257] = {0};
  for (size_t i = 0; i < n; ++i) c[((in[i] >> shift) & 0xff) + 1]++;
  for (int b = 0; b < 256; ++b) c[b + 1] += c[b];
  for (size_t i = 0; i < n; ++i) out[c[(in[i] >> shift) & 0xff]++] = in[i];
}

}  // extern "C"
