/* pa_hip.h -- C ABI of the MI355X (gfx950) radix-sort library libpa_hip.so.
 *
 * Every entry point is extern "C" with plain pointers and sizes.  It is the
 * boundary that the drop-in C++ headers in this directory (RadixSortLsdParallel.h,
 * HistogramParallel.h, RadixSelect.h, ParallelMerge.h) forward to, and what an
 * FFI (ctypes, cgo, JNI, N-API) would bind.
 *
 * Conventions
 *  - Array pointers may be device memory (hipMalloc) or host memory.  With
 *    device pointers a call only enqueues work on `stream` and returns; the
 *    caller synchronises.  When any array argument is host memory the library
 *    stages it through device memory and returns after the result is back in
 *    host memory (the reference's blocking semantics).
 *  - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *  - Calls are thread safe.  Calls on one stream are serialised by the
 *    library; device scratch is held per stream, so calls on different streams
 *    run concurrently without sharing scratch.
 *  - Return value: PA_OK (0) or a pa_status code; nothing is retried or
 *    silently redirected to the CPU.
 *  - Sizes: sorts (keys-only and key-value) and segments of a segmented sort
 *    take any n that fits device memory: arrays of 2^30 keys or more (past a
 *    look-back word's 30-bit count) are sorted in chunks of fewer and merged
 *    pairwise, stably; histograms, select, partition and the counting sort
 *    take n < 2^32.  Beyond a limit the call returns PA_ERR_TOO_LARGE.
 *
 * Each declaration names the reference interface it replaces
 * (DragonSpit/ParallelAlgorithms, file:line).
 */
#ifndef PA_HIP_H
#define PA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* pa_stream_t;

enum pa_status {
  PA_OK = 0,
  PA_ERR_ARG = 1,       /* bad argument (null pointer with n > 0, k >= n, misaligned key pointer, ...) */
  PA_ERR_OOM = 2,       /* device or pinned host allocation failed */
  PA_ERR_HIP = 3,       /* a HIP runtime call failed; see pa_last_error_string() */
  PA_ERR_TOO_LARGE = 4, /* n beyond the documented limit of the entry point */
  PA_ERR_NO_DEVICE = 5  /* no usable gfx950 device */
};

/* Library info and errors. */
const char* pa_status_string(int status);
const char* pa_last_error_string(void);
int pa_version(void);
/* Selects the device used by subsequent calls on this host thread (default 0). */
int pa_set_device(int device);
/* Blocks until all work enqueued on `stream` has finished. */
int pa_synchronize(pa_stream_t stream);
/* A stream owned by the calling host thread on its current device, created
 * on first use (ordered after null-stream work, like the null stream) and
 * reused by later threads once this one ends.  The drop-in headers enqueue on
 * it, so concurrent callers on different threads run concurrently, each with
 * its own scratch.  NULL (the null stream) if a stream cannot be created.
 * No reference counterpart (the reference runs on TBB's task pool). */
pa_stream_t pa_thread_stream(void);
/* 1 if p points into device memory (hipMalloc or managed), else 0. */
int pa_is_device_pointer(const void* p);
/* Copies `bytes` bytes between any mix of host and device memory, stream-
 * ordered (blocking when host memory is involved).  Used by the drop-in
 * headers where a reference call puts its result in a second buffer. */
int pa_copy(void* dst, const void* src, size_t bytes, pa_stream_t stream);

/* ---- LSD radix sort (ascending, stable) --------------------------------- */

/* Replaces ParallelAlgorithms::SortRadixPar(unsigned* a, unsigned* tmp, size_t n,
 * size_t thr) (RadixSortLsdParallel.h:367-387) and SortRadixPar(unsigned* a,
 * size_t n, size_t thr) (:340-364).  Result in `keys`; `tmp` (n elements) is
 * scratch whose contents are unspecified afterwards; tmp may be NULL, then the
 * library uses its own device scratch. */
int pa_sort_u32(uint32_t* keys, uint32_t* tmp, size_t n, pa_stream_t stream);

/* Allocation-free form of pa_sort_u32.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  keys and tmp
 * must be device memory of n elements each (no host staging), n <= 2^27.  The
 * sort's scratch is a static device array of 22 MiB that every GPU the library
 * runs on holds until the process ends; no call allocates or frees device
 * memory.  All calls on one GPU share that array: calls that overlap on one GPU
 * (different streams or host threads) race, so callers serialise them.
 * Returns PA_ERR_TOO_LARGE for n > 2^27, or when the pass shape chosen for n
 * (PA_SORT_VARIANT, the autotuner) needs more scratch than the array holds. */
int pa_sort_u32_noalloc(uint32_t* keys, uint32_t* tmp, size_t n, pa_stream_t stream);

/* Replaces RadixSortLSDPowerOf2RadixParallel_unsigned_TwoPhase(unsigned* a,
 * unsigned* b, unsigned long n) (RadixSortLsdParallel.h:86-106): sorted result
 * in BOTH a and b (the reference copies it to b, :80-82).  Also serves
 * ..._TwoPhase_DeRandomize (:491-511), which is a no-op stub for n >= 100 in the
 * reference. */
int pa_sort_two_phase_u32(uint32_t* a, uint32_t* b, size_t n, pa_stream_t stream);

/* Merge sort with radix-sorted leaves: replaces parallel_merge_sort_hybrid_radix
 * (ParallelMergeSort.h:255-295: leaves of parallelThreshold keys sorted by the
 * LSD radix sort, then merged upward with merge_parallel_L5).  The leaves
 * keys[i * leaf_keys, (i + 1) * leaf_keys) are sorted (u32: as one segmented
 * batch; u64: one sort each) and merged pairwise, round after round, by the
 * merge-path kernels, ping-ponging with tmp (device or host memory, n keys;
 * NULL: library scratch).  leaf_keys = 0 picks max(65,536, n / 8).  The
 * result is the sorted array, in keys; one radix sort (pa_sort_u32) reaches
 * it with fewer passes over memory. */
int pa_merge_sort_u32(uint32_t* keys, uint32_t* tmp, size_t n, size_t leaf_keys, pa_stream_t stream);
int pa_merge_sort_u64(uint64_t* keys, uint64_t* tmp, size_t n, size_t leaf_keys, pa_stream_t stream);

/* Pass shape per size class (n <= 2^16, <= 2^20, <= 2^24, larger).  The
 * reference fixes its work split by constants (SortRadixPar's quantum,
 * RadixSortLsdParallel.h:374-379); here each class has a built-in tile shape,
 * and pa_autotune_sort_u32 times every candidate shape on this GPU at one size
 * per class (up to max_n keys, device memory it allocates and frees) and makes
 * later u32 sorts in this process use the fastest.  pa_sort_variant_u32
 * returns the candidate index a sort of n u32 keys will use;
 * pa_autotune_reset restores the built-in choice.  The environment variable
 * PA_SORT_VARIANT=i forces candidate i for every size. */
int pa_autotune_sort_u32(size_t max_n, pa_stream_t stream);
int pa_sort_variant_u32(size_t n);
int pa_autotune_reset(void);

/* Tuning targets and candidate shapes.  No reference counterpart: the
 * reference fixes 8-bit digits (SortRadixInnerPar<256, 8>,
 * RadixSortLsdParallel.h:215-216,249-253; the digit loop runs
 * sizeof(key) * 8 / Log2ofPowerOfTwoRadix times) and lists persisting tuned
 * settings as future work (TODO.txt:43).
 *   PA_TUNE_U32      keys-only sorts of 32-bit keys (u32, i32, f32), 4 size
 *                    classes as above;
 *   PA_TUNE_U64      keys-only sorts of 64-bit keys, same classes;
 *   PA_TUNE_SEG_U32  the segmented sort (one class): the shape of its
 *                    onesweep class and the largest segment of its persistent
 *                    queue class.
 * A shape is a tile (block x items), look-back width, staging rounds, lanes
 * per ranking unit and digit width (4 to 8 bits: ceil(key bits / width)
 * passes; an odd count ends with a copy back, which the tuner's timings
 * include). */
enum pa_tune_target { PA_TUNE_U32 = 0, PA_TUNE_U64 = 1, PA_TUNE_SEG_U32 = 2 };
typedef struct pa_sort_shape {
  int block, items, lookback, split, lanes_per_unit, digit_bits, passes;
  uint64_t queue_max_keys; /* PA_TUNE_SEG_U32: largest segment of the queue class */
} pa_sort_shape;
/* Number of candidate shapes of a target (0 for an unknown target). */
int pa_sort_shape_count(int target);
int pa_sort_shape_info(int target, int index, pa_sort_shape* out);
/* Index of the shape a sort of n keys will use (PA_TUNE_SEG_U32: n ignored). */
int pa_sort_shape_choice(int target, size_t n);
/* Sets the shape of one size class (index -1: the built-in choice). */
int pa_sort_shape_set(int target, int size_class, int index);
/* Times every candidate at one size per class (up to max_n keys, on device
 * memory it allocates and frees) and keeps the fastest for this process. */
int pa_autotune_sort(int target, size_t max_n, pa_stream_t stream);
/* Tunes the segmented sort on a batch with these segment lengths (random
 * keys in device memory it allocates and frees; the caller's arrays are not
 * touched). */
int pa_autotune_segmented_u32(const uint64_t* lens, size_t num_segments, pa_stream_t stream);
/* Tuning caches: a JSON file holding the chosen shapes (by name) per device
 * key, the GPU architecture and the HIP runtime version (pa_tune_device_key).
 * Save keeps other devices' entries; load applies only the entry of this
 * device and leaves the rest of the state alone.  With PA_TUNE_CACHE=path in
 * the environment the first sort loads it and every tuning run saves it. */
int pa_tune_cache_save(const char* path);
int pa_tune_cache_load(const char* path);
const char* pa_tune_device_key(void);

/* 64-bit keys, 8 passes.  No reference counterpart on the parallel path (the
 * serial u64 LSD at RadixSortLSD.h:356-400 mis-sorts keys >= 2^32). */
int pa_sort_u64(uint64_t* keys, uint64_t* tmp, size_t n, pa_stream_t stream);

/* float32 keys (IEEE bits) through the order-preserving transform
 * f(u) = u ^ (sign ? 0xFFFFFFFF : 0x80000000), inverse applied on the last
 * pass.  Order: -NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN, NaN payloads
 * kept.  No reference counterpart (RadixSortCommon.h:62-68 is the only order
 * transform in the reference, for signed integers, and is unused). */
int pa_sort_f32(float* keys, float* tmp, size_t n, pa_stream_t stream);

/* float64 keys through the 64-bit form of the same transform, and signed
 * 32/64-bit integers through a sign-bit flip (the order transform the
 * reference writes for signed digits, extractDigitNegate,
 * RadixSortCommon.h:62-68).  These are the element types of the reference's
 * comparison sorts on this path (sort_par, SortParallel.h:24-94, and the
 * merge sorts of ParallelMergeSort.h, which its benchmark runs on doubles,
 * ParallelMergeSortBenchmark.cpp:61). */
int pa_sort_f64(double* keys, double* tmp, size_t n, pa_stream_t stream);
int pa_sort_i32(int32_t* keys, int32_t* tmp, size_t n, pa_stream_t stream);
int pa_sort_i64(int64_t* keys, int64_t* tmp, size_t n, pa_stream_t stream);

/* Stable key-value sorts: keys ascending, vals permuted with their keys, ties
 * keep input order.  No reference counterpart (new component). */
int pa_sort_kv_u32_u32(uint32_t* keys, uint32_t* vals, uint32_t* ktmp, uint32_t* vtmp, size_t n, pa_stream_t stream);
int pa_sort_kv_u64_u32(uint64_t* keys, uint32_t* vals, uint64_t* ktmp, uint32_t* vtmp, size_t n, pa_stream_t stream);

/* ---- Histograms ----------------------------------------------------------- */

/* All 4 byte-digit histograms of u32 keys in one read.  counts: 4 x 256
 * uint64 (row d = byte d).  Replaces HistogramByteComponentsParallel<256,8>
 * (HistogramParallel.h:249-329), whose 8-row return shape the drop-in header
 * rebuilds. */
int pa_hist_digits_u32(const uint32_t* keys, size_t n, uint64_t* counts, pa_stream_t stream);

/* Histogram of (key >> shift) & (2^bits - 1), bits in 1..12; counts has 2^bits
 * uint64.  Replaces HistogramOneByteComponentParallel<R,L>(unsigned[], l, r,
 * shift, thr) (HistogramParallel.h:535-576) and the _2 variant (:466-533). */
int pa_hist_u32(const uint32_t* keys, size_t n, unsigned shift, unsigned bits, uint64_t* counts, pa_stream_t stream);

/* As pa_hist_u32 over uint64 keys, shift in 0..63.  Extension of
 * HistogramOneByteComponentParallel (HistogramParallel.h:535-576), which is
 * 32-bit only, for the served uint64 jobs. */
int pa_hist_u64(const uint64_t* keys, size_t n, unsigned shift, unsigned bits, uint64_t* counts, pa_stream_t stream);

/* Per-work-quantum histogram of byte `which_byte` over keys[first, last):
 * counts[q][b] (nq rows of 256 uint64) with q = index / quantum on the
 * absolute index; rows outside [first/quantum, (last-1)/quantum] stay zero.
 * Replaces HistogramByteComponentsQCPar<256,8>(in, l, r, workQuanta,
 * numberOfQuantas, whichByte, thr) (HistogramParallel.h:444-464, with
 * QCParInner :399-442 and AcrossWorkQuantasQC :331-397): first = l,
 * last = r + 1. */
int pa_hist_quanta_u32(const uint32_t* keys, size_t first, size_t last, size_t quantum, size_t nq, unsigned which_byte,
                       uint64_t* counts, pa_stream_t stream);

/* start[q][b] = sum_{b'<b} size[b'] + sum_{q'<q} count[q'][b] for byte `digit`
 * of keys[0, n) (nq rows of 256 uint64).  Replaces ComputeStartOfBinsPar<256,8>
 * (in, size, workQuanta, numberOfQuantas, digit, thr)
 * (RadixSortLsdParallel.h:108-165). */
int pa_start_of_bins_u32(const uint32_t* keys, size_t n, size_t quantum, size_t nq, unsigned digit, uint64_t* start,
                         pa_stream_t stream);

/* 256-bin histogram of a byte array.  Replaces
 * HistogramOneByteComponentParallel{,_2,_3,_4}<256>(unsigned char[], l, r, thr)
 * (HistogramParallel.h:16-247). */
int pa_hist_u8(const uint8_t* bytes, size_t n, uint64_t* counts, pa_stream_t stream);

/* ---- Counting sort and fill ------------------------------------------------ */

/* Sorts a byte array in place by counting: one histogram read, then each bin's
 * value written over its range.  Replaces counting_sort_parallel(a, a_size)
 * and counting_sort_parallel_inner<256>(a, l, r, threshold_count,
 * threshold_fill) (CountingSortParallel.h:29-90).  n < 2^32. */
int pa_counting_sort_u8(uint8_t* bytes, size_t n, pa_stream_t stream);

/* Writes n copies of the elem_bytes-byte value *value (a host pointer) to a,
 * which must be aligned to elem_bytes (1, 2, 4, 8 or 16).  Replaces
 * parallel_fill(src, value, l, r, thr) (FillParallel.h:22-66) with
 * a = src + l, n = r - l. */
int pa_fill(void* a, size_t elem_bytes, const void* value, size_t n, pa_stream_t stream);

/* ---- Radix select ---------------------------------------------------------- */

/* *result = k-th smallest (0-based) of keys[0..n).  keys is left unmodified.
 * Replaces SelectRadix(unsigned[], size_t len, size_t k) and
 * SelectRadix(unsigned[], size_t start, size_t len, size_t k)
 * (RadixSelect.h:139-168).  result may be host or device memory. */
int pa_select_u32(const uint32_t* keys, size_t n, size_t k, uint32_t* result, pa_stream_t stream);

/* As pa_select_u32 over uint64 keys (digits 56 .. 0).  Extension of SelectRadix
 * (RadixSelect.h:139-168), which is 32-bit only, for the served uint64 jobs. */
int pa_select_u64(const uint64_t* keys, size_t n, size_t k, uint64_t* result, pa_stream_t stream);

/* As pa_select_u32, and also stores the selected value at keys[k] -- the
 * post-condition SelectRadix's callers rely on (they read a[k] after the call,
 * RadixSelectBenchmark.cpp:57-58).  The rest of keys is unchanged (the
 * reference leaves it reordered and partly overwritten, RadixSelect.h:52). */
int pa_select_radix_u32(uint32_t* keys, size_t n, size_t k, uint32_t* result, pa_stream_t stream);

/* ---- Radix partition (nth_element) ---------------------------------------- */

/* Partitions keys[0..n) in place so that keys[k] holds the k-th smallest key,
 * no key before it is larger and no key after it is smaller; *result = that
 * key.  Replaces PartitionRadix(unsigned[], size_t aLength, size_t k) and
 * PartitionRadix(unsigned[], size_t aStart, size_t aLength, size_t k)
 * (RadixPartition.h:94-145).  The order of the other keys within each side is
 * their input order (the reference's is its in-place swap order). */
int pa_partition_u32(uint32_t* keys, size_t n, size_t k, uint32_t* result, pa_stream_t stream);

/* The same for several ranks ks[0..num_ks) (ascending, each < n) at once.
 * Replaces PartitionRadix(unsigned[], size_t aLength, size_t k[], size_t
 * kLength) and the aStart/kStart form (RadixPartition.h:147-184). */
int pa_partition_multi_u32(uint32_t* keys, size_t n, const size_t* ks, size_t num_ks, pa_stream_t stream);

/* ---- Merge ----------------------------------------------------------------- */

/* out[0..na+nb) = stable merge of sorted a[0..na) and b[0..nb) (ties take a
 * first).  Replaces merge_parallel_L5<T>(t, p1, r1, p2, r2, a, p3, thr)
 * (ParallelMerge.h:285-315) and the rest of the ParallelMerge.h family, which
 * the drop-in header maps onto it.  out must not overlap a or b. */
int pa_merge_u32(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, uint32_t* out, pa_stream_t stream);
int pa_merge_u64(const uint64_t* a, size_t na, const uint64_t* b, size_t nb, uint64_t* out, pa_stream_t stream);

/* In-place merge of the adjacent sorted runs t[0..n1) and t[n1..n1+n2):
 * merged into device scratch, then copied back over t.  Replaces the in-place
 * family of ParallelMerge.h (merge_truly_in_place :400-446, merge_in_place
 * :448-494, p_merge_in_place_2 :499-557, p_merge_truly_in_place :591-651,
 * p_merge_in_place_adaptive :653-668 and the preventative-adaptive forms
 * :670-726), whose result is the same merged run. */
int pa_merge_inplace_u32(uint32_t* t, size_t n1, size_t n2, pa_stream_t stream);
int pa_merge_inplace_u64(uint64_t* t, size_t n1, size_t n2, pa_stream_t stream);

/* ---- Segmented (batched) sort ---------------------------------------------- */

/* Sorts num_segments independent arrays in one batch.  Segment i is
 * keys[offsets[i] .. offsets[i+1]); offsets (num_segments + 1 entries,
 * non-decreasing) is host memory; keys/tmp are device memory.  Each segment's
 * result equals pa_sort_u32 on that segment.  No reference counterpart (the
 * reference only benchmarks bundling small tasks, BundleSmallWorkItems.cpp:36-144). */
int pa_segmented_sort_u32(uint32_t* keys, uint32_t* tmp, const uint64_t* offsets, size_t num_segments,
                          pa_stream_t stream);

/* As pa_segmented_sort_u32 for segments that live in separate device
 * buffers: segment i is keys[i][0 .. lens[i]) (keys and lens are host arrays;
 * each keys[i] is device memory).  Used to batch independent requests. */
int pa_segmented_sort_ptrs_u32(uint32_t* const* keys, const uint64_t* lens, size_t num_segments, pa_stream_t stream);

/* Reusable plan for sorting the same device arrays again and again (a
 * serving loop that re-sorts buffers of fixed shapes): the segment tables and
 * the library's scratch are set up once, and the launch sequence is captured
 * into a HIP graph, so each run is one graph launch that sorts the arrays'
 * current contents in place.  No reference counterpart (the reference sorts
 * one array per call).  The arrays must stay allocated while the plan lives;
 * runs of one plan must not overlap. */
typedef struct pa_segplan pa_segplan;
pa_segplan* pa_segmented_plan_create_u32(uint32_t* const* keys, const uint64_t* lens, size_t num_segments);
int pa_segmented_plan_run(pa_segplan* plan, pa_stream_t stream);
void pa_segmented_plan_destroy(pa_segplan* plan);

/* ---- Serving: continuous batching of sort / select / histogram jobs --------- */
/* No reference counterpart (the reference is a blocking call-per-array
 * library; BundleSmallWorkItems.cpp:36-144 and MemoryUsage.cpp:20-86 hold
 * the nearest ideas).  Jobs name device-resident arrays; results are what the
 * same call returns alone. */

enum pa_op { PA_OP_SORT = 0, PA_OP_SELECT = 1, PA_OP_HIST = 2 };
enum pa_key_type { PA_KEY_U32 = 0, PA_KEY_U64 = 1 };

typedef struct pa_job {
  int op;              /* pa_op */
  int key_type;        /* pa_key_type */
  void* keys;          /* device memory, n keys; sorted in place for PA_OP_SORT */
  size_t n;
  size_t k;            /* PA_OP_SELECT: rank */
  unsigned shift, bits; /* PA_OP_HIST: digit (key >> shift) & (2^bits - 1) */
  void* out;           /* device memory: one key (select) or 2^bits uint64 (hist) */
} pa_job;

typedef struct pa_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t launches;           /* units launched (a coalesced batch counts once) */
  uint64_t coalesced_batches;  /* batches of two or more small sorts */
  uint64_t coalesced_jobs;     /* jobs that ran inside such batches */
  uint64_t peak_bytes_in_flight; /* peak device scratch held by the server's streams */
  uint64_t pending;            /* jobs waiting for admission */
  uint64_t admission_waits;    /* units that waited for units in flight to fit the budget */
  uint64_t scratch_releases;   /* idle streams' scratch blocks given back to fit a unit */
  uint64_t oversize_units;     /* units larger than the budget alone, run by themselves */
  uint64_t bytes_held;         /* device scratch the server's streams hold now */
} pa_server_stats;

typedef struct pa_server pa_server;

/* hbm_budget_bytes: bound on the device scratch the library holds for the
 * server's streams at once (0 = 64 GiB); the job arrays are the caller's and
 * are not counted;
 * num_streams: concurrent HIP streams; coalesce_max_keys: u32 sorts up to
 * this size are batched together (0 = 1 Mi keys).  NULL on failure. */
pa_server* pa_server_create(size_t hbm_budget_bytes, int num_streams, size_t coalesce_max_keys);
int pa_server_submit(pa_server* server, const pa_job* job, uint64_t* ticket);
/* Blocks until the job is done; returns its status and its latency
 * (completion - submission, host steady clock).  Each ticket is waited once. */
int pa_server_wait(pa_server* server, uint64_t ticket, double* latency_ms);
int pa_server_drain(pa_server* server);
int pa_server_get_stats(pa_server* server, pa_server_stats* out);
void pa_server_destroy(pa_server* server);
/* Device scratch (bytes) the library allocates for this job run alone on a
 * stream holding none: the figure the server's admission charges.  Computed
 * from the same layout functions the sort, select, histogram entry points
 * allocate with; the job's arrays are not included. */
int pa_job_scratch_bytes(const pa_job* job, size_t* bytes);

/* ---- Measurement helpers (no reference counterpart) ------------------------ */

/* Plain device copy kernel, dst[i] = src[i]; the known-size calibration point
 * for HBM byte counters and the achievable-bandwidth probe. */
int pa_copy_u32(uint32_t* dst, const uint32_t* src, size_t n, pa_stream_t stream);

/* Per-kernel timing with HIP events recorded on the launch stream around each
 * kernel the library launches.  enable=1 starts recording, 0 stops. */
int pa_profile_enable(int enable);
/* Restricts kernel timing to kernels whose name starts with `prefix` (NULL or
 * "" = every kernel), so that a benchmark times its dominant kernel without
 * event records around every other launch. */
int pa_profile_only(const char* prefix);
/* Totals for kernels whose name starts with `prefix` since the last reset;
 * waits for the recorded events. */
int pa_profile_query(const char* prefix, double* total_ms, uint64_t* launches);
int pa_profile_reset(void);

/* ---- Deterministic host-side input generators ------------------------------ */

/* std::mt19937_64(seed) outputs truncated to 32 bits (the reference's
 * generator: ParallelAlgorithms.cpp:91,114-115).  Host memory. */
void pa_gen_mt19937_u32(uint64_t seed, uint32_t* out, size_t n);
/* Full 64-bit std::mt19937_64(seed) outputs. */
void pa_gen_mt19937_u64(uint64_t seed, uint64_t* out, size_t n);
/* Device-side uniform keys for large synthetic workloads: dst[i] =
 * splitmix64(seed + i * 0x9E3779B97F4A7C15) (truncated for u32). */
int pa_fill_random_u32(uint32_t* dst, size_t n, uint64_t seed, pa_stream_t stream);
int pa_fill_random_u64(uint64_t* dst, size_t n, uint64_t seed, pa_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PA_HIP_H */
