"""GPU parity of the segmented (batched) sort: every segment must equal the
oracle's SortRadixPar of that segment and the plain per-array GPU call, bit
for bit.  Covers every size class (empty, one key, LDS-resident small
segments, the persistent queue's chunked segments, single- and multi-tile
onesweep segments) and presorted segments."""
import numpy as np
import pytest

from gpu_util import dev_empty, to_dev, to_host

pytestmark = pytest.mark.gpu


def _segments(seed, count, lo, hi, presorted_frac=0.05):
    import parallelalgorithms_amd as P
    r = P.gen_u64(seed, 2 * count + 1)
    u = (r[:count] >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    lens = np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.uint64)
    offs = np.zeros(count + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    keys = P.gen_u32(seed + 1, int(offs[-1]))
    pres = ((r[count:2 * count] >> np.uint64(11)).astype(np.float64) / float(1 << 53)) < presorted_frac
    for i in np.nonzero(pres)[0]:
        a, b = int(offs[i]), int(offs[i + 1])
        keys[a:b] = np.sort(keys[a:b])
    return keys, offs


def _check(keys, offs, got, orc=None, oracle_every=0):
    for i in range(len(offs) - 1):
        a, b = int(offs[i]), int(offs[i + 1])
        expect = np.sort(keys[a:b])
        if orc is not None and oracle_every and i % oracle_every == 0:
            expect = orc.sort_radix_par_u32(keys[a:b])
        np.testing.assert_array_equal(got[a:b], expect, err_msg=f"segment {i} [{a},{b})")


def test_segmented_edge_classes(pa, orc):
    # Around every class edge: LDS 4096 / 8192 / 16384, onesweep tiles of 24576.
    lens = [0, 1, 2, 3, 63, 64, 4095, 4096, 4097, 8191, 8192, 8193, 16_383, 16_384, 16_385, 20_000, 24_576, 24_577,
            0, 1, 100_003, 7]
    offs = np.zeros(len(lens) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    import parallelalgorithms_amd as P
    keys = P.gen_u32(77, int(offs[-1]))
    keys[offs[5]:offs[6]] = 0x5A5A5A5A  # a constant segment
    keys[offs[9]:offs[10]] |= 0xFFFFFF00  # keys sharing the padding's high digits (8,191 keys, LDS class)
    keys[offs[17]:offs[18]] = 0xFFFFFFFF  # every key equal to the padding (24,577 keys: a full and a one-key onesweep tile)
    d = to_dev(keys)
    pa.segmented_sort_u32(d, dev_empty(keys.size, np.uint32), offs)
    _check(keys, offs, to_host(d, np.uint32), orc, oracle_every=1)


def _launches(P, fn, *prefixes):
    """Launch counts of the named kernels during fn(), from the library's
    event profiler."""
    P.profile_reset()
    P.profile_only(None)
    P.profile_enable(True)
    try:
        fn()
    finally:
        P.profile_enable(False)
    n = {p: P.profile_query(p)[1] for p in prefixes}
    P.profile_reset()
    return n


def test_segmented_queue_class(pa, orc):
    # The persistent queue sorts segments of 16,384 < len <= 65,536 keys in
    # 8,192-key chunks when the batch balances over its 512 workgroups (the
    # longest queued segment at most twice the average load).  Edge lengths:
    # chunk edges, a ragged last chunk, 65,536 +- 1, with constant, presorted,
    # descending and shared-digit segments (passes where one digit holds
    # every key), among 1,600 fillers of 20K-30K keys so the queue is used;
    # the longer ones go to the onesweep class.
    import parallelalgorithms_amd as P
    edges = [16_385, 24_576, 32_767, 32_768, 32_769, 57_345, 65_535, 65_536, 65_537, 140_000,
             40_000, 40_000, 40_000, 40_000, 16_385, 60_000, 600_000]
    r = P.gen_u64(80, 1600)
    fill = (20_000 + (r % np.uint64(10_000))).astype(np.int64).tolist()
    lens = edges[:8] + fill[:800] + edges[8:] + fill[800:]
    offs = np.zeros(len(lens) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    keys = P.gen_u32(79, int(offs[-1]))
    seg = lambda i: slice(int(offs[i]), int(offs[i + 1]))
    c = 808  # index of edges[8] in lens; c + 9 onwards are queued fillers
    keys[seg(c + 10)] = 0xA5A5A5A5                                       # constant
    keys[seg(c + 11)] = np.sort(keys[seg(c + 11)])                       # presorted
    keys[seg(c + 12)] = np.sort(keys[seg(c + 12)])[::-1]                 # descending
    keys[seg(c + 13)] = (keys[seg(c + 13)] & 0xFF00FF00) | 0x00120034    # digits 0 and 2 constant
    keys[seg(c + 14)] = keys[seg(c + 14)] & 0x3                          # many equal keys
    keys[seg(c + 15)] = keys[seg(c + 15)] | 0xFFFFFF00                   # real keys equal to the padding's digits
    keys[seg(c + 16)] = 0xFFFFFFFF                                       # every key equal to the padding
    for tmp in (dev_empty(keys.size, np.uint32), None):
        d = to_dev(keys)
        counts = _launches(P, lambda: pa.segmented_sort_u32(d, tmp, offs), "k_seg_queue", "k_onesweep_seg")
        assert counts == {"k_seg_queue": 1, "k_onesweep_seg": 4}, counts
        _check(keys, offs, to_host(d, np.uint32), orc, oracle_every=97)


def test_segmented_queue_needs_balance(pa):
    # Too few mid-size segments to balance the workgroups: all of them go to
    # the onesweep class, and the result is the same.
    import parallelalgorithms_amd as P
    lens = [20_000, 131_072, 70_001, 3]
    offs = np.zeros(len(lens) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    keys = P.gen_u32(81, int(offs[-1]))
    d = to_dev(keys)
    assert _launches(P, lambda: pa.segmented_sort_u32(d, None, offs), "k_seg_queue") == {"k_seg_queue": 0}
    _check(keys, offs, to_host(d, np.uint32))


def test_segmented_tmp_null_and_offset_base(pa):
    import parallelalgorithms_amd as P
    keys = P.gen_u32(78, 300_000)
    offs = np.array([1000, 1000, 5000, 70_000, 290_000], dtype=np.uint64)  # starts past 0
    d = to_dev(keys)
    pa.segmented_sort_u32(d, None, offs)
    got = to_host(d, np.uint32)
    np.testing.assert_array_equal(got[:1000], keys[:1000])        # outside: untouched
    np.testing.assert_array_equal(got[290_000:], keys[290_000:])
    _check(keys, offs, got)


def test_segmented_equals_per_array_calls(pa, orc):
    keys, offs = _segments(1239, 400, 64, 400_000)
    d = to_dev(keys)
    pa.segmented_sort_u32(d, dev_empty(keys.size, np.uint32), offs)
    got = to_host(d, np.uint32)
    single = to_dev(keys)
    tmp = dev_empty(keys.size, np.uint32)
    for i in range(len(offs) - 1):
        a, b = int(offs[i]), int(offs[i + 1])
        pa.sort_u32(single[a:b], tmp[a:b])
    np.testing.assert_array_equal(got, to_host(single, np.uint32))
    _check(keys, offs, got, orc, oracle_every=37)


def test_segmented_rejects_decreasing_offsets(pa):
    import parallelalgorithms_amd as P
    d = dev_empty(10, np.uint32)
    with pytest.raises(P.PAError):
        pa.segmented_sort_u32(d, None, np.array([0, 5, 3], dtype=np.uint64))


def test_segmented_pointer_table(pa):
    """Separate device buffers sorted in one batch (the server's coalescing path)."""
    import parallelalgorithms_amd as P
    lens = [0, 1, 5, 4096, 4097, 70_001, 3, 250_000]
    hosts = [P.gen_u32(90 + i, m) for i, m in enumerate(lens)]
    devs = [to_dev(h) if h.size else dev_empty(1, np.uint32)[:0] for h in hosts]
    pa.segmented_sort_ptrs_u32(devs)
    for h, d in zip(hosts, devs):
        np.testing.assert_array_equal(to_host(d, np.uint32) if h.size else h, np.sort(h))


def test_segmented_plan_graph_replays(pa):
    """The graph-captured plan sorts the arrays' current contents on every run:
    three runs with fresh keys written eagerly (and synchronised) between them,
    each checked against np.sort and against the one-shot batched call."""
    import torch
    import parallelalgorithms_amd as P
    lens = [1, 7, 4096, 4097, 30_000, 24_577, 100_000, 2, 333_333]
    devs = [torch.empty(m, dtype=torch.int32, device="cuda") for m in lens]
    plan = P.SegmentedPlan(devs)
    try:
        for it in range(3):
            for i, d in enumerate(devs):
                P.fill_random_u32(d, 1000 * it + i)
            if it == 1:  # a presorted array in the mix
                devs[6].copy_(torch.arange(lens[6], dtype=torch.int32, device="cuda"))
            torch.cuda.synchronize()
            before = [to_host(d, np.uint32).copy() for d in devs]
            copies = [d.clone() for d in devs]
            plan.run()
            torch.cuda.synchronize()
            pa.segmented_sort_ptrs_u32(copies)
            for h, d, c in zip(before, devs, copies):
                got = to_host(d, np.uint32)
                np.testing.assert_array_equal(got, np.sort(h))
                np.testing.assert_array_equal(got, to_host(c, np.uint32))
    finally:
        plan.close()


def test_segmented_plan_graph_replays_queue(pa):
    """A captured plan whose segments go through the persistent queue (1,200
    arrays of 20K-40K keys, enough to balance its workgroups): the queue's
    counter is cleared by a kernel inside the graph, so every replay sorts the
    arrays' current contents.  Three replays with fresh keys written eagerly
    between them, each checked against np.sort and the one-shot batched call."""
    import torch
    import parallelalgorithms_amd as P
    r = P.gen_u64(82, 1200)
    lens = (20_000 + (r % np.uint64(20_000))).astype(np.int64)
    offs = np.zeros(len(lens) + 1, dtype=np.int64)
    offs[1:] = np.cumsum(lens)
    buf = torch.empty(int(offs[-1]), dtype=torch.int32, device="cuda")
    devs = [buf[int(offs[i]):int(offs[i + 1])] for i in range(len(lens))]
    plan = P.SegmentedPlan(devs)
    try:
        for it in range(3):
            P.fill_random_u32(buf, 500 + it)
            if it == 2:
                devs[5].copy_(torch.arange(int(lens[5]), dtype=torch.int32, device="cuda"))  # presorted
            torch.cuda.synchronize()
            before = to_host(buf, np.uint32).copy()
            copy = buf.clone()
            plan.run()
            torch.cuda.synchronize()
            pa.segmented_sort_ptrs_u32([copy[int(offs[i]):int(offs[i + 1])] for i in range(len(lens))])
            torch.cuda.synchronize()
            got = to_host(buf, np.uint32)
            np.testing.assert_array_equal(got, to_host(copy, np.uint32))
            for i in range(len(lens)):
                a, b = int(offs[i]), int(offs[i + 1])
                np.testing.assert_array_equal(got[a:b], np.sort(before[a:b]), err_msg=f"replay {it} segment {i}")
    finally:
        plan.close()
