"""GPU parity of radix select and merge against the oracle.

Select mirrors RadixSelectBenchmark.cpp:33-67 (SelectRadix vs nth_element at
k = n - 21) plus the configured ranks k = n/2 and n/100; the selected value is
unique, so equality is exact.  Merge mirrors merge_parallel_L5
(ParallelMerge.h:285-315): for scalar keys the merged output is unique.
"""
import numpy as np
import pytest

from gpu_util import dev_empty, make_u32, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dist", ["uniform", "presorted", "constant", "reversed", "nearly", "few"])
@pytest.mark.parametrize("n", [1, 2, 7, 100, 4097, 1_000_003])
def test_select_matches_oracle(pa, orc, dist, n):
    a = make_u32(dist, n)
    d = to_dev(a)
    for k in sorted({0, n // 100, n // 2, max(0, n - 21), n - 1}):
        expect, _ = orc.select_radix_u32(a, k)
        res = np.zeros(1, dtype=np.uint32)
        pa.select_u32(d, k, res)  # host result: synchronous
        assert int(res[0]) == expect, (dist, n, k)
    np.testing.assert_array_equal(to_host(d, np.uint32), a)  # input untouched


@pytest.mark.parametrize("low_bits", [1, 8, 12, 16, 24, 31])
def test_select_shared_high_bytes(pa, orc, low_bits):
    """Keys that share their high bytes (only the low `low_bits` vary): the
    select fixes the shared bytes from the keys' minimum and maximum and
    starts its passes at the first byte that differs."""
    import parallelalgorithms_amd as P
    n = 200_003
    r = P.gen_u32(31 + low_bits, n)
    a = (np.uint32(0xA5C3E1F7) & ~np.uint32((1 << low_bits) - 1)) | (r & np.uint32((1 << low_bits) - 1))
    d = to_dev(a)
    for k in sorted({0, n // 100, n // 2, n - 21, n - 1}):
        expect, _ = orc.select_radix_u32(a, k)
        res = np.zeros(1, dtype=np.uint32)
        pa.select_u32(d, k, res)
        assert int(res[0]) == expect, (low_bits, k)


def test_select_device_result_and_host_keys(pa, orc):
    n = 300_001
    a = make_u32("uniform", n, seed=21)
    k = 12345
    expect, _ = orc.select_radix_u32(a, k)
    r = dev_empty(1, np.uint32)
    pa.select_u32(to_dev(a), k, r)
    assert int(to_host(r, np.uint32)[0]) == expect
    res = np.zeros(1, dtype=np.uint32)
    pa.select_u32(a, k, res)
    assert int(res[0]) == expect


def test_select_sample_band_miss_falls_back(pa, orc):
    """The select predicts the answer's top-byte bins from runs of keys at an
    even stride.  Keys at exactly those positions are drawn from the low
    quarter of the key range and every other key from the high quarter, so
    the predicted bins miss the answer (and, for other ranks, hold it): the
    result must not depend on the prediction."""
    n = 1_000_003
    runs, run = 256, 128  # the sampling layout: 256 runs of 128 keys
    import parallelalgorithms_amd as P
    rnd = P.gen_u32(77, n)
    a = (rnd & np.uint32(0x3FFFFFFF)) | np.uint32(0xC0000000)
    for i in range(runs):
        base = i * (n - run) // (runs - 1)
        a[base:base + run] = rnd[base:base + run] & np.uint32(0x3FFFFFFF)
    d = to_dev(a)
    for k in (0, 100, n // 100, n // 2, n - 1):
        expect, _ = orc.select_radix_u32(a, k)
        res = np.zeros(1, dtype=np.uint32)
        pa.select_u32(d, k, res)
        assert int(res[0]) == expect, k


def test_select_radix_host_keys_device_result(pa, orc):
    """SelectRadix's post-condition a[k] == k-th smallest with the keys in
    host memory and the result in device memory: the value reaches both."""
    n = 200_003
    a = make_u32("uniform", n, seed=22)
    for k in (0, n // 2, n - 1):
        expect, _ = orc.select_radix_u32(a, k)
        h = a.copy()
        r = dev_empty(1, np.uint32)
        pa.select_radix_u32(h, k, r)
        assert int(to_host(r, np.uint32)[0]) == expect, k
        assert int(h[k]) == expect, k
        mask = np.ones(n, dtype=bool)
        mask[k] = False
        np.testing.assert_array_equal(h[mask], a[mask])  # the other keys untouched


def test_select_rejects_bad_k(pa):
    import parallelalgorithms_amd as P
    a = to_dev(np.arange(10, dtype=np.uint32))
    res = np.zeros(1, dtype=np.uint32)
    with pytest.raises(P.PAError):
        pa.select_u32(a, 10, res)


def _make_u64(dist, n, seed=77):
    import parallelalgorithms_amd as P
    if dist == "uniform":
        return P.gen_u64(seed, n)
    if dist == "presorted":
        return np.arange(n, dtype=np.uint64) * np.uint64(0x100000001)
    if dist == "constant":
        return np.full(n, 0x5A5A5A5A_A5A5A5A5, dtype=np.uint64)
    if dist == "high":  # only the top bytes vary: every low digit is one bin
        return (P.gen_u64(seed, n) % np.uint64(5)) << np.uint64(56)
    raise ValueError(dist)


@pytest.mark.parametrize("dist", ["uniform", "presorted", "constant", "high"])
@pytest.mark.parametrize("n", [1, 5, 4097, 1_000_003])
def test_select_u64(pa, dist, n):
    """u64 select has no reference counterpart (SelectRadix is 32-bit): the
    answer is fixed by the order statistic itself, np.sort(a)[k]."""
    a = _make_u64(dist, n)
    srt = np.sort(a)
    d = to_dev(a)
    for k in sorted({0, n // 100, n // 2, max(0, n - 21), n - 1}):
        res = np.zeros(1, dtype=np.uint64)
        pa.select_u64(d, k, res)
        assert int(res[0]) == int(srt[k]), (dist, n, k)
    r = dev_empty(1, np.uint64)
    pa.select_u64(d, n // 2, r)
    assert int(to_host(r, np.uint64)[0]) == int(srt[n // 2])
    np.testing.assert_array_equal(to_host(d, np.uint64), a)


@pytest.mark.parametrize("dist", ["uniform", "presorted", "few", "nearly"])
@pytest.mark.parametrize("n", [32_767, 32_768, 32_769, 33_000, 65_536 + 129])
def test_select_sample_boundary(pa, orc, dist, n):
    """Sizes around the 32,768-key sample: below it every key is sampled and
    the band is off; from it on the sample is 256 runs of 128 keys read by 16
    blocks (at n = 32,768 exactly every key once, just above it runs that
    nearly touch), and the band the first read derives from their counts
    decides which keys are compacted.  u32 against the oracle, u64 against
    the order statistic."""
    a = make_u32(dist, n)
    d = to_dev(a)
    ks = sorted({0, 1, n // 100, n // 3, n // 2, n - 21, n - 1})
    for k in ks:
        expect, _ = orc.select_radix_u32(a, k)
        res = np.zeros(1, dtype=np.uint32)
        pa.select_u32(d, k, res)
        assert int(res[0]) == expect, (dist, n, k)
    np.testing.assert_array_equal(to_host(d, np.uint32), a)
    a64 = (a.astype(np.uint64) << np.uint64(32)) | a[::-1].astype(np.uint64)
    srt = np.sort(a64)
    d64 = to_dev(a64)
    for k in ks:
        res = np.zeros(1, dtype=np.uint64)
        pa.select_u64(d64, k, res)
        assert int(res[0]) == int(srt[k]), ("u64", dist, n, k)


def _sorted_u32(seed, n, mod=None):
    import parallelalgorithms_amd as P
    x = P.gen_u32(seed, n)
    if mod:
        x = x % np.uint32(mod)
    return np.sort(x)


@pytest.mark.parametrize("na,nb", [(0, 0), (0, 9), (9, 0), (1, 1), (4096, 1), (5000, 70_001), (1_000_003, 999_999)])
@pytest.mark.parametrize("mod", [None, 100])
def test_merge_u32_matches_oracle(pa, orc, na, nb, mod):
    x, y = _sorted_u32(37, na, mod), _sorted_u32(38, nb, mod)
    t = np.concatenate([x, y]).astype(np.uint32)
    expect = np.zeros(na + nb, dtype=np.uint32)
    if na + nb:
        wrap = lambda p, m: (p + m - 1) % (1 << 64)
        orc.merge_parallel_l5_u32(t, 0, wrap(0, na), na, wrap(na, nb), expect, 0)
    out = dev_empty(max(1, na + nb), np.uint32)
    dx = to_dev(x) if na else dev_empty(1, np.uint32)
    dy = to_dev(y) if nb else dev_empty(1, np.uint32)
    import parallelalgorithms_amd as P
    P.call("pa_merge_u32", dx.data_ptr(), na, dy.data_ptr(), nb, out.data_ptr(), P._stream(None))
    got = to_host(out, np.uint32)[: na + nb]
    np.testing.assert_array_equal(got, expect)


def test_merge_u64_and_host_pointers(pa, orc):
    import parallelalgorithms_amd as P
    x = np.sort(P.gen_u64(39, 300_000))
    y = np.sort(P.gen_u64(40, 200_001))
    t = np.concatenate([x, y])
    expect = np.zeros(t.size, dtype=np.uint64)
    orc.merge_parallel_l5_u64(t, 0, x.size - 1, x.size, t.size - 1, expect, 0)
    out = dev_empty(t.size, np.uint64)
    pa.merge_u64(to_dev(x), to_dev(y), out)
    np.testing.assert_array_equal(to_host(out, np.uint64), expect)
    hout = np.zeros(t.size, dtype=np.uint64)
    pa.merge_u64(x, y, hout)  # host memory: staged, synchronous
    np.testing.assert_array_equal(hout, expect)


def _skewed_runs(kind, na, nb):
    """Sorted u32 runs whose values are spread so that a line through two
    probed points predicts the merge split badly (the tile splits must still
    be exact): cubic growth, one run entirely below the other, a lone huge
    outlier at the end of each run, long runs of one repeated value, and
    runs that interleave only in a narrow band."""
    i = np.arange(max(na, nb), dtype=np.float64)
    if kind == "cubic":
        x = (i[:na] ** 3 / max(1.0, (na - 1)) ** 3 * 4.0e9).astype(np.uint32)
        y = (np.sqrt(i[:nb] / max(1, nb - 1)) * 4.0e9).astype(np.uint32)
    elif kind == "disjoint":
        x = np.arange(na, dtype=np.uint32)
        y = np.arange(nb, dtype=np.uint32) + np.uint32(na + 7)
    elif kind == "outlier":
        x = np.arange(na, dtype=np.uint32)
        y = np.arange(nb, dtype=np.uint32) * np.uint32(2)
        x[-1] = 0xFFFFFFFF
        y[-1] = 0xFFFFFFFE
    elif kind == "plateaus":
        x = (np.arange(na) // 50_000).astype(np.uint32) * np.uint32(3)
        y = (np.arange(nb) // 70_001).astype(np.uint32) * np.uint32(3) + np.uint32(1)
    else:  # band: the runs overlap only in the middle tenth
        x = np.sort(np.concatenate([np.zeros(na * 9 // 20, np.uint32),
                                    np.random.default_rng(41).integers(1000, 2000, na - na * 9 // 20 - na * 9 // 20, dtype=np.uint32),
                                    np.full(na * 9 // 20, 0xFFFFFFFF, np.uint32)]))
        y = np.sort(np.random.default_rng(42).integers(900, 2100, nb, dtype=np.uint32))
    return np.sort(x), np.sort(y)


@pytest.mark.parametrize("kind", ["cubic", "disjoint", "outlier", "plateaus", "band"])
def test_merge_split_search_skewed_values(pa, kind):
    """The tile splits are found by an interpolation-placed search in HBM;
    whatever the value spread, the merged output must be the stable merge
    (checked against numpy's stable sort of the concatenation)."""
    na, nb = 1_000_003, 777_777
    x, y = _skewed_runs(kind, na, nb)
    expect = np.sort(np.concatenate([x, y]), kind="stable")
    out = dev_empty(na + nb, np.uint32)
    pa.merge_u32(to_dev(x), to_dev(y), out)
    np.testing.assert_array_equal(to_host(out, np.uint32), expect)


def test_merge_u64_split_search_near_top(pa):
    """u64 keys near 2^64, where differences of key values lose their low bits
    as doubles: the split search may only use them to place probes."""
    base = np.uint64(0xFFFFFFFF00000000)
    x = base + np.arange(0, 2_000_000, 2, dtype=np.uint64)
    y = base + np.arange(1, 1_500_001, 3, dtype=np.uint64)
    y[-5:] = np.uint64(0xFFFFFFFFFFFFFFFF)
    expect = np.sort(np.concatenate([x, y]), kind="stable")
    out = dev_empty(x.size + y.size, np.uint64)
    pa.merge_u64(to_dev(x), to_dev(y), out)
    np.testing.assert_array_equal(to_host(out, np.uint64), expect)


def test_select_band_overflow_falls_back(pa):
    """The first read compacts the keys of a band of top-digit bins predicted
    from strided sample runs (256 runs of 128 keys at positions
    run * (n - 128) / 255) when the sample puts at most half of the keys in
    it.  Here the sampled runs hold 45% band keys and the rest of the array
    70%, so the band is taken and overflows the n/2-key buffer: the select
    must fall back and still return sorted[k]."""
    n = 4_000_000
    rng = np.random.default_rng(77)
    sampled = np.zeros(n, dtype=bool)
    for run in range(256):
        b0 = run * (n - 128) // 255
        sampled[b0:b0 + 128] = True
    band = np.uint32(5 << 24)
    other = np.uint32(200 << 24)
    low = rng.integers(0, 1 << 24, n, dtype=np.uint32)
    p_band = np.where(sampled, 0.45, 0.70)
    a = np.where(rng.random(n) < p_band, band | low, other | low).astype(np.uint32)
    srt = np.sort(a)
    assert (srt >> 24 == 5).sum() > n // 2  # the band really overflows the buffer
    d = to_dev(a)
    for k in (n * 3 // 10, n // 2, n - 1):
        res = np.zeros(1, dtype=np.uint32)
        pa.select_u32(d, k, res)
        assert int(res[0]) == int(srt[k]), k
