"""GPU parity of the LSD radix sort entry points against the oracle.

Mirrors the reference's self-checking benchmark (RadixSortLsdBenchmark.cpp:78-114:
sort, then compare with std::stable_sort and exit on mismatch), with the
oracle's SortRadixPar restatement as the expected output.  Sorted output of
scalar keys is unique, so equality is bit-exact.
"""
import numpy as np
import pytest

from gpu_util import dev_empty, make_u32, to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 50, 99, 100, 101, 4095, 8191, 8192, 8193, 65537, 1_000_003]
DISTS = ["uniform", "presorted", "constant", "reversed", "nearly", "few"]


@pytest.mark.parametrize("dist", DISTS)
@pytest.mark.parametrize("n", SIZES)
def test_sort_u32_matches_oracle(pa, orc, dist, n):
    a = make_u32(dist, n)
    expect = orc.sort_radix_par_u32(a)
    d = to_dev(a)
    tmp = dev_empty(n, np.uint32)
    pa.sort_u32(d, tmp)
    got = to_host(d, np.uint32)
    np.testing.assert_array_equal(got, expect)


def test_sort_u32_config1_10M(pa, orc):
    """10,000,000 uniform u32 keys from mt19937_64(1234), 4 passes of 8 bits."""
    n = 10_000_000
    a = make_u32("uniform", n)
    expect = orc.sort_radix_par_u32(a)
    d = to_dev(a)
    pa.sort_u32(d, dev_empty(n, np.uint32))
    np.testing.assert_array_equal(to_host(d, np.uint32), expect)


def test_sort_u32_tmp_null_and_host_pointers(pa):
    n = 300_001
    a = make_u32("uniform", n, seed=7)
    expect = np.sort(a)
    d = to_dev(a)
    pa.sort_u32(d, None)  # library scratch
    np.testing.assert_array_equal(to_host(d, np.uint32), expect)
    h = a.copy()
    pa.sort_u32(h, None)  # host memory: staged, synchronous
    np.testing.assert_array_equal(h, expect)


@pytest.mark.parametrize("n", [(4 << 20) + 1, (16 << 20) + 5, 5 * (16 << 20) + 3])
def test_sort_host_memory_staging(pa, n):
    """Host arrays above 16 MB go through the pipelined pinned staging: one
    chunk, two chunks, and more chunks than pinned slots (slot reuse both ways)."""
    import parallelalgorithms_amd as P
    a = P.gen_u32(n, n)
    h = a.copy()
    pa.sort_u32(h)
    np.testing.assert_array_equal(h, np.sort(a))
    if n < (64 << 20):
        k = P.gen_u64(n + 1, n)
        v = np.arange(n, dtype=np.uint32)
        pa.sort_kv_u64_u32(k, v)  # keys and payload both staged
        order = np.argsort(P.gen_u64(n + 1, n), kind="stable")
        np.testing.assert_array_equal(v, order.astype(np.uint32))
        np.testing.assert_array_equal(k, np.sort(P.gen_u64(n + 1, n)))


def test_two_phase_result_in_both(pa):
    n = 123_457
    a = make_u32("uniform", n, seed=11)
    da, db = to_dev(a), dev_empty(n, np.uint32)
    pa.sort_two_phase_u32(da, db)
    np.testing.assert_array_equal(to_host(da, np.uint32), np.sort(a))
    np.testing.assert_array_equal(to_host(db, np.uint32), np.sort(a))


@pytest.mark.parametrize("n", [1, 100, 4097, 1_000_003])
def test_sort_u64_matches_oracle(pa, orc, n):
    import parallelalgorithms_amd as P
    a = P.gen_u64(1235, n)
    expect = orc.sort_u64(a)
    d = to_dev(a)
    pa.sort_u64(d, dev_empty(n, np.uint64))
    np.testing.assert_array_equal(to_host(d, np.uint64), expect)


@pytest.mark.parametrize("n", [1, 100, 5000, 1_000_003])
def test_sort_kv_u64_u32_stable(pa, orc, n):
    import parallelalgorithms_amd as P
    k = P.gen_u64(1235, n)
    k[::3] = k[0]  # many equal keys: stability is visible in the payload
    v = np.arange(n, dtype=np.uint32)
    ek, ev = orc.sort_kv_u64_u32(k, v)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(ev, order.astype(np.uint32))  # pins the oracle too
    dk, dv = to_dev(k), to_dev(v)
    pa.sort_kv_u64_u32(dk, dv, dev_empty(n, np.uint64), dev_empty(n, np.uint32))
    np.testing.assert_array_equal(to_host(dk, np.uint64), ek)
    np.testing.assert_array_equal(to_host(dv, np.uint32), ev)


@pytest.mark.parametrize("n", [1, 777, 1_000_003])
def test_sort_kv_u32_u32_stable(pa, orc, n):
    k = make_u32("few", n, seed=5)
    v = np.arange(n, dtype=np.uint32)
    ek, ev = orc.sort_kv_u32_u32(k, v)
    dk, dv = to_dev(k), to_dev(v)
    pa.sort_kv_u32_u32(dk, dv, dev_empty(n, np.uint32), dev_empty(n, np.uint32))
    np.testing.assert_array_equal(to_host(dk, np.uint32), ek)
    np.testing.assert_array_equal(to_host(dv, np.uint32), ev)


def _float_specials_input(n, seed=1236):
    import parallelalgorithms_amd as P
    r = P.gen_u64(seed, 2 * n + 16)
    u1 = (r[0:n] >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    u2 = (r[n:2 * n] >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    bits = z.astype(np.float32).view(np.uint32).copy()
    specials = np.array([0x7F800000, 0xFF800000, 0x00000000, 0x80000000, 0x7FC00000, 0xFFC00000, 0x7F800001],
                        dtype=np.uint32)
    pos = (r[:n] % np.uint64(100)) == np.uint64(0)
    bits[pos] = specials[(r[n:2 * n][pos] % np.uint64(len(specials))).astype(np.int64)]
    return bits


@pytest.mark.parametrize("n", [1, 1000, 1_000_003])
def test_sort_f32_order_transform(pa, orc, n):
    bits = _float_specials_input(n)
    expect = orc.sort_f32_bits(bits)
    # Closed form: stable order of the transformed bits.
    f = np.where(bits >> 31, ~bits, bits | np.uint32(0x80000000)).astype(np.uint32)
    np.testing.assert_array_equal(expect, bits[np.argsort(f, kind="stable")])
    d = to_dev(bits.view(np.float32))
    pa.sort_f32(d, dev_empty(n, np.float32))
    np.testing.assert_array_equal(to_host(d, np.float32).view(np.uint32), expect)


def _double_specials_input(n, seed=17):
    import parallelalgorithms_amd as P
    r = P.gen_u64(seed, 2 * n + 16)
    z = (r[:n] >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2e6 - 1e6
    bits = z.view(np.uint64).copy()
    specials = np.array([0x7FF0000000000000, 0xFFF0000000000000, 0, 1 << 63, 0x7FF8000000000000,
                         0xFFF8000000000000, 0x7FF0000000000001, 1, (1 << 63) | 1], dtype=np.uint64)
    pos = (r[:n] % np.uint64(50)) == np.uint64(0)
    bits[pos] = specials[(r[n:2 * n][pos] % np.uint64(len(specials))).astype(np.int64)]
    return bits


@pytest.mark.parametrize("n", [1, 99, 100, 1000, 1_000_003])
def test_sort_f64_order_transform(pa, orc, n):
    bits = _double_specials_input(n)
    expect = orc.sort_f64_bits(bits)
    d = to_dev(bits.view(np.float64))
    pa.sort_f64(d)
    np.testing.assert_array_equal(to_host(d, np.float64).view(np.uint64), expect)
    finite = np.isfinite(bits.view(np.float64))
    got = to_host(d, np.float64)
    np.testing.assert_array_equal(got[np.isfinite(got)], np.sort(bits.view(np.float64)[finite]))


@pytest.mark.parametrize("n", [1, 99, 100, 4097, 1_000_003])
def test_sort_signed(pa, orc, n):
    import parallelalgorithms_amd as P
    a32 = P.gen_u32(31 + n, n).view(np.int32)
    a64 = P.gen_u64(32 + n, n).view(np.int64)
    if n > 10:  # the extremes and zero
        a32[:3] = [np.iinfo(np.int32).min, 0, np.iinfo(np.int32).max]
        a64[:3] = [np.iinfo(np.int64).min, 0, np.iinfo(np.int64).max]
    for a, fn, ref, dt in ((a32, pa.sort_i32, orc.sort_i32, np.int32), (a64, pa.sort_i64, orc.sort_i64, np.int64)):
        expect = ref(a)
        np.testing.assert_array_equal(expect, np.sort(a))
        d = to_dev(a)
        fn(d, dev_empty(n, dt))
        np.testing.assert_array_equal(to_host(d, dt), expect)
    h = a64.copy()
    pa.sort_i64(h)  # host keys: staged through the device
    np.testing.assert_array_equal(h, np.sort(a64))


def test_profile_filter_times_only_the_named_kernels(pa):
    """pa_profile_only restricts event timing to kernels whose name starts with
    the prefix: one 4-pass sort records four k_onesweep launches and nothing
    else; without a filter the histogram step is timed too."""
    import parallelalgorithms_amd as P
    n = 3_000_017
    d = to_dev(make_u32("uniform", n, seed=5))
    P.profile_reset()
    P.profile_only("k_onesweep")
    P.profile_enable(True)
    try:
        pa.sort_u32(d)
    finally:
        P.profile_enable(False)
        P.profile_only(None)
    assert P.profile_query("k_onesweep")[1] == 4
    assert P.profile_query("k_hist")[1] == 0
    P.profile_reset()
    P.profile_enable(True)
    try:
        pa.sort_u32(d)
    finally:
        P.profile_enable(False)
    assert P.profile_query("k_onesweep")[1] == 4 and P.profile_query("k_hist")[1] >= 1
    P.profile_reset()


def test_autotune_then_sort(pa):
    """After the tuner records a pass shape per size class, sorts in every
    class still match np.sort; the choice is a valid candidate index."""
    import parallelalgorithms_amd as P
    P.autotune_sort_u32(1 << 20)
    try:
        for n in (1000, 70_001, 1 << 20, (1 << 20) + 3):
            v = P.sort_variant_u32(n)
            assert v >= 0  # an index into the candidate table
            a = P.gen_u32(7 + n, n)
            d = to_dev(a)
            pa.sort_u32(d)
            np.testing.assert_array_equal(to_host(d, np.uint32), np.sort(a))
    finally:
        P.autotune_reset()
