"""GPU parity of radix partition (PartitionRadix, RadixPartition.h:94-184).

The reference's contract, checked by its benchmark against nth_element
(RadixPartitionBenchmark.cpp:50-66), is the value at k plus the partition
property.  The GPU arranges the other keys as one stable three-way split
around v = the k-th smallest -- (keys < v | keys == v | keys > v), each group
in input order -- so its output is fixed exactly and compared bit for bit
with that closed form; v itself comes from the oracle's PartitionRadix, and
the oracle's own output pins the property the reference promises.
"""
import numpy as np
import pytest

from gpu_util import dev_empty, make_u32, to_dev, to_host

pytestmark = pytest.mark.gpu


def _expected(a, k, orc):
    v = orc.partition_radix_u32(a, [k])[k]
    return v, np.concatenate([a[a < v], a[a == v], a[a > v]])


@pytest.mark.parametrize("dist", ["uniform", "presorted", "constant", "reversed", "nearly", "few"])
@pytest.mark.parametrize("n", [1, 2, 7, 100, 4096, 4097, 1_000_003])
def test_partition_matches_stable_split(pa, orc, dist, n):
    a = make_u32(dist, n)
    for k in sorted({0, n // 100, n // 2, max(0, n - 21), n - 1}):
        v, expect = _expected(a, k, orc)
        d = to_dev(a)
        res = np.zeros(1, dtype=np.uint32)
        pa.partition_u32(d, k, res)
        out = to_host(d, np.uint32)
        assert int(res[0]) == v, (dist, n, k)
        np.testing.assert_array_equal(out, expect, err_msg=f"{dist} n={n} k={k}")
        assert out[k] == v


def test_partition_large_and_host_keys(pa, orc):
    n = 10_000_019
    a = make_u32("uniform", n, seed=31)
    k = n - 21  # the reference benchmark's rank
    v, expect = _expected(a, k, orc)
    d = to_dev(a)
    r = dev_empty(1, np.uint32)
    pa.partition_u32(d, k, r)
    assert int(to_host(r, np.uint32)[0]) == v
    np.testing.assert_array_equal(to_host(d, np.uint32), expect)
    h = a.copy()
    res = np.zeros(1, dtype=np.uint32)
    pa.partition_u32(h, k, res)  # host keys: synchronous, in place
    assert int(res[0]) == v
    np.testing.assert_array_equal(h, expect)


def test_partition_repeated_calls_reuse_scratch(pa, orc):
    # A larger call, then a smaller one on the same stream: nothing from the
    # first call's scratch leaks into the second.
    for n, k in ((300_001, 150_000), (5_003, 17)):
        a = make_u32("few", n, seed=n)
        v, expect = _expected(a, k, orc)
        d = to_dev(a)
        res = np.zeros(1, dtype=np.uint32)
        pa.partition_u32(d, k, res)
        assert int(res[0]) == v
        np.testing.assert_array_equal(to_host(d, np.uint32), expect)


def test_partition_multi_ranks(pa, orc):
    n = 1_000_003
    a = make_u32("uniform", n, seed=32)
    ks = [n // 100, n // 2, n - 21]
    d = to_dev(a)
    pa.partition_multi_u32(d, ks)
    out = to_host(d, np.uint32)
    ref = orc.partition_radix_u32(a, ks)
    for k in ks:
        assert out[k] == ref[k]
        assert out[:k].max() <= out[k] <= out[k + 1:].min(initial=0xFFFFFFFF)
    np.testing.assert_array_equal(np.sort(out), np.sort(a))


def test_partition_rejects_bad_ranks(pa):
    import parallelalgorithms_amd as P
    d = to_dev(np.arange(10, dtype=np.uint32))
    res = np.zeros(1, dtype=np.uint32)
    with pytest.raises(P.PAError):
        pa.partition_u32(d, 10, res)
    with pytest.raises(P.PAError):
        pa.partition_multi_u32(d, [5, 3])
    with pytest.raises(P.PAError):
        pa.partition_multi_u32(d, [3, 10])
    np.testing.assert_array_equal(to_host(d, np.uint32), np.arange(10, dtype=np.uint32))
