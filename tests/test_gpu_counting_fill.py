"""GPU parity of the byte counting sort, parallel fill and the copy helper.

Counting sort mirrors counting_sort_parallel (CountingSortParallel.h:29-90)
and its benchmark's check against std::sort
(CountingSortParallelBenchmark.cpp:70-92); the oracle restates the histogram +
fill and is pinned to np.sort.  Fill mirrors parallel_fill
(FillParallel.h:22-66): every element of [l, r) equals the value, nothing
outside changes.  Sizes cover the 16-byte vector body, unaligned heads and
tails, and ranges shorter than one vector.
"""
import numpy as np
import pytest

from gpu_util import dev_empty, to_dev, to_host

pytestmark = pytest.mark.gpu


def _bytes(kind, n, seed=61):
    import parallelalgorithms_amd as P
    a = (P.gen_u32(seed + n, n) & np.uint32(0xFF)).astype(np.uint8)
    if kind == "constant":
        a[:] = 0x5A
    elif kind == "two":  # most bins empty, two long runs
        a = np.where(a & 1, 3, 250).astype(np.uint8)
    elif kind == "top":  # only bin 255 and bin 0
        a = np.where(a < 8, 0, 255).astype(np.uint8)
    return a


@pytest.mark.parametrize("kind", ["uniform", "constant", "two", "top"])
@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 65_537, 10_000_019])
def test_counting_sort_matches_oracle(pa, orc, kind, n):
    a = _bytes(kind, n)
    expect = orc.counting_sort_u8(a)
    d = to_dev(a)
    pa.counting_sort_u8(d)
    np.testing.assert_array_equal(to_host(d, np.uint8), expect)


def test_counting_sort_unaligned_and_host(pa, orc):
    n = 1_000_003
    a = _bytes("uniform", n + 7)
    d = to_dev(a)
    view = d[3:3 + n]  # 3 bytes past a 16-byte boundary
    pa.counting_sort_u8(view)
    got = to_host(d, np.uint8)
    np.testing.assert_array_equal(got[3:3 + n], orc.counting_sort_u8(a[3:3 + n]))
    np.testing.assert_array_equal(got[:3], a[:3])
    np.testing.assert_array_equal(got[3 + n:], a[3 + n:])
    h = a.copy()
    pa.counting_sort_u8(h[5:5 + n])  # host memory, odd offset
    np.testing.assert_array_equal(h[5:5 + n], np.sort(a[5:5 + n]))
    np.testing.assert_array_equal(h[:5], a[:5])


@pytest.mark.parametrize("dtype,value", [(np.uint8, 0xA5), (np.uint16, 0xBEEF), (np.uint32, 0xDEADBEEF),
                                         (np.uint64, 0x0123456789ABCDEF), (np.float32, -1.5), (np.float64, 2.0 ** -30)])
@pytest.mark.parametrize("lo,hi", [(0, 1), (1, 2), (3, 40), (0, 1_000_000), (5, 3_000_001)])
def test_fill_range(pa, dtype, value, lo, hi):
    import parallelalgorithms_amd as P
    n = hi + 9
    base = P.gen_u64(71, n).view(np.uint8)[: n * np.dtype(dtype).itemsize].view(dtype).copy()
    if dtype in (np.float32, np.float64):
        base = np.nan_to_num(base)
    t = to_dev(base) if np.dtype(dtype) in (np.dtype(np.uint8), np.dtype(np.uint32), np.dtype(np.uint64),
                                            np.dtype(np.float32), np.dtype(np.float64)) else None
    v = np.array(value, dtype=dtype)
    if t is None:  # uint16: fill a host array (staged through the device)
        h = base.copy()
        pa.fill(h[lo:hi], v)
        got = h
    else:
        pa.fill(t[lo:hi], v)
        got = to_host(t, dtype)
    expect = base.copy()
    expect[lo:hi] = v
    np.testing.assert_array_equal(got.view(np.uint8), expect.view(np.uint8))


def test_fill_rejects_misaligned(pa):
    import parallelalgorithms_amd as P
    d = dev_empty(64, np.uint8)
    with pytest.raises(P.PAError):
        P.call("pa_fill", d.data_ptr() + 1, 4, np.array([1], np.uint32).ctypes.data, 4, None)


def test_copy_between_memories(pa):
    import parallelalgorithms_amd as P
    a = P.gen_u32(81, 100_001)
    d = dev_empty(a.size, np.uint32)
    P.call("pa_copy", d.data_ptr(), a.ctypes.data, a.nbytes, None)
    e = dev_empty(a.size, np.uint32)
    P.call("pa_copy", e.data_ptr(), d.data_ptr(), a.nbytes, None)
    h = np.zeros_like(a)
    P.call("pa_copy", h.ctypes.data, e.data_ptr(), a.nbytes, None)
    np.testing.assert_array_equal(h, a)
