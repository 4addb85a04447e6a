"""GPU parity of the histogram entry points against the oracle.

Reference: HistogramParallel.h (HistogramByteComponentsParallel :249-329,
HistogramOneByteComponentParallel :535-576 and _2 :466-533,
HistogramByteComponentsQCPar :444-464, byte-array histograms :16-247) and
ComputeStartOfBinsPar (RadixSortLsdParallel.h:108-165).  Counts are exact
integers: equality is bit-exact.
"""
import numpy as np
import pytest

from gpu_util import make_u32, to_dev

pytestmark = pytest.mark.gpu

SIZES = [1, 3, 100, 4097, 1_000_003]
DISTS = ["uniform", "presorted", "constant", "few"]


def _dev_u64(n):
    import torch
    return torch.zeros(n, dtype=torch.int64, device="cuda")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("dist", DISTS)
@pytest.mark.parametrize("n", SIZES)
def test_hist_digits(pa, orc, dist, n):
    a = make_u32(dist, n)
    expect = orc.hist_byte_components_u32(a, 0, n - 1)
    out = _dev_u64(1024)
    pa.hist_digits_u32(to_dev(a), out)
    got = _host(out).reshape(4, 256)
    np.testing.assert_array_equal(got, expect[:4])
    assert not expect[4:].any()


@pytest.mark.parametrize("shift,bits", [(24, 8), (0, 8), (8, 8), (16, 8), (20, 12), (3, 1), (28, 4)])
@pytest.mark.parametrize("dist", DISTS)
def test_hist_one_digit(pa, orc, dist, shift, bits):
    n = 1_000_003
    a = make_u32(dist, n)
    expect = orc.hist_one_byte_u32(a, 0, n - 1, shift, radix=1 << bits)
    out = _dev_u64(1 << bits)
    pa.hist_u32(to_dev(a), shift, bits, out)
    np.testing.assert_array_equal(_host(out), expect)


@pytest.mark.parametrize("shift,bits", [(56, 8), (0, 8), (32, 8), (28, 12), (63, 1)])
@pytest.mark.parametrize("n", [1, 3, 4097, 1_000_003])
def test_hist_u64(pa, shift, bits, n):
    """u64 one-digit histogram (an extension: the reference's is 32-bit);
    the counts are fixed by np.bincount of the digit."""
    import parallelalgorithms_amd as P
    a = P.gen_u64(91 + n, n)
    a[: n // 3] = a[0]  # a constant run for the wave-uniform path
    expect = np.bincount(((a >> np.uint64(shift)) & np.uint64((1 << bits) - 1)).astype(np.int64),
                         minlength=1 << bits).astype(np.uint64)
    out = _dev_u64(1 << bits)
    pa.hist_u64(to_dev(a), shift, bits, out)
    np.testing.assert_array_equal(_host(out), expect)
    host_out = np.zeros(1 << bits, dtype=np.uint64)
    pa.hist_u64(a, shift, bits, host_out)
    np.testing.assert_array_equal(host_out, expect)


def test_hist_one_digit_host_pointers(pa, orc):
    n = 77_777
    a = make_u32("uniform", n, seed=3)
    out = np.zeros(256, dtype=np.uint64)
    pa.hist_u32(a, 24, 8, out)  # host in, host out: staged and synchronous
    np.testing.assert_array_equal(out, orc.hist_one_byte_u32(a, 0, n - 1, 24))


@pytest.mark.parametrize("n", [1, 15, 16, 17, 1_000_003])
def test_hist_bytes(pa, orc, n):
    import parallelalgorithms_amd as P
    data = (P.gen_u32(9, n) & 0xFF).astype(np.uint8)
    data[: n // 3] = 0x5A  # a constant run
    expect = orc.hist_bytes(data, 0, n)
    out = _dev_u64(256)
    pa.hist_u8(to_dev(data), out)
    np.testing.assert_array_equal(_host(out), expect)


@pytest.mark.parametrize("quantum", [1, 7, 1000, 8192, 100_000, 512 * 1024])
@pytest.mark.parametrize("byte", [0, 3])
def test_hist_quanta(pa, orc, quantum, byte):
    n = 1_000_003 if quantum >= 1000 else 20_011  # the output has n/quantum rows
    a = make_u32("uniform", n, seed=4)
    nq = (n + quantum - 1) // quantum
    l, r = 5, n - 2  # inclusive bounds not aligned to quanta
    expect = orc.hist_qc_par_u32(a, l, r, quantum, nq, byte)
    out = _dev_u64(nq * 256)
    pa.hist_quanta_u32(to_dev(a), l, r + 1, quantum, nq, byte, out)
    np.testing.assert_array_equal(_host(out).reshape(nq, 256), expect)


@pytest.mark.parametrize("quantum", [3, 4096, 65536, 1_000_003])
@pytest.mark.parametrize("digit", [0, 1, 2, 3])
def test_start_of_bins(pa, orc, quantum, digit):
    n = 1_000_003 if quantum >= 1000 else 20_011
    a = make_u32("uniform", n, seed=8)
    nq = (n + quantum - 1) // quantum
    expect = orc.compute_start_of_bins_u32(a, quantum, nq, digit)
    out = _dev_u64(nq * 256)
    pa.start_of_bins_u32(to_dev(a), quantum, nq, digit, out)
    np.testing.assert_array_equal(_host(out).reshape(nq, 256), expect)


def test_hist_bytes_back_to_back_on_two_streams(pa, orc):
    """The byte histogram keeps its running counts in per-stream control words
    that its last block reads and clears: calls in a row, of every size
    (empty, one byte, unaligned, large), alternating between two streams,
    each give exactly their own counts."""
    import torch
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    sizes = [0, 1, 15, 17, 100_003, 5_000_000, 1, 0, 33_333_333, 64]
    for i, n in enumerate(sizes):
        data = (pa.gen_u32(50 + i, max(n, 1)) % 251).astype(np.uint8)[:n]
        d = to_dev(np.concatenate([np.zeros(3, np.uint8), data]))[3:]  # odd start address
        out = _dev_u64(256)
        st = streams[i % 2]
        with torch.cuda.stream(st):
            pa.hist_u8(d, out, stream=st.cuda_stream)
        st.synchronize()
        expect = orc.hist_bytes(data, 0, n) if n else np.zeros(256, np.uint64)
        np.testing.assert_array_equal(_host(out), expect, err_msg=f"call {i} n={n}")
