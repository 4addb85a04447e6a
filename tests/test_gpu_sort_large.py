"""Keys-only sorts past the 30-bit count of a look-back status word.

Arrays of 2^30 keys or more are sorted in equal chunks of at most 2^30 - 1
keys and the sorted runs merged pairwise (the reference's radix-leaf merge sort shape,
ParallelMergeSort.h:255-295).  At these sizes the host oracle would take
minutes, so each case is checked on the device against torch.sort of the same
keys (values only: for scalar keys the ascending sequence is unique, and the
inputs hold no NaN or signed zero, whose order torch leaves open), which fixes
the output bit for bit, and by sortedness.  The chunk and merge logic itself
is the same code at every size: 2^30 + 4099 keys make 2 chunks and one merge
round; 2^31 + 5 keys make 3 chunks, whose odd run is copied through the first
round.
"""
import pytest

pytestmark = pytest.mark.gpu

N = (1 << 30) + 4099


def _random_i32(n, seed):
    import torch
    import parallelalgorithms_amd as P
    d = torch.empty(n, dtype=torch.int32, device="cuda")
    P.call("pa_fill_random_u32", d.data_ptr(), n, seed, None)
    return d


def test_sort_u32_two_to_the_30_plus(pa):
    import torch
    d = _random_i32(N, 101)
    ref = (d.to(torch.int64) & 0xFFFFFFFF).sort().values
    print(f"[large] u32 n={N}", flush=True)
    pa.sort_u32(d)
    got = d.to(torch.int64) & 0xFFFFFFFF
    assert bool((got[1:] >= got[:-1]).all())
    assert torch.equal(got, ref)


def test_sort_f32_two_to_the_30_plus(pa):
    import torch
    r = _random_i32(N, 102).to(torch.int64)
    # finite, nonzero, both signs: sign and mantissa random, exponent 64..191
    bits = (r & 0x807FFFFF) | ((((r >> 8) & 127) + 64) << 23)
    del r
    d = bits.to(torch.int32).view(torch.float32)
    del bits
    ref = d.sort().values
    print(f"[large] f32 n={N}", flush=True)
    pa.sort_f32(d)
    assert torch.equal(d.view(torch.int32), ref.view(torch.int32))


def test_sort_u32_three_chunks(pa):
    # 2^31 + 5 keys: past what torch.sort and masked selection index, so the
    # check goes by the top 4 bits -- the output's keys of bucket b form one run, which must equal
    # torch.sort of the input's bucket-b keys (about 134M each).
    import torch
    n = (1 << 31) + 5
    d = _random_i32(n, 104)
    inp = d.to(torch.int64) & 0xFFFFFFFF
    print(f"[large] u32 n={n}", flush=True)
    pa.sort_u32(d)
    got = d.to(torch.int64) & 0xFFFFFFFF
    del d
    assert bool((got[1:] >= got[:-1]).all())
    start = 0
    h = n // 2  # masked selection stays below 2^31 elements per call
    for b in range(16):
        sel = torch.cat([part[(part >> 28) == b] for part in (inp[:h], inp[h:])])
        m = sel.numel()
        assert torch.equal(got[start:start + m], sel.sort().values), b
        start += m
    assert start == n


def test_sort_i64_two_to_the_30_plus(pa):
    import torch
    import parallelalgorithms_amd as P
    n = (1 << 30) + 7
    d = torch.empty(n, dtype=torch.int64, device="cuda")
    P.call("pa_fill_random_u64", d.data_ptr(), n, 103, None)
    ref = d.sort().values
    print(f"[large] i64 n={n}", flush=True)
    pa.sort_i64(d)
    assert torch.equal(d, ref)
