"""Sorts past the 30-bit count of a look-back status word: keys-only and
key-value sorts, and a segment of a segmented sort.

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


@pytest.mark.parametrize("wide", [False, True])
def test_sort_kv_two_to_the_30_plus_is_stable(pa, wide):
    """Key-value sorts past 2^30 keys: chunks sorted stably and merged with
    ties taken from the earlier run.  Keys drawn from 1,000,003 values (about
    a thousand copies each) and payload = input index: the payload must be
    torch's stable argsort, which pins both the order and the stability."""
    import torch
    import parallelalgorithms_amd as P
    n = (1 << 30) + 4099
    if wide:
        k = torch.empty(n, dtype=torch.int64, device="cuda")
        P.call("pa_fill_random_u64", k.data_ptr(), n, 105, None)
        k = (k & 0x7FFFFFFFFFFFFFFF) % 1_000_003 * 0x100000001
        key_view = k
    else:
        r = _random_i32(n, 106).to(torch.int64) & 0xFFFFFFFF
        key_view = r % 1_000_003 * 4099
        k = key_view.to(torch.int32)
        del r
    v = torch.arange(n, dtype=torch.int32, device="cuda")
    ref = torch.sort(key_view, stable=True)
    print(f"[large] kv {'u64' if wide else 'u32'} n={n}", flush=True)
    (P.sort_kv_u64_u32 if wide else P.sort_kv_u32_u32)(k, v)
    got_k = k if wide else k.to(torch.int64) & 0xFFFFFFFF
    assert torch.equal(got_k, ref.values)
    assert torch.equal(v.to(torch.int64), ref.indices)


def test_segmented_with_a_segment_past_two_to_the_30(pa):
    """A segment of 2^30 + 13 keys between small, queue-class and onesweep-class
    segments: the large one is sorted like a plain array of that size."""
    import numpy as np
    import torch
    lens = [1000, (1 << 30) + 13, 70_000, 5, 300_000, 1]
    offs = np.zeros(len(lens) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    total = int(offs[-1])
    d = _random_i32(total, 107)
    src = d.to(torch.int64) & 0xFFFFFFFF
    print(f"[large] segmented with one segment of {lens[1]} keys", flush=True)
    pa.segmented_sort_u32(d, None, offs)
    got = d.to(torch.int64) & 0xFFFFFFFF
    del d
    for i in range(len(lens)):
        a, b = int(offs[i]), int(offs[i + 1])
        assert torch.equal(got[a:b], src[a:b].sort().values), i
