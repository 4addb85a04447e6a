"""The benchmark configurations at their full sizes, checked on the device.

At 100M-200M keys the host oracle takes too long per case, so each result is
checked against an independent device computation that fixes it bit for bit:
torch.sort of the same keys (values; for scalar keys the ascending sequence is
unique), torch's stable sort indices for the key-value payload, torch.bincount
for histograms, and the sorted array's k-th element for select.  The
histogram invariants of the presorted and constant inputs are the closed forms
the reference's configuration implies.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 100_000_000
SIGN32 = -(1 << 31)


def _keys(dist, n=N):
    import torch
    import parallelalgorithms_amd as P
    if dist == "uniform":
        return torch.from_numpy(P.gen_u32(1234, n).view(np.int32)).cuda()
    if dist == "presorted":
        return torch.arange(n, dtype=torch.int32, device="cuda")
    return torch.full((n,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")


def _u(t):  # int32 view of u32 keys -> exact int64 values
    import torch
    return t.to(torch.int64) & 0xFFFFFFFF


@pytest.mark.parametrize("dist", ["uniform", "presorted", "constant"])
def test_config2_sort_hist_select_100M(pa, dist):
    import torch
    d = _keys(dist)
    ref = _u(d).sort().values
    # 256-bin histogram of the top byte
    counts = torch.zeros(256, dtype=torch.int64, device="cuda")
    pa.hist_u32(d, 24, 8, counts)
    expect = torch.bincount((_u(d) >> 24), minlength=256)
    assert torch.equal(counts, expect)
    if dist == "presorted":
        c = counts.cpu().numpy()
        assert list(c[:5]) == [16_777_216] * 5 and c[5] == 16_113_920 and c[6:].sum() == 0
    if dist == "constant":
        assert int(counts[0x5A]) == N
    # radix select of k = n/2 and k = n/100
    for k in (N // 2, N // 100):
        r = torch.zeros(1, dtype=torch.int32, device="cuda")
        pa.select_u32(d, k, r)
        assert int(_u(r)[0]) == int(ref[k]), (dist, k)
    # the sort itself
    pa.sort_u32(d)
    assert torch.equal(_u(d), ref)


def test_config3_kv_u64_200M(pa):
    import torch
    import parallelalgorithms_amd as P
    n = 200_000_000
    k = torch.from_numpy(P.gen_u64(1235, n).view(np.int64)).cuda()
    # Unsigned order of int64 bit patterns = signed order with the sign flipped.
    flipped = k ^ (-(1 << 63))
    ref = flipped.sort(stable=True)
    v = torch.arange(n, dtype=torch.int32, device="cuda")
    pa.sort_kv_u64_u32(k, v)
    assert torch.equal(k ^ (-(1 << 63)), ref.values)
    assert torch.equal(v.to(torch.int64), ref.indices)  # stable: ties keep input order


def test_kv_u32_100M_stable(pa):
    """Key-value u32 sort at 100M keys (the large-array pass shape), payload =
    original index: keys equal torch.sort's, payload equals torch's stable
    sort indices."""
    import torch
    import parallelalgorithms_amd as P
    k = torch.from_numpy(P.gen_u32(1239, N).view(np.int32)).cuda()
    k &= 0x7FFFFF0F  # many ties, so stability is exercised
    ref = k.sort(stable=True)  # non-negative int32: signed order is the unsigned order
    v = torch.arange(N, dtype=torch.int32, device="cuda")
    pa.sort_kv_u32_u32(k, v)
    assert torch.equal(k, ref.values)
    assert torch.equal(v.to(torch.int64), ref.indices)


def test_config3_f32_specials_100M(pa):
    import torch
    from parallelalgorithms_amd import workloads
    bits = workloads.f32_with_specials(N, 1236)
    # The order: transformed bits f(u) = u ^ (sign ? ~0 : 0x80000000), sorted
    # as unsigned, then transformed back.
    b = _u(bits)
    f = torch.where(b >> 31 == 1, b ^ 0xFFFFFFFF, b | 0x80000000)
    fs = f.sort().values
    expect = torch.where(fs >> 31 == 1, fs ^ 0x80000000, fs ^ 0xFFFFFFFF)
    del b, f, fs
    assert torch.equal(workloads.sorted_f32_bits(bits), expect)  # the bench's check agrees
    d = bits.view(torch.float32)
    pa.sort_f32(d)
    assert torch.equal(_u(d.view(torch.int32)), expect)


def test_config3_merge_2x100M(pa):
    import torch
    import parallelalgorithms_amd as P
    a = _u(torch.from_numpy(P.gen_u32(1237, N).view(np.int32)).cuda()).sort().values.to(torch.int32)
    b = _u(torch.from_numpy(P.gen_u32(1238, N).view(np.int32)).cuda()).sort().values.to(torch.int32)
    out = torch.empty(2 * N, dtype=torch.int32, device="cuda")
    pa.merge_u32(a, b, out)
    ref = torch.cat([_u(a), _u(b)]).sort().values
    assert torch.equal(_u(out), ref)
