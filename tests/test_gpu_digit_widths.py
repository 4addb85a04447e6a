"""Every candidate shape of the sort -- tile, staging, ranking unit and digit
width of 4 to 8 bits -- gives the oracle's result, for u32, f32, i32 and u64
keys and for the segmented sort; the tuners pick valid shapes; tuning caches
round-trip and apply only to their own device.

Each plain shape is forced on every size class with pa_sort_shape_set, so the
sizes below run it in the onesweep path (the LDS single-workgroup path takes
u32 arrays of at most 16,384 keys whatever the shape).  The expected output is
the oracle's SortRadixInnerPar at the same digit width."""
import json
import os

import numpy as np
import pytest

from gpu_util import dev_empty, to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES_32 = [20_011, 70_001, 1_300_007, 5_000_011, 17_000_003]
SIZES_64 = [70_001, 3_000_017]


def _force(P, target, index):
    for c in range(4):
        P.sort_shape_set(target, c, index)


def _reset(P):
    P.autotune_reset()


@pytest.fixture(scope="module")
def inputs_32():
    import parallelalgorithms_amd as P
    out = {}
    for n in SIZES_32:
        a = P.gen_u32(4000 + n % 97, n)
        a[::11] = a[5]  # ties
        out[n] = a
    return out


def test_every_u32_shape_matches_oracle(pa, orc, inputs_32):
    P = pa
    shapes = P.sort_shapes(P.TUNE_U32)
    assert {s["digit_bits"] for s in shapes} == {4, 5, 6, 7, 8}
    assert any(s["passes"] % 2 for s in shapes)  # odd pass counts are covered
    expect = {n: orc.sort_radix_bits(a, 8) for n, a in inputs_32.items()}
    dev = {n: to_dev(a) for n, a in inputs_32.items()}
    try:
        for i, sh in enumerate(shapes):
            _force(P, P.TUNE_U32, i)
            for n, a in inputs_32.items():
                assert P.sort_shape_choice(P.TUNE_U32, n) == i
                print(f"[widths] u32 shape {i} {sh['block']}x{sh['items']} rb{sh['digit_bits']} n={n}", flush=True)
                d = dev[n].clone()
                P.sort_u32(d, dev_empty(n, np.uint32))
                got = to_host(d, np.uint32)
                e = orc.sort_radix_bits(a, sh["digit_bits"]) if n == SIZES_32[1] else expect[n]
                np.testing.assert_array_equal(got, e, err_msg=f"shape {i} ({sh}) n={n}")
    finally:
        _reset(P)


@pytest.mark.parametrize("rb", [4, 5, 6, 7, 8])
def test_float_and_signed_keys_at_every_width(pa, orc, rb):
    """The key transform is applied on the first pass and undone on the last,
    which with an odd pass count is the pass before the copy back."""
    import torch
    from parallelalgorithms_amd import workloads
    P = pa
    idx = next(i for i, s in enumerate(P.sort_shapes(P.TUNE_U32)) if s["digit_bits"] == rb and s["items"] == 16)
    n = 1_000_003
    d = workloads.f32_with_specials(n, seed=1236 + rb)  # device int32 bit patterns
    bits = to_host(d, np.uint32).copy()
    s = P.gen_u32(99 + rb, n).view(np.int32)
    try:
        _force(P, P.TUNE_U32, idx)
        P.sort_f32(d.view(torch.float32))
        np.testing.assert_array_equal(to_host(d, np.uint32), orc.sort_f32_bits(bits))
        di = to_dev(s)
        P.sort_i32(di)
        np.testing.assert_array_equal(to_host(di, np.int32), orc.sort_i32(s))
    finally:
        _reset(P)


def test_every_u64_shape_matches_oracle(pa, orc):
    P = pa
    shapes = P.sort_shapes(P.TUNE_U64)
    assert {s["digit_bits"] for s in shapes} >= {4, 5, 6, 7, 8}
    data = {n: P.gen_u64(600 + n % 89, n) for n in SIZES_64}
    expect = {n: orc.sort_u64(a) for n, a in data.items()}
    try:
        for i, sh in enumerate(shapes):
            _force(P, P.TUNE_U64, i)
            for n, a in data.items():
                print(f"[widths] u64 shape {i} rb{sh['digit_bits']} n={n}", flush=True)
                d = to_dev(a)
                P.sort_u64(d)
                np.testing.assert_array_equal(to_host(d, np.uint64), expect[n], err_msg=f"shape {i} ({sh}) n={n}")
    finally:
        _reset(P)


def _seg_batch(P, count=300, seed=77):
    from parallelalgorithms_amd import workloads
    lens = workloads.log_uniform(P.gen_u64(seed, count), 64, 400_000)
    lens[::37] = 70_000 + np.arange(len(lens[::37]), dtype=np.uint64) * 1001  # above the default queue class
    offs = np.zeros(count + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    keys = P.gen_u32(seed, int(offs[-1]))
    return lens.astype(np.uint64), offs, keys


def test_every_segmented_shape_matches_oracle(pa, orc):
    P = pa
    shapes = P.sort_shapes(P.TUNE_SEG_U32)
    assert {s["digit_bits"] for s in shapes} >= {4, 6, 8}
    lens, offs, keys = _seg_batch(P)
    expect = [orc.sort_radix_par_u32(keys[offs[i]:offs[i + 1]]) for i in range(len(lens))]
    try:
        for i, sh in enumerate(shapes):
            P.sort_shape_set(P.TUNE_SEG_U32, 0, i)
            assert P.sort_shape_choice(P.TUNE_SEG_U32) == i
            print(f"[widths] segmented shape {i} {sh}", flush=True)
            d = to_dev(keys)
            P.segmented_sort_u32(d, None, offs)
            got = to_host(d, np.uint32)
            for j in range(len(lens)):
                np.testing.assert_array_equal(got[offs[j]:offs[j + 1]], expect[j], err_msg=f"shape {i} segment {j}")
    finally:
        _reset(P)


def test_tuners_and_cache_round_trip(pa, tmp_path):
    P = pa
    key = P.tune_device_key()
    assert key.startswith("gfx950")
    _reset(P)
    builtin = [P.sort_shape_choice(P.TUNE_U32, n) for n in (1 << 16, 1 << 20, 1 << 24, 1 << 26)]
    try:
        P.autotune_sort(P.TUNE_U32, 4_000_000)
        P.autotune_sort(P.TUNE_U64, 2_000_000)
        lens, _, _ = _seg_batch(P, count=200, seed=5)
        P.autotune_segmented_u32(lens)
        tuned32 = [P.sort_shape_choice(P.TUNE_U32, n) for n in (1 << 16, 1 << 20, 1 << 24, 1 << 26)]
        tuned64 = [P.sort_shape_choice(P.TUNE_U64, n) for n in (1 << 16, 1 << 20, 1 << 24, 1 << 26)]
        tseg = P.sort_shape_choice(P.TUNE_SEG_U32)
        n32, n64, nseg = (len(P.sort_shapes(t)) for t in (P.TUNE_U32, P.TUNE_U64, P.TUNE_SEG_U32))
        assert all(0 <= t < n32 for t in tuned32) and all(0 <= t < n64 for t in tuned64) and 0 <= tseg < nseg
        print(f"[widths] tuned u32 {tuned32} u64 {tuned64} seg {tseg}", flush=True)
        path = tmp_path / "tune.json"
        P.tune_cache_save(str(path))
        doc = json.loads(path.read_text())
        assert key in doc["devices"] and len(doc["devices"][key]["u32"]) == 4
        _reset(P)
        assert [P.sort_shape_choice(P.TUNE_U32, n) for n in (1 << 16, 1 << 20, 1 << 24, 1 << 26)] == builtin
        P.tune_cache_load(str(path))
        assert [P.sort_shape_choice(P.TUNE_U32, n) for n in (1 << 16, 1 << 20, 1 << 24, 1 << 26)] == tuned32
        assert [P.sort_shape_choice(P.TUNE_U64, n) for n in (1 << 16, 1 << 20, 1 << 24, 1 << 26)] == tuned64
        assert P.sort_shape_choice(P.TUNE_SEG_U32) == tseg
        # Another device's entry is kept on save and ignored on load.
        p2 = tmp_path / "other.json"
        p2.write_text(json.dumps({"format": "pa-tune-1", "devices": {"gfx942|hip 1": {"u32": ["512x16/lb4/sp1/ul16/rb4/w8"] * 4}}}))
        _reset(P)
        P.tune_cache_load(str(p2))
        assert [P.sort_shape_choice(P.TUNE_U32, n) for n in (1 << 16, 1 << 20, 1 << 24, 1 << 26)] == builtin
        P.tune_cache_save(str(p2))
        doc2 = json.loads(p2.read_text())
        assert "gfx942|hip 1" in doc2["devices"] and key in doc2["devices"]
        # A sort under the loaded choices is still correct.
        P.tune_cache_load(str(path))
        a = P.gen_u32(31, 3_000_001)
        d = to_dev(a)
        P.sort_u32(d)
        np.testing.assert_array_equal(to_host(d, np.uint32), np.sort(a))
    finally:
        _reset(P)
        os.environ.pop("PA_TUNE_CACHE", None)
