"""Sorts whose keys leave some digit passes nothing to do (a digit equal for
every key): those passes are skipped, except one that still runs when the
other passes would otherwise leave the data in the other buffer.  Each pattern
below makes a different set of passes trivial; the result must equal a plain
sort, and key-value sorts must stay stable.

Patterns over 8-bit digits of u32 keys (pass p sorts bits 8p..8p+7; with an
odd number of non-trivial passes one trivial pass still runs):
  lt8    a & 0xFF:                   passes 1-3 trivial, two of them skipped
  lt16   a & 0xFFFF:                 passes 2-3 trivial, both skipped
  lt24   a & 0xFFFFFF:               pass 3 trivial, still run
  mid    (a & 0xFF00FFFF) | AB << 16: pass 2 trivial, still run
  low    (a & ~0xFF) | 0x5A:         pass 0 trivial (the prefix-ready pass), still run
  two    a & 0x00FF00FF:             passes 1 and 3 trivial, both skipped (pass 2 reads
                                     the temporary array)
  const  one value:                  every pass trivial and skipped
"""
import numpy as np
import pytest

from gpu_util import dev_empty, to_dev, to_host

pytestmark = pytest.mark.gpu

PATTERNS = {
    "lt8": lambda a: a & np.uint32(0xFF),
    "lt16": lambda a: a & np.uint32(0xFFFF),
    "lt24": lambda a: a & np.uint32(0xFFFFFF),
    "mid": lambda a: (a & np.uint32(0xFF00FFFF)) | np.uint32(0xAB << 16),
    "low": lambda a: (a & np.uint32(0xFFFFFF00)) | np.uint32(0x5A),
    "two": lambda a: a & np.uint32(0x00FF00FF),
    "const": lambda a: np.full_like(a, 0x12345678),
}
# Sizes in the three onesweep size classes (512 x 16 tiles with 32- and
# 16-lane units, then 768 x 32 tiles above 2^24 keys).
SIZES = [70_001, 1_300_007, 20_000_003]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("pattern", list(PATTERNS))
def test_sort_u32_trivial_passes(pa, pattern, n):
    a = PATTERNS[pattern](pa.gen_u32(7000 + n % 101, n))
    print(f"[trivial] u32 {pattern} n={n}", flush=True)
    d = to_dev(a)
    pa.sort_u32(d, dev_empty(n, np.uint32))
    np.testing.assert_array_equal(to_host(d, np.uint32), np.sort(a))


def test_sort_u32_trivial_passes_odd_pass_count(pa):
    """7-bit digits: 5 passes for u32.  Keys below 2^14 make passes 2-4
    trivial, so two passes permute and one trivial pass runs as well (the
    data must end in the temporary array, which the host copies back, as with
    no trivial pass)."""
    P = pa
    shapes = P.sort_shapes(P.TUNE_U32)
    idx = next(i for i, s in enumerate(shapes) if s["digit_bits"] == 7)
    try:
        for c in range(4):
            P.sort_shape_set(P.TUNE_U32, c, idx)
        for n, mask in ((1_300_007, 0x3FFF), (1_300_007, 0x1FFFFF), (5_000_011, 0x3FFF)):
            a = P.gen_u32(n, n) & np.uint32(mask)
            print(f"[trivial] u32 rb7 n={n} mask={mask:#x}", flush=True)
            d = to_dev(a)
            P.sort_u32(d, dev_empty(n, np.uint32))
            np.testing.assert_array_equal(to_host(d, np.uint32), np.sort(a))
    finally:
        P.autotune_reset()


@pytest.mark.parametrize("pattern", ["lt16", "low", "lt24", "const"])
def test_sort_kv_u32_trivial_passes_stable(pa, pattern):
    n = 2_000_003
    k = PATTERNS[pattern](pa.gen_u32(31, n))
    v = np.arange(n, dtype=np.uint32)
    print(f"[trivial] kv u32 {pattern} n={n}", flush=True)
    dk, dv = to_dev(k), to_dev(v)
    pa.sort_kv_u32_u32(dk, dv, dev_empty(n, np.uint32), dev_empty(n, np.uint32))
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(to_host(dk, np.uint32), k[order])
    np.testing.assert_array_equal(to_host(dv, np.uint32), v[order])


@pytest.mark.parametrize("bits", [32, 40, 8])
def test_sort_u64_trivial_passes(pa, bits):
    """u64 keys below 2^bits: 8 - bits / 8 trivial passes (4: an even number
    of permuting passes; 3 and 7 trivial: odd, so one trivial pass runs)."""
    n = 1_500_007
    a = pa.gen_u64(77, n) & np.uint64((1 << bits) - 1)
    print(f"[trivial] u64 below 2^{bits} n={n}", flush=True)
    d = to_dev(a)
    pa.sort_u64(d, dev_empty(n, np.uint64))
    np.testing.assert_array_equal(to_host(d, np.uint64), np.sort(a))
