"""The radix-leaf merge sort (pa_merge_sort_u32 / _u64): leaves sorted as
one segmented batch (u32) or one sort each (u64), merged upward by the
merge-path kernels -- the shape of the reference's
parallel_merge_sort_hybrid_radix (ParallelMergeSort.h:255-295).  Leaf sizes
cover every segmented class (LDS, queue, onesweep), leaves that do not divide
n, an odd number of runs in a round and one leaf for the whole array; the
expected output is the oracle's SortRadixPar."""
import numpy as np
import pytest

from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,leaf", [(1_000_003, 0), (1_000_003, 4096), (5_000_011, 70_001), (3_000_000, 3_000_000),
                                    (777, 100), (2, 1), (2_500_001, 300_000)])
def test_merge_sort_u32_matches_oracle(pa, orc, n, leaf):
    a = pa.gen_u32(900 + n % 1000, n)
    a[::13] = a[min(7, n - 1)]
    d = to_dev(a)
    print(f"[msort] u32 n={n} leaf={leaf}", flush=True)
    pa.merge_sort_u32(d, leaf)
    np.testing.assert_array_equal(to_host(d, np.uint32), orc.sort_radix_par_u32(a))


@pytest.mark.parametrize("n,leaf", [(2_000_003, 300_000), (1_000_000, 0), (99_999, 99_999)])
def test_merge_sort_u64_matches_oracle(pa, orc, n, leaf):
    a = pa.gen_u64(77 + n % 1000, n)
    d = to_dev(a)
    print(f"[msort] u64 n={n} leaf={leaf}", flush=True)
    pa.merge_sort_u64(d, leaf)
    np.testing.assert_array_equal(to_host(d, np.uint64), orc.sort_u64(a))


def test_merge_sort_host_memory_and_caller_tmp(pa, orc):
    a = pa.gen_u32(31, 600_001)
    h = a.copy()
    pa.merge_sort_u32(h, 50_000)  # host array: staged
    np.testing.assert_array_equal(h, orc.sort_radix_par_u32(a))
    import torch
    d = to_dev(a)
    t = torch.empty_like(d)
    pa.merge_sort_u32(d, 50_000, tmp=t)
    np.testing.assert_array_equal(to_host(d, np.uint32), orc.sort_radix_par_u32(a))
