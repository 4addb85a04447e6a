"""Synthetic inputs of the batched configuration shared by bench.py and the
tests: 10,000 u32 arrays with lengths log-uniform in [64, 4,000,000] drawn
from mt19937_64(1239), 5% of them presorted (the decision drawn from the same
generator), keys from the counter-based device generator."""
import numpy as np


def unit_floats(u64):
    """[0, 1) doubles from the top 53 bits of mt19937_64 outputs."""
    return (u64 >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def log_uniform(u64, lo, hi):
    return np.floor(np.exp(np.log(lo) + unit_floats(u64) * (np.log(hi) - np.log(lo)))).astype(np.uint64)


def segmented_batch(count=10_000, seed=1239, lo=64, hi=4_000_000, presorted_frac=0.05):
    """(lengths, offsets, presorted mask) of the batched-sort configuration:
    offsets has count + 1 entries, offsets[i]..offsets[i+1] is array i."""
    from . import gen_u64
    r = gen_u64(seed, 2 * count)
    lens = log_uniform(r[:count], lo, hi)
    presorted = unit_floats(r[count:]) < presorted_frac
    offs = np.zeros(count + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    return lens, offs, presorted


def fill_segmented_batch(keys, offs, presorted, seed):
    """Writes the batch's keys into the device tensor `keys` (int32 view of
    u32): uniform keys from the device generator, the presorted arrays sorted
    in place by one segmented call."""
    from . import fill_random_u32, segmented_sort_ptrs_u32
    fill_random_u32(keys, seed)
    pre = [keys[int(offs[i]):int(offs[i + 1])] for i in np.nonzero(presorted)[0]]
    if pre:
        segmented_sort_ptrs_u32(pre)
