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


def f32_with_specials(n, seed=1236):
    """Device int32 bit patterns of the float configuration: normal(0, 1) by
    Box-Muller on mt19937_64(seed) (not std::normal_distribution, whose
    algorithm is implementation-defined), with about 0.8% of the keys (low 7
    bits of the first draw zero) replaced by one of +inf, -inf, +0, -0, a
    quiet NaN of each sign and a signalling NaN."""
    import torch
    from . import gen_u64
    r = torch.from_numpy(gen_u64(seed, 2 * n).view(np.int64)).cuda()
    u1 = ((r[:n] >> 11) & ((1 << 53) - 1)).to(torch.float64) / float(1 << 53)
    u2 = ((r[n:] >> 11) & ((1 << 53) - 1)).to(torch.float64) / float(1 << 53)
    z = torch.sqrt(-2.0 * torch.log1p(-u1)) * torch.cos(2.0 * np.pi * u2)
    bits = z.to(torch.float32).view(torch.int32).clone()
    del u1, u2, z
    specials = torch.tensor([0x7F800000, 0xFF800000 - (1 << 32), 0, -(1 << 31), 0x7FC00000, 0xFFC00000 - (1 << 32),
                             0x7F800001], dtype=torch.int32, device="cuda")
    pick = (r[:n] & 127) == 0
    bits[pick] = specials[(r[n:][pick] % 7).to(torch.int64)]
    return bits


def sorted_f32_bits(bits):
    """The sorted order of float bit patterns (int32 device tensor) under the
    library's total order, computed with torch: transform f(u) = u ^ (sign ?
    ~0 : 0x80000000), sort as unsigned, transform back.  Returns int64 values
    of the unsigned patterns."""
    import torch
    b = bits.to(torch.int64) & 0xFFFFFFFF
    f = torch.where(b >> 31 == 1, b ^ 0xFFFFFFFF, b | 0x80000000)
    fs = f.sort().values
    return torch.where(fs >> 31 == 1, fs ^ 0x80000000, fs ^ 0xFFFFFFFF)
