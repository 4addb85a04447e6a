"""Helpers shared by the GPU tests: seeded inputs and host<->device transfer.

Unsigned arrays travel as same-width signed torch tensors (bit-identical
views), so no value conversion happens on the way.
"""
import numpy as np

_VIEW = {np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64, np.dtype(np.float32): np.float32,
         np.dtype(np.uint8): np.uint8, np.dtype(np.int32): np.int32, np.dtype(np.int64): np.int64,
         np.dtype(np.float64): np.float64}


def to_dev(a):
    import torch
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(_VIEW[a.dtype]).copy()).to("cuda")


def to_host(t, dtype):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)


def dev_empty(n, dtype):
    import torch
    tmap = {np.dtype(np.uint32): torch.int32, np.dtype(np.uint64): torch.int64, np.dtype(np.float32): torch.float32,
            np.dtype(np.uint8): torch.uint8, np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
            np.dtype(np.float64): torch.float64}
    return torch.empty(n, dtype=tmap[np.dtype(dtype)], device="cuda")


def make_u32(dist, n, seed=1234):
    """Seeded u32 inputs: the reference's generator (mt19937_64, truncated),
    presorted 0..n-1, constant 0x5A5A5A5A, reversed, and few distinct values."""
    import parallelalgorithms_amd as pa
    if dist == "uniform":
        return pa.gen_u32(seed, n)
    if dist == "presorted":
        return np.arange(n, dtype=np.uint32)
    if dist == "constant":
        return np.full(n, 0x5A5A5A5A, dtype=np.uint32)
    if dist == "reversed":
        return np.arange(n, 0, -1, dtype=np.uint64).astype(np.uint32)
    if dist == "nearly":  # ParallelAlgorithms.cpp:151-162: a[i]=i, every 100th = 0x1010
        a = np.arange(n, dtype=np.uint32)
        a[::100] = 0x1010
        return a
    if dist == "few":
        return (pa.gen_u32(seed, n) % 7).astype(np.uint32) * np.uint32(0x01010101)
    raise ValueError(dist)
