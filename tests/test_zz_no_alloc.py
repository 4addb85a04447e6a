"""pa_sort_u32_noalloc, the allocation-free form of pa_sort_u32 (scratch in a
static device array): called through the C ABI, its output must equal
pa_sort_u32's bit for bit, on consecutive calls of decreasing size (the
second call finds the first call's contents in the shared array)."""
import ctypes

import numpy as np
import pytest

from gpu_util import dev_empty, to_dev

pytestmark = pytest.mark.gpu

PA_ERR_ARG, PA_ERR_TOO_LARGE = 1, 4


def _noalloc(lib, keys, tmp, n, stream):
    return lib.pa_sort_u32_noalloc(ctypes.c_void_p(keys.data_ptr()), ctypes.c_void_p(tmp.data_ptr()), n, stream)


def test_noalloc_equals_sort_on_consecutive_calls(pa):
    import torch
    lib = pa.load()
    stream = torch.cuda.current_stream().cuda_stream
    # 100M (the headline size, 24,576-key tiles), then smaller: 1M + 3
    # (8,192-key tiles), then 12,345 (one LDS workgroup).
    for n, seed in ((100_000_000, 71), (1_000_003, 72), (12_345, 73)):
        print(f"[noalloc] n={n} seed={seed}", flush=True)
        a = pa.gen_u32(seed, n)
        ref = to_dev(a)
        pa.sort_u32(ref, dev_empty(n, np.uint32))
        got = to_dev(a)
        tmp = dev_empty(n, np.uint32)
        assert _noalloc(lib, got, tmp, n, stream) == 0, lib.pa_last_error_string()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), n
        del ref, got, tmp
        torch.cuda.empty_cache()


def test_noalloc_rejects_what_it_cannot_do(pa):
    import torch
    lib = pa.load()
    stream = torch.cuda.current_stream().cuda_stream
    keys = dev_empty(16, np.uint32)
    # Sizes past the static array's range are refused before any device work.
    assert _noalloc(lib, keys, keys, (1 << 27) + 1, stream) == PA_ERR_TOO_LARGE
    # Host memory would need staging buffers: refused.
    h = np.zeros(1000, dtype=np.uint32)
    assert lib.pa_sort_u32_noalloc(h.ctypes.data, h.ctypes.data, h.size, stream) == PA_ERR_ARG
    assert lib.pa_sort_u32_noalloc(None, None, 10, stream) == PA_ERR_ARG


def test_noalloc_sort_in_a_captured_graph(pa):
    """The entry point is for code under HIP graph capture: a 10M-key sort is
    captured once (kernels only; the sort enqueues no memset or allocation)
    and replayed three times, each time on fresh keys written eagerly into
    the captured buffer, with the program's own work and synchronisations
    between replays; every replay must equal torch.sort of its keys."""
    import torch
    lib = pa.load()
    n = 10_000_000
    keys = dev_empty(n, np.uint32)
    tmp = dev_empty(n, np.uint32)
    # One eager call first: looks up the static array's address and loads the
    # kernels outside the capture.
    keys.copy_(to_dev(pa.gen_u32(80, n)))
    assert _noalloc(lib, keys, tmp, n, torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        rc = _noalloc(lib, keys, tmp, n, torch.cuda.current_stream().cuda_stream)
    assert rc == 0, lib.pa_last_error_string()
    for seed in (81, 82, 83):
        print(f"[noalloc graph] replay seed={seed}", flush=True)
        fresh = to_dev(pa.gen_u32(seed, n))
        keys.copy_(fresh)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        ref = torch.sort(fresh.to(torch.int64) & 0xFFFFFFFF).values.to(torch.int32)
        assert torch.equal(keys, ref), seed
