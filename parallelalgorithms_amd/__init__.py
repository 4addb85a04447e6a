"""ctypes binding of libpa_hip.so, the gfx950 radix-sort library.

The product is the C ABI declared in include/pa_hip.h (and the drop-in C++
headers above it).  This module binds that ABI for Python callers: tests and
bench.py pass device tensors (torch, ROCm) or host numpy arrays straight
through.  Nothing here computes on the CPU: if the shared library is missing,
every call raises.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libpa_hip.so")
HEADER = os.path.join(ROOT, "include", "pa_hip.h")

PA_OK = 0
STATUS = {0: "PA_OK", 1: "PA_ERR_ARG", 2: "PA_ERR_OOM", 3: "PA_ERR_HIP", 4: "PA_ERR_TOO_LARGE", 5: "PA_ERR_NO_DEVICE"}

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u = ctypes.c_uint
_i = ctypes.c_int
_u64 = ctypes.c_uint64

# name -> (restype, argtypes)
_SIGS = {
    "pa_status_string": (ctypes.c_char_p, [_i]),
    "pa_last_error_string": (ctypes.c_char_p, []),
    "pa_version": (_i, []),
    "pa_set_device": (_i, [_i]),
    "pa_synchronize": (_i, [_vp]),
    "pa_sort_u32": (_i, [_vp, _vp, _sz, _vp]),
    "pa_sort_u32_noalloc": (_i, [_vp, _vp, _sz, _vp]),
    "pa_sort_two_phase_u32": (_i, [_vp, _vp, _sz, _vp]),
    "pa_autotune_sort_u32": (_i, [_sz, _vp]),
    "pa_sort_variant_u32": (_i, [_sz]),
    "pa_autotune_reset": (_i, []),
    "pa_sort_shape_count": (_i, [_i]),
    "pa_sort_shape_info": (_i, [_i, _i, _vp]),
    "pa_sort_shape_choice": (_i, [_i, _sz]),
    "pa_sort_shape_set": (_i, [_i, _i, _i]),
    "pa_autotune_sort": (_i, [_i, _sz, _vp]),
    "pa_autotune_segmented_u32": (_i, [_vp, _sz, _vp]),
    "pa_tune_cache_save": (_i, [ctypes.c_char_p]),
    "pa_tune_cache_load": (_i, [ctypes.c_char_p]),
    "pa_tune_device_key": (ctypes.c_char_p, []),
    "pa_merge_sort_u32": (_i, [_vp, _vp, _sz, _sz, _vp]),
    "pa_merge_sort_u64": (_i, [_vp, _vp, _sz, _sz, _vp]),
    "pa_sort_u64": (_i, [_vp, _vp, _sz, _vp]),
    "pa_sort_f32": (_i, [_vp, _vp, _sz, _vp]),
    "pa_sort_f64": (_i, [_vp, _vp, _sz, _vp]),
    "pa_counting_sort_u8": (_i, [_vp, _sz, _vp]),
    "pa_copy": (_i, [_vp, _vp, _sz, _vp]),
    "pa_fill": (_i, [_vp, _sz, _vp, _sz, _vp]),
    "pa_sort_i32": (_i, [_vp, _vp, _sz, _vp]),
    "pa_sort_i64": (_i, [_vp, _vp, _sz, _vp]),
    "pa_sort_kv_u32_u32": (_i, [_vp, _vp, _vp, _vp, _sz, _vp]),
    "pa_sort_kv_u64_u32": (_i, [_vp, _vp, _vp, _vp, _sz, _vp]),
    "pa_hist_digits_u32": (_i, [_vp, _sz, _vp, _vp]),
    "pa_hist_u32": (_i, [_vp, _sz, _u, _u, _vp, _vp]),
    "pa_hist_u64": (_i, [_vp, _sz, _u, _u, _vp, _vp]),
    "pa_hist_quanta_u32": (_i, [_vp, _sz, _sz, _sz, _sz, _u, _vp, _vp]),
    "pa_start_of_bins_u32": (_i, [_vp, _sz, _sz, _sz, _u, _vp, _vp]),
    "pa_hist_u8": (_i, [_vp, _sz, _vp, _vp]),
    "pa_select_u32": (_i, [_vp, _sz, _sz, _vp, _vp]),
    "pa_select_u64": (_i, [_vp, _sz, _sz, _vp, _vp]),
    "pa_partition_u32": (_i, [_vp, _sz, _sz, _vp, _vp]),
    "pa_partition_multi_u32": (_i, [_vp, _sz, _vp, _sz, _vp]),
    "pa_select_radix_u32": (_i, [_vp, _sz, _sz, _vp, _vp]),
    "pa_merge_u32": (_i, [_vp, _sz, _vp, _sz, _vp, _vp]),
    "pa_merge_u64": (_i, [_vp, _sz, _vp, _sz, _vp, _vp]),
    "pa_merge_inplace_u32": (_i, [_vp, _sz, _sz, _vp]),
    "pa_merge_inplace_u64": (_i, [_vp, _sz, _sz, _vp]),
    "pa_segmented_sort_u32": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "pa_segmented_sort_ptrs_u32": (_i, [_vp, _vp, _sz, _vp]),
    "pa_segmented_plan_create_u32": (_vp, [_vp, _vp, _sz]),
    "pa_segmented_plan_run": (_i, [_vp, _vp]),
    "pa_segmented_plan_destroy": (None, [_vp]),
    "pa_copy_u32": (_i, [_vp, _vp, _sz, _vp]),
    "pa_server_create": (_vp, [_sz, _i, _sz]),
    "pa_server_submit": (_i, [_vp, _vp, _vp]),
    "pa_server_wait": (_i, [_vp, _u64, _vp]),
    "pa_server_drain": (_i, [_vp]),
    "pa_server_get_stats": (_i, [_vp, _vp]),
    "pa_server_destroy": (None, [_vp]),
    "pa_job_scratch_bytes": (_i, [_vp, _vp]),
    "pa_thread_stream": (_vp, []),
    "pa_is_device_pointer": (_i, [_vp]),
    "pa_profile_enable": (_i, [_i]),
    "pa_profile_only": (_i, [ctypes.c_char_p]),
    "pa_profile_query": (_i, [ctypes.c_char_p, _vp, _vp]),
    "pa_profile_reset": (_i, []),
    "pa_gen_mt19937_u32": (None, [_u64, _vp, _sz]),
    "pa_gen_mt19937_u64": (None, [_u64, _vp, _sz]),
    "pa_fill_random_u32": (_i, [_vp, _sz, _u64, _vp]),
    "pa_fill_random_u64": (_i, [_vp, _sz, _u64, _vp]),
}

_lib = None


class PAError(RuntimeError):
    def __init__(self, name, status, detail):
        super().__init__(f"{name} -> {STATUS.get(status, status)}: {detail}")
        self.status = status


def header_symbols(path=HEADER):
    """Names of the functions declared in include/pa_hip.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pa_[a-z0-9_]+)\s*\(", text)))


def load():
    """Loads libpa_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python __graft_entry__.py` "
                           "or parallelalgorithms_amd/build.py")
    try:  # share torch's HIP runtime when torch is present (same soname)
        import torch  # noqa: F401
    except Exception:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(x):
    """Raw address of a torch tensor, numpy array, int or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(f"cannot take the address of {type(x)}")


def _stream(stream):
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return torch.cuda.current_stream().cuda_stream
        except Exception:
            pass
        return None
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return stream


def call(name, *args):
    lib = load()
    st = getattr(lib, name)(*args)
    if st != PA_OK:
        raise PAError(name, st, lib.pa_last_error_string().decode())
    return st


def numel(x):
    return x.numel() if hasattr(x, "numel") else x.size


def _check(x, itemsize, name, need=None):
    """Raises ValueError unless x (torch tensor or numpy array; None passes)
    has elements of itemsize bytes, is contiguous and holds >= need elements.
    The library takes raw pointers and counts: a wrong element size, a strided
    view or a short buffer would make its kernels read or write the wrong
    memory, so they are refused here before any call."""
    if x is None:
        return
    if hasattr(x, "element_size"):
        size, contig = x.element_size(), x.is_contiguous()
    elif hasattr(x, "itemsize"):
        size, contig = x.itemsize, bool(x.flags["C_CONTIGUOUS"])
    else:
        raise TypeError(f"{name}: expected a torch tensor or a numpy array, got {type(x)}")
    if size != itemsize:
        raise ValueError(f"{name}: elements must be {itemsize} bytes, got {size}")
    if not contig:
        raise ValueError(f"{name}: the array must be contiguous")
    if need is not None and numel(x) < need:
        raise ValueError(f"{name}: needs at least {need} elements, got {numel(x)}")


def sort_u32(keys, tmp=None, stream=None):
    _check(keys, 4, "keys")
    _check(tmp, 4, "tmp", numel(keys))
    call("pa_sort_u32", _ptr(keys), _ptr(tmp), numel(keys), _stream(stream))


def sort_u32_noalloc(keys, tmp, stream=None):
    """pa_sort_u32_noalloc: keys and tmp on the device, no device allocation."""
    _check(keys, 4, "keys")
    _check(tmp, 4, "tmp", numel(keys))
    call("pa_sort_u32_noalloc", _ptr(keys), _ptr(tmp), numel(keys), _stream(stream))


def sort_two_phase_u32(a, b, stream=None):
    _check(a, 4, "a")
    _check(b, 4, "b", numel(a))
    call("pa_sort_two_phase_u32", _ptr(a), _ptr(b), numel(a), _stream(stream))


def autotune_sort_u32(max_n, stream=None):
    call("pa_autotune_sort_u32", max_n, _stream(stream))


def sort_variant_u32(n):
    return load().pa_sort_variant_u32(n)


TUNE_U32, TUNE_U64, TUNE_SEG_U32 = 0, 1, 2


class SortShape(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("items", ctypes.c_int), ("lookback", ctypes.c_int), ("split", ctypes.c_int),
                ("lanes_per_unit", ctypes.c_int), ("digit_bits", ctypes.c_int), ("passes", ctypes.c_int),
                ("queue_max_keys", ctypes.c_uint64)]


def sort_shapes(target):
    """Candidate shapes of a tuning target (pa_sort_shape_info), as dicts."""
    lib = load()
    out = []
    for i in range(lib.pa_sort_shape_count(target)):
        sh = SortShape()
        call("pa_sort_shape_info", target, i, ctypes.addressof(sh))
        out.append({name: getattr(sh, name) for name, _ in SortShape._fields_})
    return out


def sort_shape_choice(target, n=0):
    return load().pa_sort_shape_choice(target, n)


def sort_shape_set(target, size_class, index):
    call("pa_sort_shape_set", target, size_class, index)


def autotune_sort(target, max_n, stream=None):
    call("pa_autotune_sort", target, max_n, _stream(stream))


def autotune_segmented_u32(lens, stream=None):
    import numpy as np
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    call("pa_autotune_segmented_u32", lens.ctypes.data, lens.size, _stream(stream))


def tune_cache_save(path):
    call("pa_tune_cache_save", os.fsencode(path))


def tune_cache_load(path):
    call("pa_tune_cache_load", os.fsencode(path))


def tune_device_key():
    return load().pa_tune_device_key().decode()


def autotune_reset():
    call("pa_autotune_reset")


def sort_u64(keys, tmp=None, stream=None):
    _check(keys, 8, "keys")
    _check(tmp, 8, "tmp", numel(keys))
    call("pa_sort_u64", _ptr(keys), _ptr(tmp), numel(keys), _stream(stream))


def merge_sort_u32(keys, leaf_keys=0, tmp=None, stream=None):
    """Radix-sorted leaves merged upward (pa_merge_sort_u32)."""
    _check(keys, 4, "keys")
    _check(tmp, 4, "tmp", numel(keys))
    call("pa_merge_sort_u32", _ptr(keys), _ptr(tmp), numel(keys), leaf_keys, _stream(stream))


def merge_sort_u64(keys, leaf_keys=0, tmp=None, stream=None):
    _check(keys, 8, "keys")
    _check(tmp, 8, "tmp", numel(keys))
    call("pa_merge_sort_u64", _ptr(keys), _ptr(tmp), numel(keys), leaf_keys, _stream(stream))


def sort_f32(keys, tmp=None, stream=None):
    _check(keys, 4, "keys")
    _check(tmp, 4, "tmp", numel(keys))
    call("pa_sort_f32", _ptr(keys), _ptr(tmp), numel(keys), _stream(stream))


def counting_sort_u8(data, stream=None):
    _check(data, 1, "data")
    call("pa_counting_sort_u8", _ptr(data), numel(data), _stream(stream))


def fill(a, value, stream=None):
    """Fills every element of a (numpy array or torch tensor) with value, a
    numpy scalar of the element's size (its bytes are copied as they are)."""
    import numpy as np
    itemsize = a.element_size() if hasattr(a, "element_size") else a.itemsize
    v = np.ascontiguousarray(np.asarray(value).reshape(1))
    if v.itemsize != itemsize:
        raise ValueError("fill: value must have the element size of a")
    call("pa_fill", _ptr(a), itemsize, v.ctypes.data, numel(a), _stream(stream))


def sort_f64(keys, tmp=None, stream=None):
    _check(keys, 8, "keys")
    _check(tmp, 8, "tmp", numel(keys))
    call("pa_sort_f64", _ptr(keys), _ptr(tmp), numel(keys), _stream(stream))


def sort_i32(keys, tmp=None, stream=None):
    _check(keys, 4, "keys")
    _check(tmp, 4, "tmp", numel(keys))
    call("pa_sort_i32", _ptr(keys), _ptr(tmp), numel(keys), _stream(stream))


def sort_i64(keys, tmp=None, stream=None):
    _check(keys, 8, "keys")
    _check(tmp, 8, "tmp", numel(keys))
    call("pa_sort_i64", _ptr(keys), _ptr(tmp), numel(keys), _stream(stream))


def sort_kv_u32_u32(keys, vals, ktmp=None, vtmp=None, stream=None):
    _check(keys, 4, "keys")
    _check(vals, 4, "vals", numel(keys))
    _check(ktmp, 4, "ktmp", numel(keys))
    _check(vtmp, 4, "vtmp", numel(keys))
    call("pa_sort_kv_u32_u32", _ptr(keys), _ptr(vals), _ptr(ktmp), _ptr(vtmp), numel(keys), _stream(stream))


def sort_kv_u64_u32(keys, vals, ktmp=None, vtmp=None, stream=None):
    _check(keys, 8, "keys")
    _check(vals, 4, "vals", numel(keys))
    _check(ktmp, 8, "ktmp", numel(keys))
    _check(vtmp, 4, "vtmp", numel(keys))
    call("pa_sort_kv_u64_u32", _ptr(keys), _ptr(vals), _ptr(ktmp), _ptr(vtmp), numel(keys), _stream(stream))


def hist_digits_u32(keys, counts, stream=None):
    _check(keys, 4, "keys")
    _check(counts, 8, "counts", 4 * 256)
    call("pa_hist_digits_u32", _ptr(keys), numel(keys), _ptr(counts), _stream(stream))


def hist_u32(keys, shift, bits, counts, stream=None):
    _check(keys, 4, "keys")
    _check(counts, 8, "counts", 1 << min(max(int(bits), 0), 12))
    call("pa_hist_u32", _ptr(keys), numel(keys), shift, bits, _ptr(counts), _stream(stream))


def hist_u64(keys, shift, bits, counts, stream=None):
    _check(keys, 8, "keys")
    _check(counts, 8, "counts", 1 << min(max(int(bits), 0), 12))
    call("pa_hist_u64", _ptr(keys), numel(keys), shift, bits, _ptr(counts), _stream(stream))


def hist_quanta_u32(keys, first, last, quantum, nq, which_byte, counts, stream=None):
    _check(keys, 4, "keys", last)
    _check(counts, 8, "counts", nq * 256)
    call("pa_hist_quanta_u32", _ptr(keys), first, last, quantum, nq, which_byte, _ptr(counts), _stream(stream))


def start_of_bins_u32(keys, quantum, nq, digit, out, stream=None):
    _check(keys, 4, "keys")
    _check(out, 8, "out", nq * 256)
    call("pa_start_of_bins_u32", _ptr(keys), numel(keys), quantum, nq, digit, _ptr(out), _stream(stream))


def hist_u8(data, counts, stream=None):
    _check(data, 1, "data")
    _check(counts, 8, "counts", 256)
    call("pa_hist_u8", _ptr(data), numel(data), _ptr(counts), _stream(stream))


def select_u32(keys, k, result, stream=None):
    _check(keys, 4, "keys")
    _check(result, 4, "result", 1)
    call("pa_select_u32", _ptr(keys), numel(keys), k, _ptr(result), _stream(stream))


def select_u64(keys, k, result, stream=None):
    _check(keys, 8, "keys")
    _check(result, 8, "result", 1)
    call("pa_select_u64", _ptr(keys), numel(keys), k, _ptr(result), _stream(stream))


def partition_u32(keys, k, result, stream=None):
    _check(keys, 4, "keys")
    _check(result, 4, "result", 1)
    call("pa_partition_u32", _ptr(keys), numel(keys), k, _ptr(result), _stream(stream))


def partition_multi_u32(keys, ks, stream=None):
    import numpy as np
    _check(keys, 4, "keys")
    ks = np.ascontiguousarray(ks, dtype=np.uint64)
    call("pa_partition_multi_u32", _ptr(keys), numel(keys), ks.ctypes.data, ks.size, _stream(stream))


def select_radix_u32(keys, k, result, stream=None):
    _check(keys, 4, "keys")
    _check(result, 4, "result", 1)
    call("pa_select_radix_u32", _ptr(keys), numel(keys), k, _ptr(result), _stream(stream))


def merge_u32(a, b, out, stream=None):
    _check(a, 4, "a")
    _check(b, 4, "b")
    _check(out, 4, "out", numel(a) + numel(b))
    call("pa_merge_u32", _ptr(a), numel(a), _ptr(b), numel(b), _ptr(out), _stream(stream))


def merge_u64(a, b, out, stream=None):
    _check(a, 8, "a")
    _check(b, 8, "b")
    _check(out, 8, "out", numel(a) + numel(b))
    call("pa_merge_u64", _ptr(a), numel(a), _ptr(b), numel(b), _ptr(out), _stream(stream))


def merge_inplace_u32(t, n1, stream=None):
    _check(t, 4, "t", n1)
    call("pa_merge_inplace_u32", _ptr(t), n1, numel(t) - n1, _stream(stream))


def merge_inplace_u64(t, n1, stream=None):
    _check(t, 8, "t", n1)
    call("pa_merge_inplace_u64", _ptr(t), n1, numel(t) - n1, _stream(stream))


def segmented_sort_u32(keys, tmp, offsets, stream=None):
    """offsets: host numpy uint64 array of num_segments + 1 entries."""
    _check(keys, 4, "keys")
    _check(tmp, 4, "tmp", numel(keys))
    _check(offsets, 8, "offsets")
    if len(offsets) and int(offsets[-1]) > numel(keys):
        raise ValueError(f"offsets: the last offset {int(offsets[-1])} is past the {numel(keys)} keys")
    call("pa_segmented_sort_u32", _ptr(keys), _ptr(tmp), _ptr(offsets), len(offsets) - 1, _stream(stream))


def segmented_sort_ptrs_u32(tensors, stream=None):
    """Sorts each device array in `tensors` (a list) in one batch."""
    import numpy as np
    ptrs = np.array([_ptr(t) or 0 for t in tensors], dtype=np.uint64)
    lens = np.array([numel(t) for t in tensors], dtype=np.uint64)
    call("pa_segmented_sort_ptrs_u32", ptrs.ctypes.data, lens.ctypes.data, len(tensors), _stream(stream))


class SegmentedPlan:
    """A graph-captured segmented sort of fixed device arrays
    (pa_segmented_plan_*): each run() sorts their current contents."""

    def __init__(self, tensors):
        import numpy as np
        self.lib = load()
        self.keep = list(tensors)  # the arrays must outlive the plan
        ptrs = np.array([_ptr(t) or 0 for t in tensors], dtype=np.uint64)
        lens = np.array([numel(t) for t in tensors], dtype=np.uint64)
        self.h = self.lib.pa_segmented_plan_create_u32(ptrs.ctypes.data, lens.ctypes.data, len(tensors))
        if not self.h:
            raise PAError("pa_segmented_plan_create_u32", 1, self.lib.pa_last_error_string().decode())

    def run(self, stream=None):
        call("pa_segmented_plan_run", self.h, _stream(stream))

    def close(self):
        if self.h:
            self.lib.pa_segmented_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


OP_SORT, OP_SELECT, OP_HIST = 0, 1, 2
KEY_U32, KEY_U64 = 0, 1


class Job(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("key_type", ctypes.c_int), ("keys", ctypes.c_void_p), ("n", ctypes.c_size_t),
                ("k", ctypes.c_size_t), ("shift", ctypes.c_uint), ("bits", ctypes.c_uint), ("out", ctypes.c_void_p)]


class ServerStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in ("submitted", "completed", "failed", "launches",
                                                      "coalesced_batches", "coalesced_jobs",
                                                      "peak_bytes_in_flight", "pending", "admission_waits",
                                                      "scratch_releases", "oversize_units", "bytes_held")]


class Server:
    """The continuous-batching server (pa_server_*): jobs on device arrays."""

    def __init__(self, budget_bytes=64 << 30, streams=4, coalesce_max_keys=1 << 20):
        self.lib = load()
        self.h = self.lib.pa_server_create(budget_bytes, streams, coalesce_max_keys)
        if not self.h:
            raise PAError("pa_server_create", 3, self.lib.pa_last_error_string().decode())

    @staticmethod
    def job_scratch_bytes(op, n, key_type=KEY_U32, bits=8):
        """Device scratch the library allocates for one job run alone
        (pa_job_scratch_bytes): what the server's admission charges."""
        job = Job(op, key_type, None, n, 0, 0, bits, None)
        b = ctypes.c_size_t(0)
        call("pa_job_scratch_bytes", ctypes.addressof(job), ctypes.addressof(b))
        return b.value

    def submit(self, op, keys, key_type=KEY_U32, k=0, shift=0, bits=8, out=None):
        job = Job(op, key_type, _ptr(keys), numel(keys), k, shift, bits, _ptr(out))
        t = ctypes.c_uint64(0)
        call("pa_server_submit", self.h, ctypes.addressof(job), ctypes.addressof(t))
        return t.value

    def wait(self, ticket):
        lat = ctypes.c_double(0.0)
        st = self.lib.pa_server_wait(self.h, ticket, ctypes.addressof(lat))
        if st != PA_OK:
            raise PAError("pa_server_wait", st, self.lib.pa_last_error_string().decode())
        return lat.value

    def drain(self):
        call("pa_server_drain", self.h)

    def stats(self):
        s = ServerStats()
        call("pa_server_get_stats", self.h, ctypes.addressof(s))
        return {name: getattr(s, name) for name, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            self.lib.pa_server_destroy(self.h)
            self.h = None


def copy_u32(dst, src, stream=None):
    _check(src, 4, "src")
    _check(dst, 4, "dst", numel(src))
    call("pa_copy_u32", _ptr(dst), _ptr(src), numel(src), _stream(stream))


def synchronize(stream=None):
    call("pa_synchronize", _stream(stream))


def profile_enable(on=True):
    call("pa_profile_enable", 1 if on else 0)


def profile_only(prefix=None):
    call("pa_profile_only", prefix.encode() if prefix else None)


def profile_reset():
    call("pa_profile_reset")


def profile_query(prefix):
    """(total_ms, launches) for kernels whose name starts with prefix."""
    tot = ctypes.c_double(0.0)
    cnt = ctypes.c_uint64(0)
    call("pa_profile_query", prefix.encode(), ctypes.addressof(tot), ctypes.addressof(cnt))
    return tot.value, cnt.value


def fill_random_u32(dst, seed, stream=None):
    _check(dst, 4, "dst")
    call("pa_fill_random_u32", _ptr(dst), numel(dst), seed, _stream(stream))


def fill_random_u64(dst, seed, stream=None):
    _check(dst, 8, "dst")
    call("pa_fill_random_u64", _ptr(dst), numel(dst), seed, _stream(stream))


def gen_u32(seed, n):
    import numpy as np
    out = np.empty(n, dtype=np.uint32)
    load().pa_gen_mt19937_u32(seed, out.ctypes.data, n)
    return out


def gen_u64(seed, n):
    import numpy as np
    out = np.empty(n, dtype=np.uint64)
    load().pa_gen_mt19937_u64(seed, out.ctypes.data, n)
    return out
