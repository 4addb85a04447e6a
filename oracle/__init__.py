"""ctypes binding of liboracle.so -- the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the timed CPU baseline.  The
product (parallelalgorithms_amd, libpa_hip.so) never imports or calls it.
All arrays are host numpy arrays.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PA_ORACLE_LIB selects another build of the same source (the sanitizer build).
LIB_PATH = os.environ.get("PA_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libpa_ref.so")

_vp, _sz, _u, _u64, _u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint64, ctypes.c_uint32

_SIGS = {
    "orc_set_threads": (None, [_u]),
    "orc_get_threads": (_u, []),
    "orc_insertion_sort_u32": (None, [_vp, _sz]),
    "orc_binary_search_u32": (_sz, [_u32, _vp, _sz, _sz]),
    "orc_sort_radix_par_u32": (None, [_vp, _vp, _sz, _sz]),
    "orc_sort_radix_par_alloc_u32": (None, [_vp, _sz, _sz]),
    "orc_sort_two_phase_u32": (None, [_vp, _vp, _sz]),
    "orc_sort_radix_par_u64": (None, [_vp, _vp, _sz, _sz]),
    "orc_sort_radix_bits_u32": (None, [_vp, _vp, _sz, _u, _u, _sz]),
    "orc_sort_radix_bits_u64": (None, [_vp, _vp, _sz, _u, _u, _sz]),
    "orc_sort_kv_u64_u32": (None, [_vp, _vp, _vp, _vp, _sz, _sz]),
    "orc_sort_kv_u32_u32": (None, [_vp, _vp, _vp, _vp, _sz, _sz]),
    "orc_f32_to_ordered": (_u32, [_u32]),
    "orc_ordered_to_f32": (_u32, [_u32]),
    "orc_f32_transform_exhaustive_check": (_u64, []),
    "orc_sort_f32": (None, [_vp, _vp, _sz, _sz]),
    "orc_sort_f64": (None, [_vp, _vp, _sz, _sz]),
    "orc_sort_i32": (None, [_vp, _vp, _sz, _sz]),
    "orc_sort_i64": (None, [_vp, _vp, _sz, _sz]),
    "orc_hist_byte_components_u32": (None, [_vp, _sz, _sz, _sz, _vp]),
    "orc_hist_qc_par_u32": (None, [_vp, _sz, _sz, _sz, _sz, _u, _sz, _vp]),
    "orc_compute_start_of_bins_u32": (None, [_vp, _sz, _sz, _sz, _u, _sz, _vp]),
    "orc_hist_one_byte_u32": (None, [_vp, _sz, _sz, _u, _u, _sz, _vp]),
    "orc_hist_one_byte_2_bugcompat_u32": (None, [_vp, _sz, _sz, _u, _sz, _vp]),
    "orc_hist_bytes": (None, [_vp, _sz, _sz, _sz, _vp]),
    "orc_counting_sort_u8": (None, [_vp, _sz, _sz, _sz]),
    "orc_select_radix_u32": (_u32, [_vp, _sz, _sz, _sz, _vp]),
    "orc_partition_radix_u32": (None, [_vp, _sz, _sz, _vp, _sz]),
    "orc_merge_parallel_l5_u32": (None, [_vp, _sz, _sz, _sz, _sz, _vp, _sz, _sz]),
    "orc_merge_parallel_l5_u64": (None, [_vp, _sz, _sz, _sz, _sz, _vp, _sz, _sz]),
    "orc_counting_pass_u32": (None, [_vp, _vp, _sz, _u]),
}

_lib = None
_ref = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _lib = lib
    return _lib


def load_ref():
    """The partial reference build (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        lib = ctypes.CDLL(REF_LIB_PATH)
        lib.ref_binary_search_u32.restype = _sz
        lib.ref_binary_search_u32.argtypes = [_u32, _vp, _sz, _sz]
        lib.ref_insertion_sort_u32.argtypes = [_vp, _sz]
        lib.ref_hist_one_component_opt_u32.argtypes = [_vp, _sz, _sz, _u, _u, _vp]
        lib.ref_hist_byte_components_serial_u32.argtypes = [_vp, _sz, _sz, _vp]
        lib.ref_extract_digit_1_u32.restype = _u32
        lib.ref_extract_digit_1_u32.argtypes = [_u32, _u32, _u32]
        _ref = lib
    return _ref


def _p(a):
    return a.ctypes.data if a is not None else None


def set_threads(n):
    load().orc_set_threads(n)


def get_threads():
    return load().orc_get_threads()


def sort_radix_par_u32(a, thr=512 * 1024):
    """SortRadixPar(a, tmp, n, thr) on a copy; returns the sorted copy."""
    a = np.array(a, dtype=np.uint32, copy=True)
    tmp = np.empty_like(a)
    load().orc_sort_radix_par_u32(_p(a), _p(tmp), a.size, thr)
    return a


def sort_radix_par_alloc_u32(a, thr=64 * 1024):
    a = np.array(a, dtype=np.uint32, copy=True)
    load().orc_sort_radix_par_alloc_u32(_p(a), a.size, thr)
    return a


def sort_two_phase_u32(a):
    a = np.array(a, dtype=np.uint32, copy=True)
    b = np.empty_like(a)
    load().orc_sort_two_phase_u32(_p(a), _p(b), a.size)
    return a, b


def sort_u64(a, thr=512 * 1024):
    a = np.array(a, dtype=np.uint64, copy=True)
    tmp = np.empty_like(a)
    load().orc_sort_radix_par_u64(_p(a), _p(tmp), a.size, thr)
    return a


def sort_radix_bits(a, bits, max_passes=64, thr=512 * 1024):
    """SortRadixInnerPar with `bits`-bit digits (4..8) on a u32 or u64 array;
    max_passes stops after that many passes (the state the tests inspect)."""
    a = np.ascontiguousarray(a).copy()
    tmp = np.empty_like(a)
    fn = load().orc_sort_radix_bits_u64 if a.dtype == np.uint64 else load().orc_sort_radix_bits_u32
    fn(_p(a), _p(tmp), a.size, bits, max_passes, thr)
    return a


def sort_kv_u64_u32(k, v, thr=512 * 1024):
    k = np.array(k, dtype=np.uint64, copy=True)
    v = np.array(v, dtype=np.uint32, copy=True)
    kt, vt = np.empty_like(k), np.empty_like(v)  # keep the scratch alive across the call
    load().orc_sort_kv_u64_u32(_p(k), _p(v), _p(kt), _p(vt), k.size, thr)
    return k, v


def sort_kv_u32_u32(k, v, thr=512 * 1024):
    k = np.array(k, dtype=np.uint32, copy=True)
    v = np.array(v, dtype=np.uint32, copy=True)
    kt, vt = np.empty_like(k), np.empty_like(v)  # keep the scratch alive across the call
    load().orc_sort_kv_u32_u32(_p(k), _p(v), _p(kt), _p(vt), k.size, thr)
    return k, v


def sort_f32_bits(bits, thr=512 * 1024):
    bits = np.array(bits, dtype=np.uint32, copy=True)
    tmp = np.empty_like(bits)
    load().orc_sort_f32(_p(bits), _p(tmp), bits.size, thr)
    return bits


def sort_f64_bits(bits, thr=512 * 1024):
    """float64 bit patterns sorted in the total order of the transformed bits."""
    bits = np.array(bits, dtype=np.uint64, copy=True)
    tmp = np.empty_like(bits)
    load().orc_sort_f64(_p(bits), _p(tmp), bits.size, thr)
    return bits


def sort_i32(a, thr=512 * 1024):
    a = np.array(a, dtype=np.int32, copy=True)
    tmp = np.empty_like(a)
    load().orc_sort_i32(_p(a), _p(tmp), a.size, thr)
    return a


def sort_i64(a, thr=512 * 1024):
    a = np.array(a, dtype=np.int64, copy=True)
    tmp = np.empty_like(a)
    load().orc_sort_i64(_p(a), _p(tmp), a.size, thr)
    return a


def f32_to_ordered(u):
    return load().orc_f32_to_ordered(u)


def ordered_to_f32(v):
    return load().orc_ordered_to_f32(v)


def hist_byte_components_u32(a, l, r, thr=64 * 1024):
    """8 x 256 counts (rows 4..7 zero), like the reference's return shape."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.zeros((8, 256), dtype=np.uint64)
    load().orc_hist_byte_components_u32(_p(a), l, r, thr, _p(out))
    return out


def hist_qc_par_u32(a, l, r, quantum, nq, which_byte, thr=16 * 1024):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.zeros((nq, 256), dtype=np.uint64)
    load().orc_hist_qc_par_u32(_p(a), l, r, quantum, nq, which_byte, thr, _p(out))
    return out


def compute_start_of_bins_u32(a, quantum, nq, digit, thr=16 * 1024):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.zeros((nq, 256), dtype=np.uint64)
    load().orc_compute_start_of_bins_u32(_p(a), a.size, quantum, nq, digit, thr, _p(out))
    return out


def hist_one_byte_u32(a, l, r, shift, radix=256, thr=64 * 1024):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.zeros(radix, dtype=np.uint64)
    load().orc_hist_one_byte_u32(_p(a), l, r, shift, radix, thr, _p(out))
    return out


def hist_one_byte_2_bugcompat_u32(a, l, r, shift, thr=64 * 1024):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.zeros(256, dtype=np.uint64)
    load().orc_hist_one_byte_2_bugcompat_u32(_p(a), l, r, shift, thr, _p(out))
    return out


def hist_bytes(a, l, r, thr=64 * 1024):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    out = np.zeros(256, dtype=np.uint64)
    load().orc_hist_bytes(_p(a), l, r, thr, _p(out))
    return out


def counting_sort_u8(a, l=0, r=None, thr=64 * 1024):
    """counting_sort_parallel_inner<256> on a copy, sorting [l, r)."""
    a = np.array(a, dtype=np.uint8, copy=True)
    if r is None:
        r = a.size
    load().orc_counting_sort_u8(_p(a), l, r, thr)
    return a


def select_radix_u32(a, k, start=0, length=None):
    """SelectRadix on a copy; returns (value, candidate counts per pass)."""
    a = np.array(a, dtype=np.uint32, copy=True)
    if length is None:
        length = a.size - start
    pc = np.zeros(4, dtype=np.uint64)
    v = load().orc_select_radix_u32(_p(a), start, length, k, _p(pc))
    return v, pc


def partition_radix_u32(a, ks, start=0, length=None):
    """PartitionRadix on a copy with ranks ks (ascending, absolute); returns the array."""
    a = np.array(a, dtype=np.uint32, copy=True)
    if length is None:
        length = a.size - start
    k = np.ascontiguousarray(ks, dtype=np.uint64)
    load().orc_partition_radix_u32(_p(a), start, length, _p(k), k.size)
    return a


def merge_parallel_l5_u32(t, p1, r1, p2, r2, out, p3, thr=32768):
    t = np.ascontiguousarray(t, dtype=np.uint32)
    load().orc_merge_parallel_l5_u32(_p(t), p1, r1, p2, r2, _p(out), p3, thr)
    return out


def merge_parallel_l5_u64(t, p1, r1, p2, r2, out, p3, thr=32768):
    t = np.ascontiguousarray(t, dtype=np.uint64)
    load().orc_merge_parallel_l5_u64(_p(t), p1, r1, p2, r2, _p(out), p3, thr)
    return out


def counting_pass_u32(a, shift):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.empty_like(a)
    load().orc_counting_pass_u32(_p(a), _p(out), a.size, shift)
    return out


def binary_search_u32(value, a, left, right):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return load().orc_binary_search_u32(value, _p(a), left, right)


def insertion_sort_u32(a):
    a = np.array(a, dtype=np.uint32, copy=True)
    load().orc_insertion_sort_u32(_p(a), a.size)
    return a
