"""The ctypes binding refuses arrays the library would misread: wrong element
size, strided views, short payload / scratch / count buffers.  These checks run
before any library call, so they need no device."""
import ctypes

import numpy as np
import pytest

import parallelalgorithms_amd as P


@pytest.mark.parametrize("call", [
    lambda: P.sort_u32(np.zeros(8, np.int64)),
    lambda: P.sort_u64(np.zeros(8, np.uint32)),
    lambda: P.sort_f32(np.zeros(8, np.float64)),
    lambda: P.sort_i64(np.zeros(8, np.int32)),
    lambda: P.counting_sort_u8(np.zeros(8, np.uint16)),
    lambda: P.select_u32(np.zeros(8, np.uint64), 0, np.zeros(1, np.uint32)),
], ids=["u32-as-int64", "u64-as-u32", "f32-as-f64", "i64-as-i32", "u8-as-u16", "select-u64-keys"])
def test_element_size_is_checked(call):
    with pytest.raises(ValueError, match="bytes"):
        call()


def test_strided_view_is_refused():
    with pytest.raises(ValueError, match="contiguous"):
        P.sort_u32(np.zeros(16, np.uint32)[::2])


@pytest.mark.parametrize("call, what", [
    (lambda: P.sort_u32(np.zeros(8, np.uint32), np.zeros(7, np.uint32)), "tmp"),
    (lambda: P.sort_kv_u32_u32(np.zeros(8, np.uint32), np.zeros(4, np.uint32)), "vals"),
    (lambda: P.sort_kv_u64_u32(np.zeros(8, np.uint64), np.zeros(8, np.uint32), np.zeros(8, np.uint64),
                               np.zeros(7, np.uint32)), "vtmp"),
    (lambda: P.hist_digits_u32(np.zeros(8, np.uint32), np.zeros(1023, np.uint64)), "counts"),
    (lambda: P.hist_u32(np.zeros(8, np.uint32), 20, 12, np.zeros(4095, np.uint64)), "counts"),
    (lambda: P.hist_u8(np.zeros(8, np.uint8), np.zeros(255, np.uint64)), "counts"),
    (lambda: P.hist_quanta_u32(np.zeros(8, np.uint32), 0, 8, 4, 2, 0, np.zeros(511, np.uint64)), "counts"),
    (lambda: P.start_of_bins_u32(np.zeros(8, np.uint32), 4, 2, 0, np.zeros(511, np.uint64)), "out"),
    (lambda: P.merge_u32(np.zeros(4, np.uint32), np.zeros(4, np.uint32), np.zeros(7, np.uint32)), "out"),
    (lambda: P.select_u32(np.zeros(8, np.uint32), 0, np.zeros(0, np.uint32)), "result"),
    (lambda: P.segmented_sort_u32(np.zeros(8, np.uint32), np.zeros(8, np.uint32),
                                  np.array([0, 4, 9], np.uint64)), "offsets"),
])
def test_short_buffers_are_refused(call, what):
    with pytest.raises(ValueError, match=what):
        call()


def test_valid_arguments_pass_the_checks():
    P._check(np.zeros(8, np.uint32), 4, "keys", 8)
    P._check(None, 4, "tmp", 8)


def test_job_scratch_charge():
    """The server's per-job charge (pa_job_scratch_bytes) needs no device: it
    is the library's own layout, rounded as the allocator rounds (+1/8, whole
    MiB).  A device sort of n u32 keys needs at least its n-key ping-pong
    buffer; select at least its n/2-key compaction buffer."""
    M = 1 << 20
    for n in (1, 1000, 10 * M, 100 * M, (1 << 30) + 5):
        b = P.Server.job_scratch_bytes(P.OP_SORT, n, P.KEY_U32)
        assert b % M == 0 and b >= 4 * n
        assert P.Server.job_scratch_bytes(P.OP_SORT, n, P.KEY_U64) >= 8 * n
        assert P.Server.job_scratch_bytes(P.OP_SELECT, n, P.KEY_U32) >= 4 * (n // 2)
    sizes = [P.Server.job_scratch_bytes(P.OP_SORT, n, P.KEY_U32) for n in (M, 2 * M, 4 * M, 8 * M, 16 * M)]
    assert sizes == sorted(sizes)
    assert P.Server.job_scratch_bytes(P.OP_HIST, 100 * M, P.KEY_U32, 12) == M  # counts only
    with pytest.raises(P.PAError):
        P.call("pa_job_scratch_bytes", None, None)


def test_partition_offsets_grid_is_checked_on_the_host():
    """The partition's offsets pass is one 1024-thread workgroup serving one
    block per thread: the host refuses any other grid before launching."""
    chk = P.load().pa_partition_offsets_grid_check
    chk.restype = ctypes.c_int
    chk.argtypes = [ctypes.c_size_t]
    assert chk(1) == P.PA_OK and chk(1024) == P.PA_OK
    for nb in (0, 1025, 2048, 8192):
        assert chk(nb) == 1  # PA_ERR_ARG
        assert "1024" in P.load().pa_last_error_string().decode()
