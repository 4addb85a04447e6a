"""The C ABI boundary, checked without a GPU: the library loads, exports every
function include/pa_hip.h declares, reports errors as status codes, and the
drop-in C++ headers compile against it."""
import ctypes
import os
import shutil
import subprocess

import pytest

import parallelalgorithms_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(P.LIB_PATH)
    declared = P.header_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(P.header_symbols()) == set(P._SIGS)


def test_status_strings():
    lib = P.load()
    assert lib.pa_status_string(0) == b"ok"
    assert lib.pa_status_string(1) == b"invalid argument"
    assert lib.pa_version() >= 10000


def test_argument_errors_need_no_device():
    lib = P.load()
    # null keys with n > 0 is rejected before any HIP call
    assert lib.pa_sort_u32(None, None, 10, None) == 1
    assert lib.pa_select_u32(None, 10, 1, None, None) == 1
    assert lib.pa_hist_u32(None, 10, 0, 13, None, None) == 1
    # n == 0 is a no-op success
    assert lib.pa_sort_u32(None, None, 0, None) == 0


def test_size_class_choice_needs_no_device(monkeypatch):
    """Built-in pass shape per size class: the 8K-key tile up to 2^24 keys
    (32-lane ranking units up to 2^20, 16-lane above), the split 24K-key tile
    as 768 x 32 with 16-lane ranking units above (unless PA_SORT_VARIANT
    forces one)."""
    import os
    if os.environ.get("PA_SORT_VARIANT"):
        pytest.skip("PA_SORT_VARIANT forces one variant")
    lib = P.load()
    assert lib.pa_autotune_reset() == 0
    assert [lib.pa_sort_variant_u32(n) for n in (1, 1 << 16, (1 << 16) + 1, 1 << 20)] == [2, 2, 2, 2]
    assert [lib.pa_sort_variant_u32(n) for n in ((1 << 20) + 1, 1 << 24, (1 << 24) + 1, 100_000_000)] == [7, 7, 8, 8]
    assert lib.pa_autotune_sort_u32(1, None) == 1  # needs max_n >= 2; rejected before any HIP call
    assert lib.pa_segmented_plan_create_u32(None, None, 0) is None


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_dropin_headers_compile(tmp_path):
    src = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
    exe = tmp_path / "dropin_test"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-o", str(exe),
           "-L", os.path.dirname(P.LIB_PATH), "-lpa_hip", f"-Wl,-rpath,{os.path.dirname(P.LIB_PATH)}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
