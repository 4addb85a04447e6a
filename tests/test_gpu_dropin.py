"""The drop-in C++ headers end to end on the GPU: compiles tests/cpp/dropin_test.cpp
(the reference's self-checking benchmarks re-run through include/*.h with host
std::vector inputs) against libpa_hip.so and runs it."""
import os
import shutil
import subprocess

import pytest

import parallelalgorithms_amd as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_dropin_headers_self_check(tmp_path):
    exe = tmp_path / "dropin_test"
    libdir = os.path.dirname(P.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp"), "-o", str(exe),
           "-L", libdir, "-lpa_hip", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    print("[dropin] running dropin_test 10000000", flush=True)
    r = subprocess.run([str(exe), "10000000"], capture_output=True, text=True, timeout=110)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_forwards_cutoffs_and_concurrent_callers(tmp_path):
    """tests/cpp/forwards_test.cpp: the forwards the benchmarks do not call,
    the host cut-offs (sorts below 100 keys, merges up to 32,768) and
    merge_parallel_L5 / SortRadixPar from 8 host threads at once."""
    exe = tmp_path / "forwards_test"
    libdir = os.path.dirname(P.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O2", "-pthread", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "forwards_test.cpp"), "-o", str(exe),
           "-L", libdir, "-lpa_hip", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    print("[dropin] running forwards_test", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=110)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
