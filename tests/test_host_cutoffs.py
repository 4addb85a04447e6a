"""The drop-in headers' host cut-offs run without a GPU: compiles
tests/cpp/host_cutoff_test.cpp against libpa_hip.so and runs it on the CPU
(sorts below 100 keys and merges up to 32,768 keys of host arrays never reach
the device)."""
import os
import shutil
import subprocess

import pytest

import parallelalgorithms_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(P.LIB_PATH), reason="no compiler or library")
@pytest.mark.parametrize("sanitize", [[], ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]],
                         ids=["plain", "asan_ubsan"])
def test_host_cutoffs_without_gpu(tmp_path, sanitize):
    """The header code (insertion sort, host merges) also runs clean under
    AddressSanitizer and UndefinedBehaviorSanitizer."""
    exe = tmp_path / "host_cutoff_test"
    libdir = os.path.dirname(P.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", *sanitize, "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "host_cutoff_test.cpp"), "-o", str(exe),
           "-L", libdir, "-lpa_hip", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
