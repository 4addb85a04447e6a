"""The oracle's tests under AddressSanitizer and UndefinedBehaviorSanitizer on
the CPU: builds oracle/_asan/liboracle.so (`make -C oracle asan`) and re-runs
tests/test_oracle.py in a child process that loads it, with the sanitizer
runtimes preloaded into the (uninstrumented) Python interpreter.  Child
processes first call a planted heap overflow and a planted over-wide shift
built with each sanitizer, to show that the set-up reports errors rather than
hiding them."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "oracle", "_asan", "liboracle.so")


def _runtime(name):
    r = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _env():
    env = dict(os.environ)
    env["LD_PRELOAD"] = ":".join([_runtime("libasan.so"), _runtime("libubsan.so")])
    env["ASAN_OPTIONS"] = "detect_leaks=0:protect_shadow_gap=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["PA_ORACLE_LIB"] = ASAN_LIB
    return env


pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None
                                or _runtime("libasan.so") is None or _runtime("libubsan.so") is None,
                                reason="no g++ / make / sanitizer runtimes")


@pytest.mark.parametrize("flag, body, report", [
    ("address", "int* p = (int*)std::malloc(4 * sizeof(int)); p[i] = 1; int v = p[0]; std::free(p); return v;",
     "heap-buffer-overflow"),
    ("undefined", "return 1 << (i + 28);", "runtime error: shift exponent"),
])
def test_sanitizers_report_a_planted_error(tmp_path, flag, body, report):
    src = tmp_path / "planted.cpp"
    src.write_text('#include <cstdlib>\nextern "C" int planted(int i) { ' + body + ' }\n')
    lib = tmp_path / "libplanted.so"
    r = subprocess.run(["g++", "-O0", "-g", "-fPIC", "-shared", f"-fsanitize={flag}", "-fno-sanitize-recover=all",
                        str(src), "-o", str(lib)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    code = f"import ctypes; ctypes.CDLL({str(lib)!r}).planted(4)"
    r = subprocess.run([sys.executable, "-c", code], env=_env(), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and report in r.stderr, r.stderr[-2000:]


def test_oracle_tests_clean_under_asan_ubsan():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    probe = subprocess.run([sys.executable, "-c", "import oracle; oracle.load(); print(oracle.LIB_PATH)"],
                           cwd=ROOT, env=_env(), capture_output=True, text=True, timeout=120)
    assert probe.returncode == 0 and probe.stdout.strip() == ASAN_LIB, probe.stdout + probe.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle.py")],
                       cwd=ROOT, env=_env(), capture_output=True, text=True, timeout=900)
    tail = r.stdout[-3000:] + r.stderr[-3000:]
    assert r.returncode == 0, tail
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
