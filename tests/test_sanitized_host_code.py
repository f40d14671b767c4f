"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code (the FP64 oracle
and the library's host geometry, xi-weight and FP32-rounding code), each driven through
its entry points on several geometries (tests/sanitize).  A sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="host compilers absent")
def test_host_code_is_clean_under_asan_and_ubsan():
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(["make", "-s", "-C", HERE, "run"], capture_output=True, text=True, env=env, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "oracle driver: 0 problems" in out and "geometry driver: 0 problems" in out, out[-2000:]
