"""The oracle and the library's host C++ (terrain generator, host table build, image
validation) under AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer:
`make -C oracle sanitize` builds oracle/san_driver.cpp with each, and the driver runs every
editor kind, a traced frame, GC, the colour codec and the image checks on a 2^8 scene with
several threads. Any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    r = subprocess.run(["make", "-C", ORACLE, "sanitize", "-j2"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:]
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("tool,env", [("asan", {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1",
                                                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}),
                                      ("tsan", {"TSAN_OPTIONS": "halt_on_error=1"})])
def test_sanitized_driver(built, tool, env):
    r = subprocess.run([os.path.join(built, tool), "4"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=900, env={**os.environ, **env}, cwd="/tmp")
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout[-4000:]
    assert "Sanitizer" not in r.stdout, r.stdout[-4000:]
