"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle (banded restatement, its vectorised port, full SW, chaining), the
synthetic generator and the product library's host-only pair-file and
packed-file I/O, built by `make -C oracle sanitize` with every sanitizer
report fatal and driven by oracle/san_driver.cpp (ragged batch sizes on every
workload generator, cross-checks between the restatements, file round
trips).  The GPU side of the host layer is not covered here: its C ABI needs a
device."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("g++") is None,
                    reason="needs make and g++")
def test_sanitized_cpu_run(tmp_path):
    b = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8", "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout[-2000:] + b.stderr[-2000:]
    exe = os.path.join(ROOT, "oracle", "_build", "san", "san_driver")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600,
                       env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert "all checks passed" in r.stdout
