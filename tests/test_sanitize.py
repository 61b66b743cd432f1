"""Host sanitizer runs of the CPU-side code: tests/sanitize/host_check.cpp
built with AddressSanitizer + UndefinedBehaviorSanitizer and with
ThreadSanitizer (g++ only, no HIP), over the generators, the text format
(malformed files included), the host filter on several threads, the thread
gang of the device-to-host copies and the oracle's worker threads.  A
sanitizer report or a failed check exits non-zero."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_under_sanitizer(kind, tmp_path):
    subprocess.run(["make", "-s", "-C", HERE, kind], check=True, capture_output=True, timeout=300)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=24"
    env["TSAN_OPTIONS"] = "halt_on_error=1:exitcode=25"
    r = subprocess.run([os.path.join(HERE, f"host_check_{kind}"), str(tmp_path), "4"], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "host_check: ok" in r.stdout
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
