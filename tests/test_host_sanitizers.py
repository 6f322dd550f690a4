"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/native/host_check.cpp, built by tests/native/build.py with the sanitizers on the host pass
only): the entry points that need no GPU here, and full tessellation, batch, server and static-entry
lifecycles on the GPU. A sanitizer report, a leak or a failed return-code check fails the test."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "native", "_build", "host_check")
# leaks are reported except those of the ROCm runtime's own long-lived allocations (lsan.supp)
ENV = {**os.environ, "ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:exitcode=23",
       "LSAN_OPTIONS": "suppressions=" + os.path.join(HERE, "native", "lsan.supp"),
       "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


def host_check():
    if not os.path.exists(EXE):
        from native import build
        build.build()
    return EXE


def run(args, timeout):
    r = subprocess.run([host_check(), *args], capture_output=True, text=True, timeout=timeout, env=ENV)
    assert r.returncode == 0, f"host_check {' '.join(args)} exited {r.returncode}:\n{r.stdout}\n{r.stderr[-4000:]}"
    return r.stdout


def test_host_entry_points_under_sanitizers():
    out = run([], 300)
    assert "host checks passed" in out


@pytest.mark.gpu
def test_device_lifecycles_under_host_sanitizers():
    out = run(["--device"], 280)
    for part in ("tessellation", "batch", "server", "static entry"):
        assert f"{part} checks passed" in out
