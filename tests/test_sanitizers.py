"""Sanitizer runs of host code on the CPU (tests/native/Makefile builds them):
the segment server's scheduler with a mock device under the thread sanitizer
and under the address and undefined-behaviour sanitizers, the CPU oracle over
seeded small cases and libfbp's host code on its no-device paths under the
address and undefined-behaviour sanitizers.  A report of any sanitizer fails
the test."""
import os
import subprocess

import pytest

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")
OUT = os.path.join(NATIVE, "_build")
REPORTS = ("ThreadSanitizer", "AddressSanitizer", "LeakSanitizer", "runtime error:")


def _run(target, env=None, timeout=300):
    b = subprocess.run(["make", "-s", "-j6", "-C", NATIVE, os.path.join(OUT, target)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=1200)
    assert b.returncode == 0, b.stdout
    r = subprocess.run([os.path.join(OUT, target)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout, env=dict(os.environ, **(env or {})))
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    for rep in REPORTS:
        assert rep not in r.stdout, r.stdout
    return r.stdout


def test_scheduler_thread_sanitizer():
    out = _run("server_core_tsan", {"TSAN_OPTIONS": "halt_on_error=1"})
    assert "all scheduler checks passed" in out


def test_scheduler_address_sanitizer():
    out = _run("server_core_asan", {"ASAN_OPTIONS": "detect_leaks=1"})
    assert "all scheduler checks passed" in out


def test_oracle_address_sanitizer():
    out = _run("oracle_asan", {"OMP_NUM_THREADS": "4"})
    assert "oracle checks passed" in out


def test_host_runtime_address_sanitizer():
    # the HIP runtime's own allocations at exit are not the library's: leak
    # checking is off, every access and UB check stays on
    out = _run("host_runtime_asan", {"ASAN_OPTIONS": "detect_leaks=0"}, timeout=120)
    assert "all host runtime checks passed" in out
