"""The library's host code and the CPU oracle under AddressSanitizer and
UndefinedBehaviorSanitizer.

tests/sanitize/Makefile builds every library source with the sanitizers on the host side
(device code built as usual, never run here) together with the oracle and
tests/sanitize/host_check.cpp, which drives the paths that run without a device:
descriptor validation and instance resolution over shapes and fields beyond their limits,
tuning-file parsing (well-formed, malformed, oversized), the batched work-queue builder
on random problem sets (every C tiled exactly once), the allocation-free entry point's
host staging at its limits, the sparse and utility argument checks, and the oracle's
routines against brute-force products.  Any sanitizer report aborts the check program.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
EXE = os.path.join(ROOT, "_build", "sanitize", "host_check")


@pytest.mark.skipif(not os.path.exists(os.path.join(SAN, "Makefile")) or
                    shutil.which("make") is None or
                    not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs hipcc and the sanitizer build recipe")
def test_host_code_and_oracle_under_asan_and_ubsan():
    r = subprocess.run(["make", "-C", SAN, "-j", str(min(8, os.cpu_count() or 4))],
                       capture_output=True, text=True, timeout=1200)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               LSAN_OPTIONS="suppressions=" + os.path.join(SAN, "lsan.supp"),
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([EXE], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert " 0 failed" in r.stdout, r.stdout[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, \
        r.stderr[-6000:]
