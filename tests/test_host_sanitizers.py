"""The device-free host code of libgsw_hip.so (gpusw_amd/csrc/gsw_host.h: planning thread
pool, explicit-pairs batch planner, serving arena and FIFO admission) built for the CPU
under the thread sanitizer and under the address and undefined-behaviour sanitizers, and run
on synthetic pair lists and request streams (tests/native/host_harness.cpp)."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

NATIVE = Path(__file__).resolve().parent / "native"


@pytest.fixture(scope="module")
def harness():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    subprocess.run(["make", "-s", "-C", str(NATIVE), "-j2"], check=True)
    return NATIVE


@pytest.mark.parametrize("kind,env", [
    ("tsan", {"TSAN_OPTIONS": "halt_on_error=1 exitcode=66"}),
    ("asan", {"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "halt_on_error=1 print_stacktrace=1"}),
])
def test_host_code_under_sanitizer(harness, kind, env):
    proc = subprocess.run([str(harness / f"host_harness_{kind}")], capture_output=True, text=True, timeout=600,
                          env={**os.environ, **env})
    report = proc.stdout + proc.stderr
    assert proc.returncode == 0, report[-4000:]
    assert "host harness ok" in proc.stdout
    assert "ThreadSanitizer" not in report and "AddressSanitizer" not in report and "runtime error" not in report
