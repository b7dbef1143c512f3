"""The C ABI under the sanitized debug build (tests/sanitize): the library's host code instrumented
by AddressSanitizer and UndefinedBehaviorSanitizer, and a device synchronization with an error
check after every kernel launch (WUST_DEBUG_SYNC).  The driver (tests/sanitize/abi_driver.cpp)
exercises every entry point; a sanitizer report or a failed expectation fails the test.

CPU: the argument checks that reject input before any device work.  GPU: every entry point on a
40 x 36 operator (create, refined solves in both directions, mass apply, virtual sources,
coefficients, oracle with and without batching, scales, curvature, release, NCG helpers, both
DTFT entries), plus the error paths reached on a live handle.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
DRIVER = HERE / "sanitize" / "_bin" / "abi_driver"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _driver():
    if not DRIVER.exists():
        sys.path.insert(0, str(HERE))
        try:
            from sanitize.build import build
            build()
        finally:
            sys.path.pop(0)
    return DRIVER


def _run(mode, timeout):
    print(f"sanitized abi_driver {mode}", flush=True)
    res = subprocess.run([str(_driver()), mode], capture_output=True, text=True, timeout=timeout, env=ENV)
    assert res.returncode == 0, f"rc={res.returncode}\n{res.stdout}\n{res.stderr[-4000:]}"
    assert f"abi_driver {mode}: ok" in res.stdout
    assert "ERROR: AddressSanitizer" not in res.stderr and "runtime error:" not in res.stderr, res.stderr[-4000:]


def test_sanitized_validation():
    _run("validate", 120)


@pytest.mark.gpu
def test_sanitized_full():
    _run("full", 110)
