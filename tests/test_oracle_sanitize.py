"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every entry point
on small, ragged, periodic and non-periodic grids (viscous orders 6 and 8, closures, the
metric terms) with exactly sized buffers runs clean (oracle/san_driver.c)."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_runs_clean_under_asan_and_ubsan():
    if shutil.which("gcc") is None and shutil.which("cc") is None:
        pytest.skip("no C compiler")
    b = subprocess.run(["make", "-s", "-C", ORACLE, "san_driver"], capture_output=True, text=True)
    if b.returncode != 0:
        if "asan" in b.stderr.lower() or "sanitizer" in b.stderr.lower():
            pytest.skip(f"sanitizer runtime unavailable: {b.stderr[-300:]}")
        raise AssertionError(b.stderr)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "san_driver")], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().endswith("ok")
