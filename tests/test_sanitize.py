"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
entry point on a chain, a comb with a vertex without digit, bond-1 edges,
rank-deficient and zero matrices and an invalid point (oracle/sanitize_main.c,
built by `make -C oracle sanitize`).  A memory or undefined-behaviour error
aborts the driver with a sanitizer report."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([os.path.join(ORACLE, "_san", "oracle_san")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "runtime error" not in p.stderr and "ERROR: AddressSanitizer" not in p.stderr, p.stderr
    assert "ran clean" in p.stdout
