"""The C restatement built with AddressSanitizer + UndefinedBehaviorSanitizer and driven once
through every function (oracle/sanitize_main.c): no out-of-bounds access, use after free, leak
or undefined behaviour on a small SinhSymTP grid, consumer path included. Host C only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(HERE, "..", "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_under_address_and_ub_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True, capture_output=True)
    # verify_asan_link_order=0: the environment may preload other libraries ahead of the ASan runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "oracle_asan")], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "adm_rc 0" in r.stdout
