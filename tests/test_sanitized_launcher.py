"""The C-ABI launcher's host code under AddressSanitizer/UBSan: tests/test_abi.py
against the host-instrumented build that tests/sanitize_launcher.py makes.
The build takes minutes, so this test uses it when it is present and up to
date and is skipped otherwise (run `python tests/sanitize_launcher.py`)."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import sanitize_launcher as SL  # noqa: E402


def test_abi_tests_clean_under_address_and_ub_sanitizers():
    if not SL.up_to_date():
        pytest.skip("no up-to-date sanitized build; run: python tests/sanitize_launcher.py")
    r = SL.run_abi_tests(SL.LIB)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
