"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: oracle/asan_driver.c runs
every oracle function over small seeded inputs and edge cases (empty inputs, NULL sentinels,
duplicate keys, STRING columns, capacity overflow) and checks each result against brute force or an
invariant; any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_clean_under_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle_asan", f"OUT={tmp_path}"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(tmp_path / "oracle_asan")], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "all checks passed" in r.stdout
