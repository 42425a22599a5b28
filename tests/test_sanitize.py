"""The library's host code (sparse numbering, neighbour packing, face rules, file readers and
writers) built with AddressSanitizer and UndefinedBehaviorSanitizer and run on random and edge
-case inputs (tests/sanitize/sanitize_host.cpp).  CPU only: no device code is built or run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", SAN], check=True, timeout=600)
    env = dict(os.environ, SANITIZE_TMP=str(tmp_path), ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    proc = subprocess.run([os.path.join(SAN, "_build", "sanitize_host")], env=env, capture_output=True, text=True,
                          timeout=600)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-4000:]
    assert "sanitized host checks passed" in proc.stdout
