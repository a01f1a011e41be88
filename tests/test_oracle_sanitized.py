"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile, target sanitize; gcc, CPU code only) runs every entry
point on small seeded inputs (oracle/sanitize_check.c) without a report."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, capture_output=True)
    exe = os.path.join(ROOT, "build", "oracle", "sanitize_check")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "oracle sanitize check ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
