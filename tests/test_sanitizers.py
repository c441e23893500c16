"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only;
GPU code is never built with a sanitizer): the oracle driven through every
stage by oracle/sanitize_check.c, and the deck driver (parser, admission,
output writers, restart admission) through the whole host test module with
the sanitized tpic_mi_run. Leak detection is on; any finding aborts the
program and fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def test_oracle_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "sanitize_check")], capture_output=True, text=True,
                       timeout=300, env=ENV)
    assert r.returncode == 0, (r.stdout, r.stderr[-4000:])
    assert "all checks passed" in r.stdout


def test_driver_host_tests_under_sanitizers():
    sys.path.insert(0, ROOT)
    from tpic_amd import build
    if not os.path.exists(build.LIB):
        pytest.fail("libtpic_mi.so is not built: run __graft_entry__.build()")
    driver = build.build_driver(sanitize=True)
    env = dict(ENV, TPIC_MI_RUN=driver)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_host.py")], capture_output=True, text=True, timeout=900,
                       env=env, cwd=ROOT)
    assert r.returncode == 0, (r.stdout[-6000:], r.stderr[-2000:])
    assert " passed" in r.stdout
