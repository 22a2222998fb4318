"""The C restatement under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only).

oracle/asan_check.c (built by oracle/sanitize.mk) drives restate.c's entry
points (half norms, assignment, the Lloyd loop with empty-cluster
relocation) on edge shapes with every array allocated at its exact size: n not a multiple of the 32-sample block,
d = 1, k = 1, n < k, centroids far from the data, max_iter = 0, float32 and
float64, one and two threads.  The sanitized build must run without a report
(any report aborts it with a non-zero status) and print the same digests of
labels, centroids and inertia as the plain build."""
import os
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
BUILD = os.path.join(ORACLE, "build")


@pytest.fixture(scope="module")
def harness():
    try:
        subprocess.run(["make", "-s", "-B", "-f", os.path.join(ORACLE, "sanitize.mk")],
                       check=True, capture_output=True, text=True, timeout=300)
    except subprocess.CalledProcessError as e:  # pragma: no cover - toolchain
        pytest.skip(f"sanitizer build unavailable: {e.stderr[-400:]}")
    return os.path.join(BUILD, "asan_check"), os.path.join(BUILD, "asan_check_plain")


def test_restatement_runs_clean_under_asan_and_ubsan(harness):
    sanitized, plain = harness
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    a = subprocess.run([sanitized], capture_output=True, text=True, env=env, timeout=300)
    assert a.returncode == 0, a.stderr[-4000:]
    assert "Sanitizer" not in a.stderr and "runtime error" not in a.stderr, a.stderr[-4000:]
    p = subprocess.run([plain], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-4000:]
    lines = a.stdout.splitlines()
    assert len(lines) == 16                     # 8 shapes x 2 dtypes, one thread
    assert a.stdout == p.stdout
    # max_iter = 0 runs no iteration; a single sample ends after one
    assert "n=31 d=2 k=3 max_iter=0 far=0: iters 0 " in a.stdout
    assert "n=1 d=2 k=3 max_iter=2 far=1: iters 1 inertia 0 " in a.stdout
