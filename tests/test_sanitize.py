"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: oracle/sanitize_main.cpp runs
every routine (extraction at five ROI sizes and both models, WLE, the fitters and their cascade,
FP64 fits, CRLB, compaction, rendering, statistics, consecutive merging, the display image, drift
correction, mixed-size extraction, FP64 MLE and full-Fisher CRLB of the 16-parameter model) on
small synthetic frames and on constant, saturated, tiny and empty inputs.  Any out-of-bounds
access, use after free, leak, signed overflow or other undefined behaviour aborts it."""
import os
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_oracle_clean_under_asan_and_ubsan():
    b = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True)
    if b.returncode != 0 and "fsanitize" in (b.stderr + b.stdout) and "cannot find" in (b.stderr + b.stdout):
        pytest.skip("sanitizer runtime libraries are not installed")
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([os.path.join(ORACLE, "_build", "sanitize_oracle")], capture_output=True, text=True, env=env,
                       timeout=600)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "sanitized oracle checks: 17" in r.stdout
