"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

`make -C oracle sanitize` builds oracle/_san/sanitize_check: the oracle's
sources with -fsanitize=address,undefined and a driver that calls every
entry point (dense kernels, the reference path in 1-D/2-D/3-D for all five
bases with its CSR lists, weights, apply and target gradient, the drop-in
restatement, the density-aware builder, the nearest-source transfer, MLS)
on small clouds. Any out-of-bounds access, leak or undefined behaviour ends
the run with a report and a non-zero status.
"""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_clean_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"),
                        "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san",
                                     "sanitize_check")],
                       capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "sanitize_check: ok" in r.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out
