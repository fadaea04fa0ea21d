"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile, target
`sanitize`) runs every entry point on a small scene without a report: substeps (FLIP and PIC),
the reverse pass, the mesh collider in both distance modes, CFL, the frame driver, SpMV and MIC."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parents[1] / "oracle"


def test_oracle_under_asan_and_ubsan():
    if shutil.which(os.environ.get("CXX", "g++")) is None:
        pytest.skip("no C++ compiler")
    subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(ORACLE / "sanitize_driver")], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
    assert "sanitized oracle run: ok" in r.stdout
