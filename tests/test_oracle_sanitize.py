"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/sanitize_check.cpp drives every exported routine over even, odd and
one-cell-thick grids, every boundary mix, pin on and off, both solver
semantics); any out-of-bounds access or undefined operation aborts the run."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                    reason="needs g++ and make")
def test_oracle_under_address_and_ub_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize-check"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
