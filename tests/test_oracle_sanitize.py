"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile `sanitize`) runs every entry point on small inputs --
direct and FMM sums, the core-size derivatives, Estr, the integrators with
the dynamic SFS, static particles and both viscous schemes -- without a
finding (any finding aborts the run)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_asan_and_ubsan():
    p = subprocess.run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "oracle sanitize run: ok" in p.stdout
