"""Host sanitizer run of the CPU oracle: `make -C oracle sanitize` builds a
driver that calls every oracle entry point (Euler LW/MDRK in both flux orders,
first-order and MUSCL-Hancock blending with the Dirichlet boundary and the
limiter, slip and Navier-Stokes walls, caller-defined boundary data, linear
advection(-diffusion), the curvilinear solvers, adaptive steps and the dual
tangents) under AddressSanitizer, LeakSanitizer and UndefinedBehaviorSanitizer,
and fails on any report or non-finite output."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_oracle_is_clean_under_asan_and_ubsan(tmp_path):
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize", f"SAN_OUT={tmp_path / 'san'}"],
                       capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "PASSED: 0 failure(s)" in out
    for marker in ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:"):
        assert marker not in out, out[-4000:]
