"""The kernels' lossless value transforms (sleek_amd/csrc/sleek_device.h,
compiled for the host) equal the oracle's on every float32 bit pattern and
on float64 patterns of every exponent, forward and inverse."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "transform_check")


def test_lossless_transform_matches_oracle_exhaustively():
    if not os.path.exists(EXE):
        if not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")):
            pytest.skip("hipcc not available to build the check")
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "transform_check"], check=True,
                       capture_output=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "transform mismatches: 0" in r.stdout
