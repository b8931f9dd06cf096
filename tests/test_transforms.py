"""The kernels' value transforms (sleek_amd/csrc/sleek_device.h) against the
oracle: compiled for the host, the lossless code and inverse and both forms
of the error-bounded code and its inverse equal the oracle's on every float32
bit pattern and on float64 patterns of every exponent; on the GPU, the float
form of the float32 error-bounded code equals the integer form on every
float32 pattern."""
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


GPU_EXE = os.path.join(ROOT, "tests", "cpp", "quantize_gpu_check")


@pytest.mark.gpu
def test_float_quantize_matches_integer_form_on_gpu():
    """The encoders' float form of the float32 error-bounded code equals the
    integer form (itself equal to the oracle on every pattern, above) on the
    GPU, over all 2^32 float32 patterns at bounds of every class."""
    assert os.path.exists(GPU_EXE), "tests/cpp/quantize_gpu_check is not built"
    print("quantize_gpu_check: all 2^32 patterns x 7 bounds", flush=True)
    r = subprocess.run([GPU_EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "float-form mismatches: 0" in r.stdout
