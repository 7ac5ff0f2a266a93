"""The float32 -log used by every log-applying kernel (csrc/log32.h).

Host build of the same header: for every positive finite float32 it must
equal (float)(-log((double)x)) with the C library's float64 log -- the
rounding of numpy's float64 log that the oracle applies to float32 data
(reconstruction.py:436 takes -np.log of the float32 array).  The device
build is checked over the same inputs in tests/test_gpu_log32.py.
"""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_neg_log_f32_host_exhaustive(tmp_path):
    exe = str(tmp_path / "log32_exhaustive")
    subprocess.run(
        [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
         "-I", os.path.join(ROOT, "algotom_amd", "csrc"),
         os.path.join(ROOT, "tests", "native", "log32_exhaustive.cpp"),
         "-o", exe], check=True, capture_output=True)
    run = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(run.stdout)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "mismatches 0" in run.stdout
