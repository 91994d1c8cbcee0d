"""Index arithmetic of the seam split (FP32 hexa8 kernel A / kernel B), checked on the
host: the multiply-shift division and the seam-layer flags
(icarat-contact_amd/csrc/icx_internal.h, host-and-device code) against brute force."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_seam_index_arithmetic(tmp_path):
    exe = str(tmp_path / "seam_index_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "seam_index_check.cpp")],
                   check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 mismatches" in out.stdout
