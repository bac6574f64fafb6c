"""The merge's claimed-position bitmaps (wamca-gpu-mi_amd/csrc/mlp_claims.h,
included by the kernels) against brute force, compiled for the host."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_claims_match_brute_force(tmp_path):
    exe = tmp_path / "claims_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "wamca-gpu-mi_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "claims_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout
