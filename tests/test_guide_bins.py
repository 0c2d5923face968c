"""The guide's direction -> bin map (lambda_amd/csrc/guide.h GuideBin, evaluated without an
arctangent) against the definition of its bins: built for the host with hipcc and run on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_guide_bin_matches_its_definition(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not found")
    exe = str(tmp_path / "guide_bins")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-o", exe,
                    os.path.join(HERE, "c", "guide_bins.cpp")], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
