"""The reference's lambda* C API shim (include/lambda_compat.h, lambda_amd/csrc/compat.hip).

CPU: the header compiles as C99 and C++17, a C program linking liblambda_gpu.so through it
builds, and its transform checks run (lambdaMakeAffineTransform: translation and scale entries,
a quarter turn about z; no GPU call). The GPU test renders through the shim.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "lambda_amd")
SRC = os.path.join(ROOT, "tests", "c", "lambda_api_smoke.c")


def build(tmp_path, *defs):
    exe = tmp_path / "lambda_api_smoke"
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror", *defs, "-I", os.path.join(ROOT, "include"), "-o", str(exe), SRC,
           "-L", LIB, "-llambda_gpu", "-Wl,-rpath," + LIB, "-lm"]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_c_program_builds_against_the_shim_and_transforms_check(tmp_path):
    if not os.path.exists(os.path.join(LIB, "liblambda_gpu.so")):
        pytest.skip("liblambda_gpu.so not built")
    build(tmp_path)  # the full smoke program links
    exe = build(tmp_path, "-DTRANSFORM_ONLY")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "transforms ok" in out.stdout


def test_header_compiles_as_cplusplus(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no C++ compiler")
    src = tmp_path / "use.cpp"
    src.write_text('#include "lambda_compat.h"\nint main() { LAMBDA_RenderProperties p{}; return (int)p.spp; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o",
                    str(tmp_path / "use"), str(src)], check=True)
