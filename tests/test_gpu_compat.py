"""Renders through the reference's lambda* C API shim on the GPU (tests/c/lambda_api_smoke.c):
device, triangle mesh, proxy, instance, scene commit, film, camera, render properties and
directive, progressive renderer with its callback; the normal pass's colour at the image centre
and an empty edge row."""
import os
import shutil
import subprocess

import pytest

from tests.test_compat import build

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_render_through_the_lambda_api(tmp_path):
    exe = build(tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "smoke ok" in out.stdout
