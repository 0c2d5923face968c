"""The ctypes mirror of include/lambda_gpu.h against the C compiler's layout.

A C translation unit that includes the public header is generated from the ctypes structures,
compiled with gcc and run: it prints sizeof of every descriptor struct and offsetof of every
field, and each must equal what ctypes computes for the mirror (lambda_amd/abi.py). A field added,
removed, reordered or retyped on one side only fails here.
"""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from lambda_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "Node": "lgpu_node", "MaterialDesc": "lgpu_material_desc", "MeshDesc": "lgpu_mesh_desc",
    "LightDesc": "lgpu_light_desc", "MediumDesc": "lgpu_medium_desc", "TextureDesc": "lgpu_texture_desc",
    "Camera": "lgpu_camera", "RenderDesc": "lgpu_render_desc", "RenderStats": "lgpu_render_stats",
    "PostDesc": "lgpu_post_desc", "DenoiseDesc": "lgpu_denoise_desc", "ObjMaterial": "lgpu_obj_material",
    "GuideDesc": "lgpu_guide_desc", "AdjointDesc": "lgpu_adjoint_desc",
    "BatchOptions": "lgpu_batch_options", "BatchStats": "lgpu_batch_stats",
}
C_FIELD = {"in_": "in"}  # Python keyword clashes


def _program():
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "lambda_gpu.h"', "int main(void) {"]
    for py, cname in STRUCTS.items():
        lines.append('  printf("%s %%zu\\n", sizeof(%s));' % (py, cname))
        for f in getattr(abi, py)._fields_:
            name = f[0]
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (py, name, cname, C_FIELD.get(name, name)))
    lines += ["  return 0;", "}"]
    return "\n".join(lines) + "\n"


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_ctypes_mirror_matches_c_layout(tmp_path):
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    src.write_text(_program())
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in out:
        if not line:
            continue
        key, val = line.split()
        val = int(val)
        if "." in key:
            py, field = key.split(".")
            assert getattr(getattr(abi, py), field).offset == val, key
        else:
            assert C.sizeof(getattr(abi, key)) == val, key
        seen += 1
    assert seen == len(STRUCTS) + sum(len(getattr(abi, p)._fields_) for p in STRUCTS)


def test_header_compiles_as_cplusplus(tmp_path):
    """The header is usable from C++ callers (extern "C" block, no C-only constructs)."""
    if shutil.which("g++") is None:
        pytest.skip("no C++ compiler")
    src = tmp_path / "use.cpp"
    src.write_text('#include "lambda_gpu.h"\nint main() { lgpu_render_stats s{}; return (int)s.waves; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o",
                    str(tmp_path / "use"), str(src)], check=True)
