"""CPU tests of the C-ABI library: it loads, exports every declared entry
point, and its host-only paths behave without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from lambda_amd import abi, gpu, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lambda_gpu.h")).read()
    return sorted(set(re.findall(r"\b(lgpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = gpu.lib()
    names = declared_symbols()
    assert len(names) >= 28
    for n in names:
        assert hasattr(L, n), n
    assert sorted(gpu.EXPORTS) == names


def test_abi_version():
    assert gpu.lib().lgpu_abi_version() == abi.ABI_VERSION


def test_struct_sizes_match_header_layout():
    # descriptors are passed by pointer; their sizes must match the C layout
    assert C.sizeof(abi.Node) == 4 + 16 + 48
    assert C.sizeof(abi.RenderStats) == 5 * 8 + 5 * 8 + 4 * 4 + 3 * 8 + 3 * 8
    assert C.sizeof(abi.Camera) == 4 + 12 * 4 + 3 * 4 + 4 + 4 + 4 + 8
    # scene, film, camera, 10 words, utility-pass fields (6 words),
    # ... padding to 8, the guide handle, the training-step count, padded to 8
    assert C.sizeof(abi.RenderDesc) == 8 + 8 + C.sizeof(abi.Camera) + 10 * 4 + 6 * 4 + 4 + 8 + 4 + 4


def test_camera_helpers_match_python_builder():
    cam = abi.Camera()
    origin = (C.c_float * 3)(0.0, 0.0, 3.8)
    gpu.lib().lgpu_camera_init(C.byref(cam), abi.CAMERA_PINHOLE, origin, 256.0, 256.0)
    gpu.lib().lgpu_camera_set_rotation(C.byref(cam), np.float32(np.pi), 0.0)
    gpu.lib().lgpu_camera_set_fov(C.byref(cam), 0.6911)
    ref = scene.cornell_camera(256, 256)
    np.testing.assert_allclose(list(cam.z_hat), list(ref.z_hat), atol=1e-6)
    np.testing.assert_allclose(list(cam.x_hat), list(ref.x_hat), atol=1e-6)
    np.testing.assert_allclose(list(cam.y_hat), list(ref.y_hat), atol=1e-6)
    assert cam.tan_fov == pytest.approx(ref.tan_fov, rel=1e-6)
    assert cam.aspect == 1.0


def test_errors_are_codes_not_exceptions():
    L = gpu.lib()
    # null arguments are rejected with LGPU_ERR_INVALID_ARGUMENT and a message
    assert L.lgpu_scene_commit(None, 0) == abi.ERR_INVALID_ARGUMENT
    assert b"null" in L.lgpu_last_error()
    assert L.lgpu_render(None, None) == abi.ERR_INVALID_ARGUMENT
    assert L.lgpu_film_read(None, None, None) == abi.ERR_INVALID_ARGUMENT


def test_device_create_without_gpu_reports_no_device():
    h = C.c_void_p()
    st = gpu.lib().lgpu_device_create(0, C.byref(h))
    if st == abi.OK:  # running on a GPU box
        gpu.lib().lgpu_device_destroy(h)
        pytest.skip("a GPU is present")
    assert st == abi.ERR_NO_DEVICE


def test_progressive_validates_arguments_without_a_device():
    """Null descriptors and handles are rejected before any device call."""
    L = gpu.lib()
    out = C.c_void_p()
    assert L.lgpu_progressive_create(None, C.byref(out)) == abi.ERR_INVALID_ARGUMENT
    assert L.lgpu_progressive_start(None) == abi.ERR_INVALID_ARGUMENT
    assert L.lgpu_progressive_read(None, None, None) == abi.ERR_INVALID_ARGUMENT
    assert L.lgpu_progressive_destroy(None) == abi.OK
    assert L.lgpu_progressive_running(None) == 0
