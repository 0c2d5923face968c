"""Progressive renderer (ProgressiveRender.cpp:13-62, Lambda.cpp:431-458) through the C ABI: passes
run on a native worker thread, the callback fires once per pass, the image read back is always a
completed pass, and after k passes the film equals one render of the same sample range (the
random numbers are keyed on the sample index). A render call on another film while the worker runs
is serialised on the device and unaffected."""
import threading
import time

import numpy as np
import pytest

from lambda_amd import abi, gpu, scene
from tests.test_gpu_parity import gpu_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def _wait(p, W, H, passes, timeout=60.0):
    t0 = time.time()
    while time.time() - t0 < timeout:
        img, n = p.read(W, H)
        if n >= passes:
            return
        time.sleep(0.01)
    raise AssertionError("progressive renderer made no progress")


def test_progressive_passes_equal_one_render(dev):
    W = H = 64
    s, keep = gpu_scene(dev, scene.cornell_box())
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    d = gpu.render_desc(s, film, cam, 0, 2, spectral=abi.SPECTRAL_HERO4, estimator=abi.ESTIMATOR_CORRECTED)
    p = gpu.Progressive(d)
    calls = []
    lock = threading.Lock()

    def cb():
        with lock:
            calls.append(1)

    p.set_callback(cb)
    p.start()
    assert p.running()
    # meanwhile a render on another film: serialised with the worker's passes, result unchanged
    other = gpu.Film(dev, W, H)
    gpu.render(gpu.render_desc(s, other, cam, 0, 4, spectral=abi.SPECTRAL_HERO4))
    _wait(p, W, H, 4)
    p.stop()
    assert not p.running()
    img, passes = p.read(W, H)
    assert passes >= 4 and len(calls) == passes
    ref = gpu.Film(dev, W, H)
    gpu.render(gpu.render_desc(s, ref, cam, 0, 2 * passes, spectral=abi.SPECTRAL_HERO4,
                               estimator=abi.ESTIMATOR_CORRECTED))
    a, ca = film.read()
    b, cb_ = ref.read()
    assert (ca == 2 * passes).all() and (cb_ == ca).all()
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=0)
    np.testing.assert_allclose(img, ref.read_rgba(), rtol=1e-6, atol=0)
    solo = gpu.Film(dev, W, H)
    gpu.render(gpu.render_desc(s, solo, cam, 0, 4, spectral=abi.SPECTRAL_HERO4))
    np.testing.assert_allclose(other.read()[0], solo.read()[0], rtol=1e-12, atol=0)
    # restart continues from the next pass; clear starts over
    p.start()
    _wait(p, W, H, passes + 1)
    p.stop()
    p.clear()
    img, n = p.read(W, H)
    assert n == 0 and not img.any()
    assert not film.read()[1].any()
    p.close()


def test_progressive_rejects_bad_descriptors(dev):
    W = H = 8
    s, keep = gpu_scene(dev, scene.cornell_box())
    film = gpu.Film(dev, W, H)
    with pytest.raises(gpu.LambdaError):
        gpu.Progressive(gpu.render_desc(s, film, scene.cornell_camera(W, H), 3, 3))  # empty pass
