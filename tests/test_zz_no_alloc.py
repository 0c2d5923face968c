"""lgpu_render_noalloc (the render call for callers that cannot allocate) against lgpu_render.

The no-allocation entry point keeps its path pool in a fixed 64 MiB device array. Two calls in a
row, the second on a smaller film, must give the films lgpu_render gives: identical sample
counts, and FP64 sums equal up to the order of the per-wave additions (the pool holds fewer
paths, so the frame is split into more waves; every path's radiance is the same).
"""
import ctypes as C

import numpy as np
import pytest

from lambda_amd import abi, gpu, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def _noalloc(desc):
    st = abi.RenderStats()
    return gpu.lib().lgpu_render_noalloc(C.byref(desc), C.byref(st)), st


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_PATH, abi.INTEGRATOR_DIRECT])
def test_noalloc_matches_render_twice(dev, integrator):
    sd = scene.cornell_box()
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    for W, H, spp in ((256, 256, 4), (64, 48, 2)):  # 262,144 paths (several pool waves), then fewer
        print("noalloc case: integrator %d, %dx%d, %d spp" % (integrator, W, H, spp), flush=True)
        cam = scene.cornell_camera(W, H)
        fa, fb = gpu.Film(dev, W, H), gpu.Film(dev, W, H)
        da = gpu.render_desc(s, fa, cam, 0, spp, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                             estimator=abi.ESTIMATOR_CORRECTED, integrator=integrator)
        db = gpu.render_desc(s, fb, cam, 0, spp, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                             estimator=abi.ESTIMATOR_CORRECTED, integrator=integrator)
        gpu.render(da)
        rc, st = _noalloc(db)
        assert rc == abi.OK, gpu.lib().lgpu_last_error()
        assert st.samples == W * H * spp
        sa, ca = fa.read()
        sb, cb = fb.read()
        assert np.array_equal(ca, cb) and (ca == spp).all()
        assert np.isfinite(sb).all()
        np.testing.assert_allclose(sb, sa, rtol=1e-12, atol=1e-12)
        fa.close()
        fb.close()
    del keep
    s.close()


def test_noalloc_rejects_capacity_beyond_scratch(dev):
    sd = scene.cornell_box()
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    f = gpu.Film(dev, 32, 32)
    d = gpu.render_desc(s, f, scene.cornell_camera(32, 32), 0, 1, capacity=1 << 24)
    rc, _ = _noalloc(d)
    assert rc == abi.ERR_INVALID_ARGUMENT
    assert b"scratch" in gpu.lib().lgpu_last_error()
    _, cnt = f.read()
    assert (cnt == 0).all()  # nothing rendered
    del keep
    f.close()
    s.close()
