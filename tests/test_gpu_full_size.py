"""The three rendering workloads at their full sizes, checked by what does not depend on size.
The many-light workload (1,204,514 triangles, 4,096 emissive triangles through the light tree,
1024x512 sky, thin lens at 1920x1080, depth 16, hero spectra, corrected): per-path radiance of a
random sample of (pixel, sample) pairs against the CPU checker on the same scene, every pixel's
sample count, finite sums, and a render split into many waves (a small path pool) equal to one
wave. The media workload (128^3 grid, 65 lights, 1024x1024, depth 32) and the Cornell box
(256x256, 64 spp, depth 8): per-path radiance of sampled paths against the CPU checker.
Bar as the other parity tests: |gpu - cpu| <= 1e-4 relative + 1e-6 on 99.9 % of paths."""
import numpy as np
import pytest

from lambda_amd import abi, gpu, scene
from tests import helpers

pytestmark = pytest.mark.gpu

REL, ABS = 1e-4, 1e-6
W, H = 1920, 1080


@pytest.fixture(scope="module")
def workload():
    dev = gpu.Device(0)
    sd = scene.many_light_scene()
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    assert s.info()["triangles"] == 1204514
    yield dev, sd, s, scene.many_light_camera(W, H)
    del keep
    s.close()
    dev.close()


def _desc(s, film, cam, b, e, capacity=0):
    return gpu.render_desc(s, film, cam, b, e, max_bounces=16, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                           estimator=abi.ESTIMATOR_CORRECTED, capacity=capacity)


def test_full_size_per_path_parity(workload):
    dev, sd, s, cam = workload
    o, k = helpers.oracle_scene(sd)
    rng = np.random.default_rng(2026)
    n = 16384
    pixels = rng.integers(0, W * H, n).astype(np.uint32)
    samples = rng.integers(0, 256, n).astype(np.uint32)
    film = gpu.Film(dev, W, H)
    g = gpu.trace_paths(_desc(s, film, cam, 0, 256), pixels, samples)
    c = o.trace_paths(helpers.desc(cam, 0, 256, 16, 3, abi.SPECTRAL_HERO4, abi.ESTIMATOR_CORRECTED), W, H, pixels,
                      samples)
    assert np.isfinite(g).all()
    ok = helpers.rel_close(g[:, :4], c[:, :4], REL, ABS).all(axis=1) & (g[:, 4] == c[:, 4])
    print("matching", ok.mean())
    assert ok.mean() >= 0.999


def test_full_size_film_properties_and_wave_split(workload):
    dev, sd, s, cam = workload
    one, many = gpu.Film(dev, W, H), gpu.Film(dev, W, H)
    st1 = gpu.render(_desc(s, one, cam, 0, 2))
    st2 = gpu.render(_desc(s, many, cam, 0, 2, capacity=1 << 20))
    assert st1.waves == 1 and st2.waves == 4
    assert st1.extension_rays == st2.extension_rays and st1.shadow_rays == st2.shadow_rays
    a, ca = one.read()
    b, cb = many.read()
    assert (ca == 2).all() and (cb == 2).all()
    assert np.isfinite(a).all()
    # a pixel's two samples land in different waves of the small pool: same paths, same sums
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=0)
    assert a.mean() > 0


def _sampled_parity(dev, sd, cam, W, H, spp, integrator, max_bounces, n, seed):
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    o, k = helpers.oracle_scene(sd)
    rng = np.random.default_rng(seed)
    pixels = rng.integers(0, W * H, n).astype(np.uint32)
    samples = rng.integers(0, spp, n).astype(np.uint32)
    film = gpu.Film(dev, W, H)
    d = gpu.render_desc(s, film, cam, 0, spp, max_bounces=max_bounces, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                        estimator=abi.ESTIMATOR_CORRECTED, integrator=integrator)
    g = gpu.trace_paths(d, pixels, samples)
    cd = helpers.desc(cam, 0, spp, max_bounces, 3, abi.SPECTRAL_HERO4, abi.ESTIMATOR_CORRECTED)
    cd.integrator = integrator
    c = o.trace_paths(cd, W, H, pixels, samples)
    assert np.isfinite(g).all()
    ok = helpers.rel_close(g[:, :4], c[:, :4], REL, ABS).all(axis=1) & (g[:, 4] == c[:, 4])
    del keep
    s.close()
    return g, c, ok


def test_full_size_media_per_path_parity(workload):
    """The media workload at its full size: 128^3 fBm density grid in a pass-through box, a spotlight
    and 64 point lights, MIS distance + equiangular volumetric integrator, 1024x1024, depth 32."""
    dev = workload[0]
    g, c, ok = _sampled_parity(dev, scene.media_scene(res=128), scene.media_camera(1024, 1024), 1024, 1024, 128,
                               abi.INTEGRATOR_MIS_VOLPATH, 32, 8192, 7)
    print("matching", ok.mean(), "nonzero", (np.abs(c[:, :3]).sum(axis=1) > 0).mean())
    assert (np.abs(c[:, :3]).sum(axis=1) > 0).mean() > 0.2
    assert ok.mean() >= 0.999


def test_full_size_cornell_per_path_parity(workload):
    """The small configuration at its full size: Cornell box, 256x256, 64 spp, depth 8."""
    dev = workload[0]
    g, c, ok = _sampled_parity(dev, scene.cornell_box(), scene.cornell_camera(256, 256), 256, 256, 64,
                               abi.INTEGRATOR_PATH, 8, 16384, 11)
    print("matching", ok.mean())
    assert ok.mean() >= 0.999
