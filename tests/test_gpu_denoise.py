"""Device denoiser (lgpu_film_denoise, the edge-avoiding A-Trous filter standing in for OIDN's
filter of Denoise::SetData/Process, Denoise.cpp:13-67) against the CPU checker on the same film
means, and the film write/read round trip it builds on.

Bar: |gpu - cpu| <= 1e-4 * |cpu| + 1e-6 per channel (device expf against glibc's, summed over
25 taps and several passes)."""
import numpy as np
import pytest

import oracle
from lambda_amd import abi, gpu, scene
from lambda_amd.gpu import LambdaError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def _means(film):
    rgb, cnt = film.read()
    inv = np.where(cnt > 0, 1.0 / np.maximum(cnt, 1), 0.0)[..., None]
    return (rgb * inv).astype(np.float32)


def _render_films(dev, W, H):
    sd = scene.cornell_box()
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    cam = scene.cornell_camera(W, H)
    colour, normal = gpu.Film(dev, W, H), gpu.Film(dev, W, H)
    gpu.render(gpu.render_desc(s, colour, cam, 0, 4, spectral=abi.SPECTRAL_HERO4))
    gpu.render(gpu.render_desc(s, normal, cam, 0, 1, integrator=abi.INTEGRATOR_NORMAL))
    return s, keep, colour, normal


def test_film_write_read_round_trip(dev):
    W, H = 13, 9
    f = gpu.Film(dev, W, H)
    rng = np.random.default_rng(2)
    rgb = rng.random((H, W, 3))
    cnt = rng.integers(0, 5, (H, W)).astype(np.uint32)
    f.write(rgb, cnt)
    r, c = f.read()
    np.testing.assert_array_equal(r, rgb)
    np.testing.assert_array_equal(c, cnt)
    f.write(rgb)
    assert (f.read()[1] == 1).all()


@pytest.mark.parametrize("guides", ["none", "normal", "albedo+normal"])
def test_denoise_matches_oracle(dev, guides):
    W, H = 48, 40
    s, keep, colour, normal = _render_films(dev, W, H)
    albedo = None
    alb = None
    if "albedo" in guides:
        # an albedo image with texture: a checker of two reflectances
        yy, xx = np.mgrid[0:H, 0:W]
        alb = np.where(((xx // 6 + yy // 6) % 2 == 0)[..., None], [0.7, 0.6, 0.5], [0.2, 0.3, 0.25])
        albedo = gpu.Film(dev, W, H)
        albedo.write(alb)
    nrm = normal if "normal" in guides else None
    d = abi.denoise_desc(iterations=5, sigma_colour=0.8, sigma_normal=0.2, sigma_albedo=0.1, demodulate=True)
    out = gpu.Film(dev, W, H)
    colour.denoise(d, albedo=albedo, normal=nrm, out=out)
    g = _means(out)
    c = oracle.denoise(_means(colour), d, albedo=None if alb is None else alb.astype(np.float32),
                       normal=None if nrm is None else _means(normal))
    assert (out.read()[1] == 1).all()
    np.testing.assert_allclose(g, c, rtol=1e-4, atol=1e-6)
    if albedo is None:  # the filter smooths (with an albedo the checker texture is put back)
        src = _means(colour)
        assert np.abs(np.diff(g, axis=1)).mean() < 0.7 * np.abs(np.diff(src, axis=1)).mean()


def test_denoise_in_place_equals_out_of_place(dev):
    W, H = 32, 24
    s, keep, colour, normal = _render_films(dev, W, H)
    d = abi.denoise_desc(iterations=3, sigma_colour=0.5, sigma_normal=0.3)
    out = gpu.Film(dev, W, H)
    colour.denoise(d, normal=normal, out=out)
    colour.denoise(d, normal=normal)
    np.testing.assert_array_equal(colour.read()[0], out.read()[0])


def test_denoise_rejects_bad_arguments(dev):
    a, b = gpu.Film(dev, 8, 8), gpu.Film(dev, 9, 8)
    with pytest.raises(LambdaError):
        a.denoise(abi.denoise_desc(iterations=0))
    with pytest.raises(LambdaError):
        a.denoise(abi.denoise_desc(iterations=13))
    with pytest.raises(LambdaError):
        a.denoise(abi.denoise_desc(sigma_colour=-1.0))
    with pytest.raises(LambdaError):
        a.denoise(abi.denoise_desc(), normal=b)
