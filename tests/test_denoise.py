"""The oracle's denoiser (edge-avoiding A-Trous wavelet filter, the library's stand-in for the
OIDN filter of Denoise::SetData/Process, Denoise.cpp:13-67) pinned by what the method fixes:
weights that renormalise, the plain B3-spline a-trous blur when every edge-stopping term is off
(checked against an independent numpy convolution), edge stopping across a normal edge, and
albedo demodulation that keeps texture detail. OIDN itself is absent: equivalence to its trained
filter is unpinned."""
import numpy as np
import pytest

import oracle
from lambda_amd import abi

B3 = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])


def atrous_blur(img, iterations):
    """Plain a-trous B3-spline smoothing, taps outside the image dropped and the kernel
    renormalised (numpy, float64)."""
    c = img.astype(np.float64)
    h, w = c.shape[:2]
    for it in range(iterations):
        step = 1 << it
        acc = np.zeros_like(c)
        wsum = np.zeros((h, w, 1))
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                k = B3[dy + 2] * B3[dx + 2]
                ys, xs = np.arange(h) + dy * step, np.arange(w) + dx * step
                vy, vx = (ys >= 0) & (ys < h), (xs >= 0) & (xs < w)
                m = np.outer(vy, vx)[..., None]
                shifted = c[np.clip(ys, 0, h - 1)][:, np.clip(xs, 0, w - 1)]
                acc += k * m * shifted
                wsum += k * m
        c = acc / wsum
    return c


def test_constant_image_is_unchanged():
    c = np.full((20, 24, 3), 0.37, np.float32)
    nrm = np.random.default_rng(1).random((20, 24, 3)).astype(np.float32)
    out = oracle.denoise(c, abi.denoise_desc(iterations=5, sigma_colour=0.5, sigma_normal=0.2), normal=nrm)
    np.testing.assert_allclose(out, 0.37, rtol=1e-6)


@pytest.mark.parametrize("iterations", [1, 2, 4])
def test_no_edge_terms_is_the_b3_atrous_blur(iterations):
    rng = np.random.default_rng(iterations)
    c = rng.random((23, 31, 3)).astype(np.float32)
    out = oracle.denoise(c, abi.denoise_desc(iterations=iterations, sigma_colour=0.0, sigma_normal=0.0,
                                            sigma_albedo=0.0, demodulate=False))
    np.testing.assert_allclose(out, atrous_blur(c, iterations), rtol=2e-5, atol=1e-6)


def test_normal_edge_stops_the_filter():
    h, w = 16, 32
    nrm = np.zeros((h, w, 3), np.float32)
    nrm[:, : w // 2] = (1, 0, 0)
    nrm[:, w // 2:] = (0, 1, 0)
    rng = np.random.default_rng(7)
    c = np.zeros((h, w, 3), np.float32)
    c[:, : w // 2] = 1.0 + 0.2 * rng.standard_normal((h, w // 2, 3))
    out = oracle.denoise(c, abi.denoise_desc(iterations=5, sigma_colour=0.0, sigma_normal=0.1), normal=nrm)
    # |dn|^2 = 2 across the edge: weight exp(-200), nothing leaks to the dark side
    assert np.abs(out[:, w // 2:]).max() < 1e-30
    # the bright side is smoothed towards its mean
    assert out[:, : w // 2].std() < 0.5 * c[:, : w // 2].std()
    np.testing.assert_allclose(out[:, : w // 2].mean(), c[:, : w // 2].mean(), rtol=0.02)


def test_unit_albedo_demodulation_is_identity():
    rng = np.random.default_rng(3)
    c = rng.random((12, 14, 3)).astype(np.float32)
    ones = np.ones_like(c)
    d1 = abi.denoise_desc(iterations=3, sigma_colour=0.4, sigma_albedo=0.1, demodulate=True)
    d0 = abi.denoise_desc(iterations=3, sigma_colour=0.4, sigma_albedo=0.1, demodulate=False)
    np.testing.assert_array_equal(oracle.denoise(c, d1, albedo=ones), oracle.denoise(c, d0, albedo=ones))


def test_demodulation_keeps_texture():
    # colour = albedo texture x constant illumination: demodulated it is constant, so the
    # filter returns the texture untouched; without demodulation the texture is blurred
    rng = np.random.default_rng(5)
    alb = (0.2 + 0.7 * rng.random((24, 24, 3))).astype(np.float32)
    c = (alb * 0.8).astype(np.float32)
    d = abi.denoise_desc(iterations=4, sigma_colour=0.5, sigma_albedo=0.0, demodulate=True)
    np.testing.assert_allclose(oracle.denoise(c, d, albedo=alb), c, rtol=1e-5)
    d.demodulate = 0
    blurred = oracle.denoise(c, d, albedo=alb)
    assert np.abs(blurred - c).max() > 0.05


def test_noise_is_reduced():
    rng = np.random.default_rng(11)
    c = (0.5 + 0.1 * rng.standard_normal((64, 64, 3))).astype(np.float32)
    out = oracle.denoise(c, abi.denoise_desc(iterations=5, sigma_colour=1.0))
    assert out.var() < c.var() / 20
