"""Closed-form pins of the checker's cameras (Camera.cpp:41-71, Aperture.cpp:10-27), the
environment light's sampling (EnvironmentLight.cpp:24-67, Piecewise.cpp:50-76) and the spotlight
falloff (Spotlight.cpp:21-42)."""
import math

import numpy as np
from scipy import stats

import oracle
from lambda_amd import abi, scene
from tests import helpers


def _basis(cam):
    return (np.array(cam.x_hat, dtype=np.float64), np.array(cam.y_hat, dtype=np.float64),
            np.array(cam.z_hat, dtype=np.float64))


def test_pinhole_centre_ray_is_z_hat_and_corners_follow_the_fov():
    """Image centre (u, v) = (1/2, 1/2) looks along zHat; image u maps to -xHat and v to -yHat with
    tan(fov/2) (aspect for y) at the edges (Camera.cpp:41-44)."""
    cam = scene.camera_pinhole((0.3, -0.2, 1.5), 320, 200, 2.6, 0.3, 0.9)
    x, y, z = _basis(cam)
    o, d = oracle.generate_rays(cam, [[0.5, 0.5], [0.0, 0.5], [1.0, 0.0]])
    np.testing.assert_allclose(o, np.tile(list(cam.origin), (3, 1)), atol=0)
    np.testing.assert_allclose(d[0], z, atol=2e-7)
    t = math.tan(0.45)
    e = z + x * t
    np.testing.assert_allclose(d[1], e / np.linalg.norm(e), atol=1e-6)
    e = z - x * t + y * t * (200 / 320)
    np.testing.assert_allclose(d[2], e / np.linalg.norm(e), atol=1e-6)


def test_thin_lens_with_zero_aperture_is_the_pinhole():
    """A circular aperture of size 0 gives the pinhole's ray for every image point (corrected
    centred disk; the as-written offset disk is also centred at size 0), and every lens ray of a
    nonzero aperture meets the pinhole ray on the focus plane (Camera.cpp:53-58)."""
    pin = scene.camera_pinhole((0.0, 0.4, 2.0), 64, 48, math.pi, -0.1, 0.8)
    rng = np.random.default_rng(1)
    uv = rng.uniform(size=(500, 2))
    _, dp = oracle.generate_rays(pin, uv)
    for corrected in (True, False):
        lens = scene.camera_thin_lens((0.0, 0.4, 2.0), 64, 48, math.pi, -0.1, 0.8, 2.5, 0.0)
        o, d = oracle.generate_rays(lens, uv, corrected=corrected)
        np.testing.assert_allclose(o, np.tile(list(pin.origin), (500, 1)), atol=0)
        np.testing.assert_allclose(d, dp, atol=2e-6)
    lens = scene.camera_thin_lens((0.0, 0.4, 2.0), 64, 48, math.pi, -0.1, 0.8, 2.5, 0.2)
    o, d = oracle.generate_rays(lens, uv)
    x, y, z = _basis(lens)
    origin = np.array(lens.origin, dtype=np.float64)
    # point on the focus plane along the pinhole ray: origin + dp * (f / (dp . z))
    fp = origin + dp * (2.5 / (dp @ z))[:, None]
    t = ((fp - o) @ z) / (d @ z)
    np.testing.assert_allclose(o + d * t[:, None], fp, atol=2e-5)
    off = o - origin
    assert np.abs(off @ z).max() < 1e-6  # lens points lie in the xHat/yHat plane
    assert (np.linalg.norm(off, axis=1) <= 0.2 + 1e-6).all()


def test_blade_aperture_points_lie_in_their_blade_parallelogram():
    """BladeAperture::Sample_p (Aperture.cpp:10-19): blade k spans angles [k, k + 1) 2 pi / n and the
    point p1 u0 + p2 u1 (u in [0, 1)^2, as written: the parallelogram, not the triangle) scaled by
    the size; every blade is drawn about equally often."""
    n, size = 6, 0.15
    cam = scene.camera_blades((0.0, 0.0, 0.0), 32, 32, math.pi, 0.0, 0.8, 3.0, size, n)
    x, y, z = _basis(cam)
    uv = np.full((30000, 2), 0.5)
    o, _ = oracle.generate_rays(cam, uv)
    ax, ay = o @ x / size, o @ y / size
    ang = np.mod(np.arctan2(ay, ax), 2 * np.pi)
    blade = np.floor(ang / (2 * np.pi / n)).astype(int)
    counts = np.bincount(blade, minlength=n)
    assert stats.chisquare(counts).pvalue > 1e-3
    t0 = blade * 2 * np.pi / n
    p1 = np.stack([np.cos(t0), np.sin(t0)], 1)
    p2 = np.stack([np.cos(t0 + 2 * np.pi / n), np.sin(t0 + 2 * np.pi / n)], 1)
    det = p1[:, 0] * p2[:, 1] - p1[:, 1] * p2[:, 0]
    u0 = (ax * p2[:, 1] - ay * p2[:, 0]) / det
    u1 = (p1[:, 0] * ay - p1[:, 1] * ax) / det
    assert (u0 >= -1e-5).all() and (u0 < 1 + 1e-5).all() and (u1 >= -1e-5).all() and (u1 < 1 + 1e-5).all()
    assert (u0 + u1 > 1).mean() > 0.4  # the parallelogram's far half is used (as written)


def test_spherical_camera_maps_uv_to_angles():
    """SphericalCamera (Camera.cpp:66-71): d = SphericalDirection(sin theta, cos theta, phi) in the
    camera basis with phi = 2 pi u + offset_phi, theta = pi v + offset_theta (y up in the basis)."""
    cam = scene.camera_spherical((0.1, 0.2, 0.3), 64, 32, phi=0.7, theta=0.2, offset_phi=0.4, offset_theta=0.1)
    x, y, z = _basis(cam)
    rng = np.random.default_rng(4)
    uv = rng.uniform(size=(400, 2))
    o, d = oracle.generate_rays(cam, uv)
    phi = 2 * np.pi * uv[:, 0] + 0.4
    th = np.pi * uv[:, 1] + 0.1
    loc = np.stack([np.sin(th) * np.cos(phi), np.cos(th), np.sin(th) * np.sin(phi)], 1)
    exp = loc[:, :1] * x + loc[:, 1:2] * y + loc[:, 2:3] * z
    np.testing.assert_allclose(d, exp, atol=2e-6)
    np.testing.assert_allclose(o, np.tile([0.1, 0.2, 0.3], (400, 1)), atol=1e-7)


def _env_scene(img):
    sd = scene.SceneDesc()
    sd.add_material(scene.lambertian((0.5, 0.5, 0.5)))
    sd.add_mesh(np.array([[50, 50, 50], [50.1, 50, 50], [50, 50.1, 50]], dtype=np.float32),
                np.array([[0, 1, 2]], dtype=np.uint32), 0)  # far away: nothing occludes the sky
    sd.add_env_light(img, intensity=1.0)
    return helpers.oracle_scene(sd, abi.LIGHTS_POWER)


def test_environment_sampling_follows_its_pdf():
    """EnvironmentLight::Sample_Li draws directions with density PDF_Li (the luminance sin(theta)
    Piecewise2D over the equirectangular map divided by 2 pi^2 sin(theta)): chi-square of the
    sampled directions against the integral of PDF_Li over (cos theta, phi) bins, and the integral
    of PDF_Li over the sphere is 1."""
    img = scene.sky(64, 32)
    o, keep = _env_scene(img)
    light = 0
    n = 100000
    wi, pdf, L = o.light_sample_li(light, (0, 0, 0), (0, 1, 0), n)
    ok = pdf > 0
    assert ok.mean() > 0.999
    nt, nph, sub = 32, 64, 8
    dc, dp = 2.0 / (nt * sub), 2 * np.pi / (nph * sub)
    c = -1 + dc * (np.arange(nt * sub) + 0.5)
    p = dp * (np.arange(nph * sub) + 0.5)
    C_, P_ = np.meshgrid(c, p, indexing="ij")
    s = np.sqrt(1 - C_ ** 2)
    w = np.stack([s * np.cos(P_), C_, s * np.sin(P_)], -1).reshape(-1, 3)
    pe = o.env_pdf(light, w).astype(np.float64)
    assert abs(pe.sum() * dc * dp - 1) < 5e-3
    it = (np.arange(nt * sub) // sub)[:, None].repeat(nph * sub, 1).reshape(-1)
    ip = (np.arange(nph * sub) // sub)[None, :].repeat(nt * sub, 0).reshape(-1)
    exp = np.bincount(it * nph + ip, weights=pe * dc * dp, minlength=nt * nph) * ok.sum()
    ct = np.clip(wi[ok, 1], -1, 1)
    ph = np.mod(np.arctan2(wi[ok, 2], wi[ok, 0]), 2 * np.pi)
    b = np.minimum(((ct + 1) / 2 * nt).astype(int), nt - 1) * nph + np.minimum((ph / (2 * np.pi) * nph).astype(int), nph - 1)
    obs = np.bincount(b, minlength=nt * nph)
    big = exp >= 5
    o_, e_ = obs[big].astype(np.float64), exp[big]
    if exp[~big].sum() > 0:
        o_, e_ = np.append(o_, obs[~big].sum()), np.append(e_, exp[~big].sum())
    e_ *= o_.sum() / e_.sum()
    assert stats.chisquare(o_, e_).pvalue > 1e-3
    # the returned pdf equals PDF_Li of the sampled direction (sample/pdf consistency)
    np.testing.assert_allclose(pdf[ok][:2000], o.env_pdf(light, wi[ok][:2000]), rtol=2e-3)


def test_spotlight_falloff_closed_form():
    """Spotlight::Falloff (Spotlight.cpp:21-27): 1 inside the falloff cone, ((cos - cos cone) /
    (cos falloff - cos cone))^4 between, 0 outside; Li = I * falloff / d^2 with pdf 1."""
    sd = scene.SceneDesc()
    sd.add_material(scene.lambertian((0.5, 0.5, 0.5)))
    sd.add_mesh(np.array([[50, 50, 50], [50.1, 50, 50], [50, 50.1, 50]], dtype=np.float32),
                np.array([[0, 1, 2]], dtype=np.uint32), 0)
    cone, fall = 0.6, 0.35
    sd.add_spot_light((0.0, 2.0, 0.0), (0.0, -1.0, 0.0), cone, fall, 3.0, rgb=(1, 1, 1))
    o, keep = helpers.oracle_scene(sd, abi.LIGHTS_POWER)
    for ang in (0.0, 0.2, 0.34, 0.4, 0.5, 0.58, 0.62, 0.9):
        r = 1.5
        p = (r * math.sin(ang), 2.0 - r * math.cos(ang), 0.0)
        wi, pdf, L = o.light_sample_li(0, p, (0.0, 1.0, 0.0), 1)
        c = math.cos(ang)
        cc, cf = math.cos(cone), math.cos(fall)
        f = 1.0 if c > cf else (0.0 if c < cc else ((c - cc) / (cf - cc)) ** 4)
        exp = 3.0 * f / (r * r)
        assert pdf[0] == 1.0
        # the shadow ray leaves from the point offset 1e-5 along the normal: about 1e-5 in the cosine
        assert abs(L[0, 0] - exp) <= 1e-4 * max(1.0, exp), (ang, L[0, 0], exp)
