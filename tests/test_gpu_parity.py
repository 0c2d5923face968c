"""GPU parity: the HIP path (through the C ABI) against the CPU checker on the
same seeded inputs.

Bars:
  - integer / index work (Philox words, hit primitive ids, light-tree structure): bit-exact;
  - per-path radiance (FP32): |gpu - cpu| <= 1e-4 * max(|gpu|, |cpu|) + 1e-6 for at least
    99.9% of paths (the remainder are paths whose discrete decisions flip on a
    transcendental ulp between the device math library and glibc);
  - converged images: per-channel mean within 1e-3 relative.
"""
import math

import numpy as np
import pytest

import oracle
from lambda_amd import abi, gpu, scene
from tests import helpers

pytestmark = pytest.mark.gpu

REL, ABS = 1e-4, 1e-6


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def gpu_scene(dev, sd, strategy=abi.LIGHTS_TREE):
    s = gpu.Scene(dev)
    keep = sd.push(s, strategy)
    return s, keep


def test_philox_bit_exact(dev):
    rng = np.random.default_rng(3)
    ctr = rng.integers(0, 2 ** 32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    key = np.array([0x12345678, 0x9ABCDEF0], dtype=np.uint32)
    out = dev.philox(ctr, key)
    for i in range(0, 4096, 97):
        assert list(out[i]) == list(oracle.philox(ctr[i], key))
    for c, k, e in [([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8])]:
        assert list(dev.philox(np.array([c], dtype=np.uint32), np.array(k, dtype=np.uint32))[0]) == e


def _soup(n_tris, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-4, 4, size=(n_tris, 1, 3))
    v = c + rng.normal(scale=0.15, size=(n_tris, 3, 3))
    sd = scene.SceneDesc()
    sd.add_material(scene.lambertian((0.5, 0.5, 0.5)))
    sd.add_mesh(v.reshape(-1, 3), np.arange(3 * n_tris).reshape(-1, 3), 0)
    return sd


def _rays(n, seed, lo=-4.5, hi=4.5, tmax=np.inf):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3] = 0.0
    r[:, 7] = tmax
    return r


@pytest.mark.parametrize("which", ["cornell", "soup"])
def test_intersect_matches_oracle(dev, which):
    sd = scene.cornell_box() if which == "cornell" else _soup(60000, 11)
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    rays = _rays(50000, 5, -0.95, 0.95) if which == "cornell" else _rays(50000, 5)
    if which == "soup":  # a third of the rays with a finite tmax
        rays[::3, 7] = 2.0
    g = s.intersect(rays)
    c = o.intersect(rays)
    assert (g[:, 3] == c[:, 3]).all()
    hit = c[:, 3] != 0xFFFFFFFF
    assert hit.sum() > 1000
    assert (g[hit, :3] == c[hit, :3]).all()  # t, u, v bit-identical


def test_occluded_matches_oracle(dev):
    sd = _soup(20000, 12)
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    rays = _rays(40000, 6, tmax=3.0)
    assert (s.occluded(rays) == o.occluded(rays)).all()


def _many_light_scene(n_quads=48, seed=3):
    rng = np.random.default_rng(seed)
    sd = scene.cornell_box()
    for q in range(n_quads):
        m = sd.add_material(scene.emitter_blackbody(float(rng.uniform(2000, 9000)), float(10 ** rng.uniform(0, 2))))
        c = rng.uniform(-0.8, 0.8, size=3)
        s = 0.05
        v = np.array([[c[0] - s, c[1], c[2] - s], [c[0] + s, c[1], c[2] - s], [c[0] + s, c[1], c[2] + s],
                      [c[0] - s, c[1], c[2] + s]], dtype=np.float32)
        sd.add_mesh(v, np.array([[0, 2, 1], [0, 3, 2]], dtype=np.uint32), m)
    sd.add_point_light((0.3, 0.5, 0.2), 2.0, temperature=4000.0)
    sd.add_spot_light((-0.3, 0.8, 0.0), (0.0, -1.0, 0.0), 0.6, 0.4, 3.0, rgb=(1.0, 0.9, 0.8))
    return sd


@pytest.mark.parametrize("which", ["cornell", "many"])
def test_light_tree_matches_oracle(dev, which):
    sd = scene.cornell_box() if which == "cornell" else _many_light_scene()
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    tg, tc = s.light_tree(), o.light_tree()
    assert tg.shape == tc.shape and len(tg) >= 1
    assert (tg == tc).all()


MODES = [(sp, est, strat) for sp in (abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4)
         for est in (abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED)
         for strat in (abi.LIGHTS_TREE, abi.LIGHTS_POWER)]


def _compare_paths(g, c):
    ok = helpers.rel_close(g[:, :4], c[:, :4], REL, ABS).all(axis=1) & (g[:, 4] == c[:, 4])
    return ok


@pytest.mark.parametrize("spectral,estimator,strategy", MODES)
def test_cornell_per_path_parity(dev, spectral, estimator, strategy):
    W = H = 64
    sd = scene.cornell_box()
    s, k1 = gpu_scene(dev, sd, strategy)
    o, k2 = helpers.oracle_scene(sd, strategy)
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, 2, max_bounces=8, min_bounces=3, spectral=spectral, estimator=estimator)
    g = gpu.trace_paths(d, pixels, samples)
    c = o.trace_paths(helpers.desc(cam, 0, 2, 8, 3, spectral, estimator), W, H, pixels, samples)
    ok = _compare_paths(g, c)
    print("paths", len(ok), "matching", ok.mean(), "nonzero", (np.abs(c[:, :3]).sum(axis=1) > 0).mean())
    assert ok.mean() >= 0.999
    np.testing.assert_allclose(g[:, :4].mean(axis=0), c[:, :4].mean(axis=0), rtol=1e-3, atol=1e-7)


def test_many_light_per_path_parity(dev):
    W = H = 48
    sd = _many_light_scene()
    for strategy in (abi.LIGHTS_TREE, abi.LIGHTS_POWER):
        for est in (abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED):
            s, k1 = gpu_scene(dev, sd, strategy)
            o, k2 = helpers.oracle_scene(sd, strategy)
            cam = scene.cornell_camera(W, H)
            film = gpu.Film(dev, W, H)
            pixels = np.arange(W * H, dtype=np.uint32)
            samples = np.zeros(W * H, dtype=np.uint32)
            d = gpu.render_desc(s, film, cam, 0, 1, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                                estimator=est)
            g = gpu.trace_paths(d, pixels, samples)
            c = o.trace_paths(helpers.desc(cam, 0, 1, 8, 3, abi.SPECTRAL_HERO4, est), W, H, pixels, samples)
            ok = _compare_paths(g, c)
            print(strategy, est, "matching", ok.mean())
            assert ok.mean() >= 0.999


@pytest.mark.parametrize("estimator", [abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED])
@pytest.mark.parametrize("spectral", [abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4])
def test_glossy_many_light_per_path_parity(dev, spectral, estimator):
    # GGX conductor, rough dielectric and Oren-Nayar spheres with smooth normals,
    # emissive quads through the light tree, HDR sky, thin-lens camera
    sd = scene.many_light_scene(n_spheres=40, sphere_level=2, ground_res=16, n_light_quads=64, seed=5)
    W, H = 96, 54
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    cam = scene.many_light_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, 2, max_bounces=16, min_bounces=3, spectral=spectral, estimator=estimator)
    g = gpu.trace_paths(d, pixels, samples)
    c = o.trace_paths(helpers.desc(cam, 0, 2, 16, 3, spectral, estimator), W, H, pixels, samples)
    ok = _compare_paths(g, c)
    print("matching", ok.mean(), "finite", np.isfinite(g).all())
    assert np.isfinite(g).all()
    assert ok.mean() >= 0.999


def test_film_render_matches_oracle(dev):
    W = H = 64
    spp = 8
    sd = scene.cornell_box()
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    d = gpu.render_desc(s, film, cam, 0, spp, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                        estimator=abi.ESTIMATOR_CORRECTED, capacity=10000)
    st = gpu.render(d)
    assert st.samples == W * H * spp and st.waves == (W * H * spp + 9999) // 10000
    rgb, cnt = film.read()
    assert (cnt == spp).all()
    crgb, ccnt = o.render(helpers.desc(cam, 0, spp, 8, 3, abi.SPECTRAL_HERO4, abi.ESTIMATOR_CORRECTED), W, H)
    np.testing.assert_allclose(rgb.mean(axis=(0, 1)), crgb.mean(axis=(0, 1)), rtol=1e-3)
    close = np.abs(rgb - crgb) <= 1e-3 * np.maximum(np.abs(rgb), np.abs(crgb)) + 1e-6
    assert close.mean() >= 0.99
    rgba = film.read_rgba()
    np.testing.assert_allclose(rgba[..., :3], rgb / spp, rtol=1e-6, atol=1e-7)


# CIE 1931 analytic fit (Wyman, Sloan & Shirley 2013): (weight, mu, sigma below, sigma above) per lobe
_CMF_LOBES = (((1.056, 599.8, 37.9, 31.0), (0.362, 442.0, 16.0, 26.7), (-0.065, 501.1, 20.4, 26.2)),
              ((0.821, 568.8, 46.9, 40.5), (0.286, 530.9, 16.3, 31.1)),
              ((1.217, 437.0, 11.8, 36.0), (0.681, 459.0, 26.0, 13.8)))
_XYZ_TO_RGB = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                        [0.055648, -0.204043, 1.057311]])


def _cmf_f32(lam):
    """The fit as the CPU checker evaluates it: float32 division and expf per lobe."""
    lam = lam.astype(np.float32)
    out = []
    for lobes in _CMF_LOBES:
        acc = np.zeros_like(lam)
        for w, mu, s1, s2 in lobes:
            t = (lam - np.float32(mu)) / np.where(lam < np.float32(mu), np.float32(s1), np.float32(s2))
            acc = acc + np.float32(w) * np.exp(np.float32(-0.5) * t * t)
        out.append(acc)
    return np.stack(out, axis=-1)


def test_hero_film_update_matches_restated_colour_matching(dev):
    """The device film update evaluates the fit with the hardware exp2 and the widths' reciprocals.
    Per lobe its value differs from expf(-t^2/2) with a division by at most about
    (t^2 + 1) e^(-t^2/2) 2^-23 <= 1.21 x 2^-23 times the lobe weight. With the restatement's own
    expf rounding (1 ulp) and the lobe sums' roundings on both sides (2 ulps), each colour-matching
    value is within 4.21 x 2^-23 x (sum of |weights|) of the restatement; the XYZ and RGB bounds
    follow from the sample's |L| and |XYZ -> RGB| (FP64 from there on). Bar: every pixel of a 1-spp
    hero film (one sample per pixel: the film value is that sample's update) within twice that
    bound, over 360-830 nm. A 1 % error in one lobe's width exceeds it by orders of magnitude."""
    W = H = 64
    sd = scene.cornell_box()
    s, k1 = gpu_scene(dev, sd)
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    d = gpu.render_desc(s, film, cam, 0, 1, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                        estimator=abi.ESTIMATOR_CORRECTED, seed=11)
    gpu.render(d)
    rgb, cnt = film.read()
    assert (cnt == 1).all()
    pix = np.arange(W * H, dtype=np.uint32)
    g = gpu.trace_paths(d, pix, np.zeros(W * H, dtype=np.uint32))
    L, l0 = g[:, :4].astype(np.float64), g[:, 4]
    lams = np.stack([l0 + np.float32(117.5) * np.float32(k) for k in range(4)], axis=1).astype(np.float32)
    lams = np.where(lams >= 830.0, lams - np.float32(470.0), lams).astype(np.float32)
    assert lams.min() >= 360.0 and lams.max() < 830.0
    cm = _cmf_f32(lams).astype(np.float64)                      # (n, 4, 3)
    scale = 470.0 / 4.0 / 106.856895
    xyz = (L[:, :, None] * cm).sum(axis=1) * scale
    ref = xyz @ _XYZ_TO_RGB.T
    wsum = np.array([sum(abs(l[0]) for l in lobes) for lobes in _CMF_LOBES])
    dxyz = np.abs(L).sum(axis=1)[:, None] * wsum[None, :] * 4.21 * 2.0 ** -23 * scale
    tol = 2.0 * dxyz @ np.abs(_XYZ_TO_RGB).T + 1e-12
    got = rgb.reshape(-1, 3)
    err = np.abs(got - ref)
    print("max err / bound %.3f, max rel err %.2e, lambda %.1f..%.1f" % (
        (err / tol).max(), (err / np.maximum(np.abs(ref), 1e-30)).max(), lams.min(), lams.max()))
    assert (err <= tol).all()
    film.close()


def test_render_is_resumable_by_sample_range(dev):
    # spp [0,4) then [4,8) into one film equals [0,8) in one call (Philox keyed on the sample index)
    W = H = 32
    sd = scene.cornell_box()
    s, k = gpu_scene(dev, sd)
    cam = scene.cornell_camera(W, H)
    f1, f2 = gpu.Film(dev, W, H), gpu.Film(dev, W, H)
    gpu.render(gpu.render_desc(s, f1, cam, 0, 8))
    gpu.render(gpu.render_desc(s, f2, cam, 0, 4))
    gpu.render(gpu.render_desc(s, f2, cam, 4, 8))
    a, ca = f1.read()
    b, cb = f2.read()
    assert (ca == cb).all()
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)


def test_furnace_on_gpu(dev):
    rho = 0.5
    sd = helpers.furnace_box(albedo=rho, emit=1.0, subdivide=2)
    s, k = gpu_scene(dev, sd)
    cam = scene.camera_pinhole((0.1, -0.2, 0.05), 64, 64, 0.3, 0.2, 1.2)
    film = gpu.Film(dev, 64, 64)
    gpu.render(gpu.render_desc(s, film, cam, 0, 64, max_bounces=64, estimator=abi.ESTIMATOR_CORRECTED))
    rgb, cnt = film.read()
    np.testing.assert_allclose((rgb / cnt[..., None]).mean(axis=(0, 1)), 1 / (1 - rho), rtol=0.01)


def test_invalid_render_arguments(dev):
    sd = scene.cornell_box()
    s, k = gpu_scene(dev, sd)
    film = gpu.Film(dev, 8, 8)
    cam = scene.cornell_camera(8, 8)
    d = gpu.render_desc(s, film, cam, 4, 4)
    with pytest.raises(gpu.LambdaError):
        gpu.render(d)
    d = gpu.render_desc(s, film, cam, 0, 1, max_bounces=0)
    with pytest.raises(gpu.LambdaError):
        gpu.render(d)
    d = gpu.render_desc(s, film, cam, 0, 1)
    with pytest.raises(gpu.LambdaError):
        gpu.trace_paths(d, np.array([64], dtype=np.uint32), np.array([0], dtype=np.uint32))


@pytest.mark.parametrize("spectral,estimator,strategy", MODES)
def test_direct_lighting_per_path_parity(dev, spectral, estimator, strategy):
    """DirectLightingIntegrator (SampleOneLight / EstimateDirect) on the Cornell box."""
    W = H = 64
    sd = scene.cornell_box()
    s, k1 = gpu_scene(dev, sd, strategy)
    o, k2 = helpers.oracle_scene(sd, strategy)
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, 2, spectral=spectral, estimator=estimator, integrator=abi.INTEGRATOR_DIRECT)
    g = gpu.trace_paths(d, pixels, samples)
    cd = helpers.desc(cam, 0, 2, 8, 3, spectral, estimator)
    cd.integrator = abi.INTEGRATOR_DIRECT
    c = o.trace_paths(cd, W, H, pixels, samples)
    ok = _compare_paths(g, c)
    print("direct matching", ok.mean(), "nonzero", (np.abs(c[:, :3]).sum(axis=1) > 0).mean())
    assert ok.mean() >= 0.999


@pytest.mark.parametrize("estimator", [abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED])
def test_direct_lighting_glossy_parity(dev, estimator):
    sd = scene.many_light_scene(n_spheres=40, sphere_level=2, ground_res=16, n_light_quads=64, seed=5)
    W, H = 96, 54
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    cam = scene.many_light_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, 2, spectral=abi.SPECTRAL_HERO4, estimator=estimator,
                        integrator=abi.INTEGRATOR_DIRECT)
    g = gpu.trace_paths(d, pixels, samples)
    cd = helpers.desc(cam, 0, 2, 8, 3, abi.SPECTRAL_HERO4, estimator)
    cd.integrator = abi.INTEGRATOR_DIRECT
    c = o.trace_paths(cd, W, H, pixels, samples)
    ok = _compare_paths(g, c)
    print("direct glossy matching", ok.mean())
    assert np.isfinite(g).all()
    assert ok.mean() >= 0.999


def test_direct_lighting_furnace_on_gpu(dev):
    rho = 0.5
    sd = helpers.furnace_box(albedo=rho, emit=1.0, subdivide=2)
    s, k = gpu_scene(dev, sd)
    cam = scene.camera_pinhole((0.1, -0.2, 0.05), 64, 64, 0.3, 0.2, 1.2)
    film = gpu.Film(dev, 64, 64)
    gpu.render(gpu.render_desc(s, film, cam, 0, 64, estimator=abi.ESTIMATOR_CORRECTED,
                               integrator=abi.INTEGRATOR_DIRECT))
    rgb, cnt = film.read()
    np.testing.assert_allclose((rgb / cnt[..., None]).mean(axis=(0, 1)), 1 + rho, rtol=0.01)


PASSES = [(i, sp, e) for i in (abi.INTEGRATOR_DEPTH, abi.INTEGRATOR_NORMAL, abi.INTEGRATOR_AOV)
          for sp in (abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4)
          for e in (abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED)]


@pytest.mark.parametrize("integrator,spectral,estimator", PASSES)
def test_utility_pass_per_path_parity(dev, integrator, spectral, estimator):
    """DepthPass / NormalPass / AOVPass on the glossy many-light scene (smooth normals, sky misses)."""
    sd = scene.many_light_scene(n_spheres=20, sphere_level=2, ground_res=8, n_light_quads=16, seed=3)
    W, H = 64, 36
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    cam = scene.many_light_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.arange(W * H, dtype=np.uint32)
    samples = np.zeros(W * H, dtype=np.uint32)
    for geometric in (False, True):
        d = gpu.render_desc(s, film, cam, 0, 1, spectral=spectral, estimator=estimator, integrator=integrator,
                            max_depth=12.0, depth_power=1.5, background=(0.1, 0.2, 0.3), geometric_normal=geometric)
        g = gpu.trace_paths(d, pixels, samples)
        cd = helpers.desc(cam, 0, 1, 8, 3, spectral, estimator)
        cd.integrator = integrator
        cd.pass_max_depth, cd.pass_power = 12.0, 1.5
        cd.pass_background[:] = [0.1, 0.2, 0.3]
        cd.pass_geometric_normal = 1 if geometric else 0
        c = o.trace_paths(cd, W, H, pixels, samples)
        ok = _compare_paths(g, c)
        print(integrator, spectral, estimator, geometric, "matching", ok.mean())
        assert ok.mean() >= 0.999


@pytest.mark.parametrize("estimator", [abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED])
def test_instanced_scene_per_path_parity(dev, estimator):
    """Instances (InstanceProxy/Instance) of a smooth sphere under rotations, non-uniform scales and
    mirrors with proxy, override and emissive materials, plus instanced quad lights."""
    sd = scene.instanced_cornell()
    W = H = 48
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    assert (s.light_tree() == o.light_tree()).all()
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, 2, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                        estimator=estimator)
    g = gpu.trace_paths(d, pixels, samples)
    c = o.trace_paths(helpers.desc(cam, 0, 2, 8, 3, abi.SPECTRAL_HERO4, estimator), W, H, pixels, samples)
    ok = _compare_paths(g, c)
    print("instanced", estimator, "matching", ok.mean())
    assert ok.mean() >= 0.999


def _write_obj(path, level=2):
    """An OBJ with a smooth icosphere (vt, vn) as object 'Ball' and a flat quad without normals."""
    v, f = scene.icosphere(level)
    n = scene._vertex_normals(v, f)
    uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi), 0.5 + v[:, 1] * 0.5], axis=1)
    lines = ["mtllib ball.mtl", "o Ball"]
    lines += ["v %.6f %.6f %.6f" % tuple(0.3 * p + np.array([0.2, -0.6, 0.1])) for p in v]
    lines += ["vt %.6f %.6f" % tuple(t) for t in uv]
    lines += ["vn %.6f %.6f %.6f" % tuple(x) for x in n]
    lines += ["usemtl Red"] + ["f %d/%d/%d %d/%d/%d %d/%d/%d" % (a + 1, a + 1, a + 1, b + 1, b + 1, b + 1, c + 1, c + 1, c + 1)
                               for a, b, c in f]
    k = len(v)
    lines += ["o Card", "v -0.8 -0.9 -0.6", "v -0.2 -0.9 -0.6", "v -0.2 -0.1 -0.7", "v -0.8 -0.1 -0.7",
              "usemtl Blue", "f %d %d %d %d" % (k + 1, k + 2, k + 3, k + 4)]
    path.with_name("ball.mtl").write_text("newmtl Red\nKd 0.7 0.2 0.2\nnewmtl Blue\nKd 0.2 0.3 0.8\n")
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("estimator", [abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED])
def test_obj_scene_per_path_parity(dev, tmp_path, estimator):
    """Meshes ingested by the library's OBJ loader (smooth sphere with tangents, flat quad) in the
    Cornell room, once directly and once as instances."""
    p = tmp_path / "ball.obj"
    _write_obj(p)
    sd = scene.cornell_box()
    sd.add_obj(p)
    sd.add_obj(p, transform=scene.affine((0, 0.7, 0), (0.8, 1.2, 0.8), (0.1, 0.5, -0.3)))
    W = H = 48
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, 2, spectral=abi.SPECTRAL_HERO4, estimator=estimator)
    g = gpu.trace_paths(d, pixels, samples)
    c = o.trace_paths(helpers.desc(cam, 0, 2, 8, 3, abi.SPECTRAL_HERO4, estimator), W, H, pixels, samples)
    ok = _compare_paths(g, c)
    print("obj", estimator, "matching", ok.mean())
    assert ok.mean() >= 0.999


@pytest.mark.parametrize("estimator", [abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED])
def test_film_post_matches_oracle(dev, tmp_path, estimator):
    """Device post-processing (convolution, ACES fit, clamp, gamma, 8-bit packing) of a rendered film
    against the CPU checker applied to the same film means; PNG round trip."""
    import oracle
    W, H = 48, 40
    s, k1 = gpu_scene(dev, scene.cornell_box())
    film = gpu.Film(dev, W, H)
    gpu.render(gpu.render_desc(s, film, scene.cornell_camera(W, H), 0, 8, spectral=abi.SPECTRAL_HERO4))
    means = film.read_rgba()
    kern = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], dtype=np.float32)
    for kernel in (None, kern):
        d, keep = abi.post_desc(estimator=estimator, kernel=kernel, intensity=1.0 / 16)
        g32, g8 = film.post(d)
        c32, c8 = oracle.post(means, d)
        np.testing.assert_allclose(g32, c32, rtol=1e-6, atol=1e-7)
        diff = np.abs(g8.astype(int) - c8.astype(int))
        assert diff.max() <= 1 and (diff == 0).mean() >= 0.999
    p = tmp_path / "film.png"
    gpu.write_png(p, g8)
    assert p.stat().st_size > W * H * 4


def _program_material():
    """A material whose albedo is a procedural texture: its sockets run shader-VM code."""
    return scene.Material([scene.node(abi.NODE_PERLIN, p=[5]), scene.node(abi.NODE_PERLIN, p=[7]),
                           scene.node(abi.NODE_PERLIN, p=[3]), scene.node(abi.NODE_MERGE_RGB, [0, 1, 2]),
                           scene.node(abi.NODE_LAMBERTIAN, [3])], bxdf=4)


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_PATH, abi.INTEGRATOR_MIS_VOLPATH])
def test_vm_free_kernels_match_vm_kernels(dev, integrator):
    """Scenes whose sockets are all constants run shade kernels compiled without the shader VM.
    Adding a material no mesh uses, with a procedural texture, makes the same scene run the
    kernels with the VM: both must trace bit-identical paths."""
    W = H = 32
    make = scene.cornell_box if integrator == abi.INTEGRATOR_PATH else (lambda: scene.fog_cornell("grid"))
    lean, full = make(), make()
    full.add_material(_program_material())
    s0, k0 = gpu_scene(dev, lean)
    s1, k1 = gpu_scene(dev, full)
    cam = scene.cornell_camera(W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    out = []
    for s in (s0, s1):
        film = gpu.Film(dev, W, H)
        d = gpu.render_desc(s, film, cam, 0, 2, spectral=abi.SPECTRAL_HERO4, estimator=abi.ESTIMATOR_CORRECTED,
                            integrator=integrator)
        out.append(gpu.trace_paths(d, pixels, samples))
    assert np.isfinite(out[0]).all()
    np.testing.assert_array_equal(out[0], out[1])


LOBE_CAMERAS = ["pinhole", "spherical", "blades"]


def _zoo_camera(kind, W, H):
    if kind == "spherical":
        return scene.camera_spherical((0.0, -0.1, 0.2), W, H, offset_phi=0.3, offset_theta=0.05)
    if kind == "blades":
        return scene.camera_blades((0.0, 0.0, 3.8), W, H, math.pi, 0.0, 0.6911, 3.9, 0.12, 6)
    return scene.cornell_camera(W, H)


@pytest.mark.parametrize("spectral,estimator", [(sp, e) for sp in (abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4)
                                                for e in (abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED)])
@pytest.mark.parametrize("cam_kind", LOBE_CAMERAS)
def test_lobe_zoo_per_path_parity(dev, cam_kind, spectral, estimator):
    """Every remaining reference lobe (FresnelBSDF, SpecularBRDF/BTDF, GhostBTDF, Oren-Nayar BTDF,
    Beckmann and dielectric-Fresnel microfacets, rough dielectric) seen through the pinhole, the
    SphericalCamera and a six-blade BladeAperture thin lens."""
    W, H = 64, 48
    sd = scene.lobe_zoo()
    s, k1 = gpu_scene(dev, sd)
    o, k2 = helpers.oracle_scene(sd)
    cam = _zoo_camera(cam_kind, W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), 2)
    samples = np.repeat(np.arange(2, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, 2, max_bounces=12, min_bounces=3, spectral=spectral, estimator=estimator)
    g = gpu.trace_paths(d, pixels, samples)
    c = o.trace_paths(helpers.desc(cam, 0, 2, 12, 3, spectral, estimator), W, H, pixels, samples)
    ok = _compare_paths(g, c)
    print("lobe zoo", cam_kind, spectral, estimator, "matching", ok.mean(), "nonzero", (np.abs(c[:, 0]) > 0).mean())
    assert np.isfinite(g[:, :4]).all()
    assert ok.mean() >= 0.999
