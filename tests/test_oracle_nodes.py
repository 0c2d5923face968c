"""Shader-graph nodes of the CPU checker (GraphTexture.cpp, GraphConverters.cpp, GraphMaths.cpp,
GraphInputs.cpp) against closed forms, lattice values and brute force."""
import math

import numpy as np
import pytest

from lambda_amd import abi, scene
from lambda_amd.scene import Material, node
from tests import helpers

N = abi


def _hash21(x, y):
    m = (1 << 32) - 1
    x = (x * 1597334677) & m
    y = (y * 3812015801) & m
    return ((x ^ y) * 1597334677) & m


def _unit(h):
    return float(np.float32(h / 4294967296.0))


@pytest.fixture(scope="module")
def nodes_scene():
    """One material per node under test; every node is the material's last node."""
    sd = scene.SceneDesc()
    rgba = np.zeros((4, 4, 4), dtype=np.float32)
    rgba[..., 0] = 0.25
    rgba[..., 3] = np.arange(4, dtype=np.float32)[None, :] / 4 + 0.1
    t_rgba = sd.add_texture(rgba)
    t_rgb = sd.add_texture(np.full((2, 2, 3), [0.2, 0.4, 0.6], dtype=np.float32))
    t_flat = sd.add_texture(np.ones((1, 8), dtype=np.float32))
    t_ramp = sd.add_texture(np.linspace(0, 1, 5, dtype=np.float32).reshape(1, 5))
    mats = {}

    def add(name, nodes):
        mats[name] = (sd.add_material(Material(nodes)), len(nodes) - 1)

    add("perlin", [node(N.NODE_PERLIN, p=[3])])
    add("value", [node(N.NODE_VALUE_NOISE, p=[2])])
    add("voronoi", [node(N.NODE_VORONOI, p=[2])])
    add("checker", [node(N.NODE_CHECKER, p=[2])])
    add("oct1", [node(N.NODE_OCTAVE_NOISE, p=[2, 1, N.NODE_VALUE_NOISE, 3])])
    add("oct2", [node(N.NODE_OCTAVE_NOISE, p=[2, 2, N.NODE_VALUE_NOISE, 3])])
    add("fresnel", [node(N.NODE_SCALAR, p=[1.5]), node(N.NODE_FRESNEL, [0])])
    add("fresnel_p", [node(N.NODE_FRESNEL, p=[2.0])])
    add("spd_flat", [node(N.NODE_SPECTRAL, p=[t_flat, 360, 830])])
    add("spd_ramp", [node(N.NODE_SPECTRAL, p=[t_ramp, 400, 800])])
    add("alpha", [node(N.NODE_IMAGE_CHANNEL, p=[t_rgba, 3])])
    add("alpha_rgb", [node(N.NODE_IMAGE_CHANNEL, p=[t_rgb, 3])])
    add("green", [node(N.NODE_IMAGE_CHANNEL, p=[t_rgb, 1])])
    add("spectex", [node(N.NODE_SPECTRAL_TEXTURE, p=[t_rgb])])
    add("normal", [node(N.NODE_NORMAL)])
    add("cross", [node(N.NODE_VEC3, p=[1, 2, 3]), node(N.NODE_VEC3, p=[-2, 0.5, 4]), node(N.NODE_CROSS, [0, 1])])
    add("sepz", [node(N.NODE_VEC3, p=[1, 2, 3]), node(N.NODE_SEPARATE_XYZ, [0], p=[2])])
    add("mergexyz", [node(N.NODE_SCALAR, p=[4]), node(N.NODE_SCALAR, p=[5]), node(N.NODE_SCALAR, p=[6]),
                     node(N.NODE_MERGE_XYZ, [0, 1, 2])])
    add("s2c_alpha", [node(N.NODE_SCALAR, p=[0.3]), node(N.NODE_SCALAR_TO_COLOUR, [0]), node(N.NODE_SEPARATE_R, [1], p=[3])])
    add("rgba_alpha", [node(N.NODE_SCALAR, p=[0.3]), node(N.NODE_SCALAR, p=[0.7]),
                       node(N.NODE_MERGE_RGB, [0, 0, 0, 1]), node(N.NODE_SEPARATE_R, [2], p=[3])])
    add("tex_alpha", [node(N.NODE_IMAGE_COLOUR, p=[t_rgba]), node(N.NODE_SEPARATE_R, [0], p=[3])])
    sd.add_material(scene.lambertian((0.5, 0.5, 0.5)))
    v, i = scene._quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1))
    sd.add_mesh(v, i, len(mats))
    o, keep = helpers.oracle_scene(sd, abi.LIGHTS_POWER)
    return o, mats, keep


def ev(o_m, name, **kw):
    o, mats, _ = o_m
    m, n = mats[name]
    return o.eval_node(m, n, **kw)


def test_perlin_is_half_at_lattice_points(nodes_scene):
    rng = np.random.default_rng(0)
    for _ in range(20):
        i, j = rng.integers(0, 12, size=2)
        t, v = ev(nodes_scene, "perlin", uv=(i / 3, j / 3))
        assert t == abi.VT_SCALAR
        assert v[0] == pytest.approx(0.5, abs=2e-6)
    vals = [ev(nodes_scene, "perlin", uv=tuple(rng.uniform(0, 4, 2)))[1][0] for _ in range(200)]
    assert min(vals) >= 0.5 - math.sqrt(0.5) and max(vals) <= 0.5 + math.sqrt(0.5)
    assert np.std(vals) > 0.05


def test_perlin_and_value_noise_vanish_in_media(nodes_scene):
    for name in ("perlin", "value"):
        assert ev(nodes_scene, name, uv=(0.3, 0.7), point=(0.2, 0.4, 0.9), medium=True)[1][0] == 0.0


def test_value_noise_interpolates_hashed_lattice(nodes_scene):
    for (i, j) in [(0, 0), (1, 3), (5, 2), (7, 7)]:
        want = _unit(_hash21(i, j))
        assert ev(nodes_scene, "value", uv=(i / 2, j / 2))[1][0] == pytest.approx(want, abs=1e-7)
        # halfway between two lattice points the smoothstep weight is exactly 1/2
        mid = 0.5 * (_unit(_hash21(i, j)) + _unit(_hash21(i + 1, j)))
        assert ev(nodes_scene, "value", uv=((i + 0.5) / 2, j / 2))[1][0] == pytest.approx(mid, abs=1e-6)


def _feature(c):
    c = np.asarray(c, dtype=np.float32)
    a = np.float32(c[0] * np.float32(127.1) + c[1] * np.float32(311.7))
    b = np.float32(c[0] * np.float32(269.5) + c[1] * np.float32(183.3))
    h = np.array([np.sin(a, dtype=np.float32), np.sin(b, dtype=np.float32)], dtype=np.float32) * np.float32(18.5453)
    return c + (h - np.floor(h))


def test_voronoi_is_distance_to_nearest_feature_point(nodes_scene):
    rng = np.random.default_rng(1)
    for _ in range(50):
        uv = rng.uniform(0, 3, size=2)
        t = uv * 2
        cell = np.floor(t)
        # brute force over a 5x5 neighbourhood (the node searches 3x3)
        d = min(np.linalg.norm(_feature(cell + [dx, dy]) - t) for dx in range(-2, 3) for dy in range(-2, 3))
        got = ev(nodes_scene, "voronoi", uv=tuple(uv))[1][0]
        assert got == pytest.approx(d, abs=2e-4) or got >= d
    # at a feature point the distance is 0
    fp = _feature([3.0, 1.0])
    assert ev(nodes_scene, "voronoi", uv=tuple(fp / 2))[1][0] == pytest.approx(0.0, abs=1e-3)


def test_voronoi_3d_in_media(nodes_scene):
    vals = [ev(nodes_scene, "voronoi", point=tuple(p), medium=True)[1][0]
            for p in np.random.default_rng(2).uniform(-2, 2, size=(100, 3))]
    assert min(vals) >= 0.0 and max(vals) <= math.sqrt(3.0) and np.std(vals) > 0.05


def test_checker_closed_forms(nodes_scene):
    rng = np.random.default_rng(3)
    for _ in range(100):
        uv = rng.uniform(0, 2, size=2)
        f = (uv * 2) % 1
        assert ev(nodes_scene, "checker", uv=tuple(uv))[1][0] == (0.0 if (f[0] > .5) == (f[1] > .5) else 1.0)
        p = rng.uniform(-2, 2, size=3)
        g = (p * 2) % 1
        want = 0.0 if ((g[0] > .5) == (g[1] > .5)) == (g[2] > .5) else 1.0
        assert ev(nodes_scene, "checker", point=tuple(p), medium=True)[1][0] == want


def test_octave_noise_sums_scaled_octaves(nodes_scene):
    o, mats, _ = nodes_scene
    for uv in [(0.13, 0.71), (1.4, 0.2), (2.9, 2.1)]:
        one = ev(nodes_scene, "oct1", uv=uv)[1][0]
        two = ev(nodes_scene, "oct2", uv=uv)[1][0]
        # base value noise of scale 3 inside an octave node of scale 2: coordinates x 2 x 2^k x 3
        b0 = ev(nodes_scene, "value", uv=(uv[0] * 3, uv[1] * 3))[1][0]        # value node has scale 2
        b1 = ev(nodes_scene, "value", uv=(uv[0] * 6, uv[1] * 6))[1][0]
        assert one == pytest.approx(0.5 * b0, abs=1e-6)
        assert two == pytest.approx(0.5 * b0 + 0.25 * b1, abs=1e-6)


def test_fresnel_node_is_schlick(nodes_scene):
    for eta in (1.0, 1.001, 1.3):
        r0 = ((eta - 1.5) / (eta + 1.5)) ** 2
        assert ev(nodes_scene, "fresnel", eta=eta, wo_y=1.0)[1][0] == pytest.approx(r0, rel=1e-5)
        assert ev(nodes_scene, "fresnel", eta=eta, wo_y=0.0)[1][0] == pytest.approx(1.0, rel=1e-6)
        c = 1 - 0.4
        assert ev(nodes_scene, "fresnel", eta=eta, wo_y=0.4)[1][0] == pytest.approx(r0 + (1 - r0) * c ** 5, rel=1e-5)
    r0 = ((1.0 - 2.0) / 3.0) ** 2
    assert ev(nodes_scene, "fresnel_p", eta=1.0, wo_y=1.0)[1][0] == pytest.approx(r0, rel=1e-6)


def test_spectral_input(nodes_scene):
    # a flat SPD of 1 is illuminant E: its RGB is XYZToRGB(1, 1, 1) up to the CMF fit's
    # integration error (the analytic CIE fit integrates to within ~1 % of the table)
    t, v = ev(nodes_scene, "spd_flat", mode=abi.SPECTRAL_RGB)
    assert t == abi.VT_SPECTRUM
    assert v[:3] == pytest.approx([1.2049, 0.9483, 0.9091], rel=1e-2)
    # hero wavelengths interpolate the samples linearly; outside the range they clamp
    lam = [380.0, 500.0, 650.0, 820.0]
    t, v = ev(nodes_scene, "spd_ramp", mode=abi.SPECTRAL_HERO4, lambdas=lam)
    want = [0.0, (500 - 400) / 400, (650 - 400) / 400, 1.0]
    assert v == pytest.approx(want, abs=1e-6)


def test_texture_channels_and_alpha(nodes_scene):
    # texel centres of the 4x4 RGBA texture: alpha = column / 4 + 0.1
    for col in range(4):
        u = (col + 0.0) / 4
        assert ev(nodes_scene, "alpha", uv=(u, 0.5))[1][0] == pytest.approx(col / 4 + 0.1, abs=1e-6)
        assert ev(nodes_scene, "tex_alpha", uv=(u, 0.5))[1][0] == pytest.approx(col / 4 + 0.1, abs=1e-6)
    assert ev(nodes_scene, "alpha_rgb", uv=(0.3, 0.3))[1][0] == 1.0
    assert ev(nodes_scene, "green", uv=(0.3, 0.3))[1][0] == pytest.approx(0.4)
    t, v = ev(nodes_scene, "spectex", uv=(0.3, 0.3), mode=abi.SPECTRAL_RGB)
    assert t == abi.VT_SPECTRUM and v[:3] == pytest.approx([0.2, 0.4, 0.6])


def test_vector_and_colour_converters(nodes_scene):
    t, v = ev(nodes_scene, "normal", normal=(0.6, 0.0, 0.8))
    assert t == abi.VT_VEC3 and v[:3] == pytest.approx([0.6, 0.0, 0.8])
    t, v = ev(nodes_scene, "cross")
    assert t == abi.VT_VEC3 and v[:3] == pytest.approx(list(np.cross([1, 2, 3], [-2, 0.5, 4])))
    assert ev(nodes_scene, "sepz")[1][0] == 3.0
    assert ev(nodes_scene, "mergexyz")[1][:3] == pytest.approx([4, 5, 6])
    assert ev(nodes_scene, "s2c_alpha")[1][0] == 1.0   # Colour(l) has alpha 1 (Colour.h:30-35)
    assert ev(nodes_scene, "rgba_alpha")[1][0] == pytest.approx(0.7)
