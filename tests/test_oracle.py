"""CPU tests that pin the checker (oracle/oracle.cpp) before it is trusted.

Each test uses an independent fact: published known-answer vectors, closed
forms, normalisation integrals, brute force, or an energy-balance (furnace)
result, chosen so that a plausible restatement mistake fails one of them.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from lambda_amd import abi, scene
from tests import helpers

L = oracle.lib()


def f3(*v):
    return (C.c_float * 3)(*v)


# ---------------------------------------------------------------- Philox4x32-10
# Known-answer vectors of the Random123 distribution (kat_vectors, philox4x32_10).
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert list(oracle.philox(ctr, key)) == expect


# ---------------------------------------------------------------- Fresnel (Fresnel.cpp:5-45)
def test_fresnel_dielectric_closed_forms():
    for eta in (1.33, 1.5, 2.4):
        assert L.oracle_fr_dielectric(1.0, 1.0, eta) == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-5)
    # total internal reflection from the dense side
    assert L.oracle_fr_dielectric(math.cos(math.radians(60)), 1.5, 1.0) == 1.0
    # Brewster angle: r_parallel = 0, F = r_perp^2 / 2
    eta = 1.5
    tb = math.atan(eta)
    ci, ct = math.cos(tb), math.cos(math.asin(math.sin(tb) / eta))
    rperp = (ci - eta * ct) / (ci + eta * ct)
    assert L.oracle_fr_dielectric(ci, 1.0, eta) == pytest.approx(rperp ** 2 / 2, rel=1e-4)
    # entering from inside swaps the indices (cos < 0)
    assert L.oracle_fr_dielectric(-0.9, 1.0, 1.5) == pytest.approx(L.oracle_fr_dielectric(0.9, 1.5, 1.0), rel=1e-6)


def test_fresnel_conductor_matches_complex_fresnel():
    # independent closed form: complex index n - i k, unpolarised average of |r_s|^2 and |r_p|^2
    eta = np.array([0.200438, 0.924033, 1.102212])
    k = np.array([3.912949, 2.452848, 2.142188])
    for cos_i in (1.0, 0.8, 0.5, 0.2, 0.05):
        out = (C.c_float * 3)()
        L.oracle_fr_conductor(cos_i, f3(1, 1, 1), f3(*eta), f3(*k), out)
        n = eta + 1j * k
        sin2 = 1 - cos_i ** 2
        cos_t = np.sqrt(1 - sin2 / n ** 2)
        rs = (cos_i - n * cos_t) / (cos_i + n * cos_t)
        rp = (n * cos_i - cos_t) / (n * cos_i + cos_t)
        ref = (np.abs(rs) ** 2 + np.abs(rp) ** 2) / 2
        np.testing.assert_allclose(np.array(out[:]), ref, rtol=2e-4)


# ---------------------------------------------------------------- microfacets (MicrofacetDistribution.cpp)
@pytest.mark.parametrize("ggx", [0, 1])
@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.7])
def test_microfacet_d_normalised(ggx, alpha):
    # integral of D(h) cos(theta_h) over the hemisphere is 1
    nt, npf = 4000, 8
    th = (np.arange(nt) + 0.5) / nt * (math.pi / 2)
    total = 0.0
    for t in th:
        for ph in (np.arange(npf) + 0.5) / npf * 2 * math.pi:
            h = (C.c_float * 3)(math.sin(t) * math.cos(ph), math.cos(t), math.sin(t) * math.sin(ph))
            total += L.oracle_microfacet_d(ggx, h, alpha, alpha) * math.cos(t) * math.sin(t)
    total *= (math.pi / 2 / nt) * (2 * math.pi / npf)
    assert total == pytest.approx(1.0, rel=2e-3)


@pytest.mark.parametrize("ggx", [0, 1])
def test_microfacet_lambda_nonnegative(ggx):
    rng = np.random.default_rng(1)
    for _ in range(500):
        t = rng.uniform(0, math.pi / 2 - 1e-3)
        ph = rng.uniform(0, 2 * math.pi)
        w = (C.c_float * 3)(math.sin(t) * math.cos(ph), math.cos(t), math.sin(t) * math.sin(ph))
        lam = L.oracle_microfacet_lambda(ggx, w, rng.uniform(0.05, 1), rng.uniform(0.05, 1))
        # G1 = 1 / (1 + Lambda) <= 1; the Beckmann rational fit (Walter 2007 / PBRT) dips to
        # -6.4e-4 just below its a = 1.6 cutoff, so it gets that tolerance.
        assert lam >= (0.0 if ggx else -7e-4)
    # GGX Lambda closed form (Heitz 2014) for isotropic alpha
    t, a = 1.0, 0.4
    w = (C.c_float * 3)(math.sin(t), math.cos(t), 0.0)
    if ggx:
        ref = (-1 + math.sqrt(1 + (a * math.tan(t)) ** 2)) / 2
        assert L.oracle_microfacet_lambda(1, w, a, a) == pytest.approx(ref, rel=1e-5)


# ---------------------------------------------------------------- spectra
def test_cmf_fit_values_and_integral():
    out = (C.c_float * 3)()
    L.oracle_cmf(555.0, out)
    assert out[1] == pytest.approx(1.0, abs=0.01)  # CIE 1931 ybar(555) = 1.000
    L.oracle_cmf(600.0, out)
    assert out[0] == pytest.approx(1.0622, abs=0.02)  # xbar(600) = 1.0622
    ys = []
    for lam in range(360, 831):
        L.oracle_cmf(float(lam), out)
        ys.append(out[1])
    assert sum(ys) == pytest.approx(106.856895, rel=3e-3)


def test_blackbody_normalised_peak():
    for T in (2000.0, 6500.0, 9000.0):
        peak = 2.8977721e-3 / T * 1e9
        assert L.oracle_blackbody_normalised(peak, T) == pytest.approx(1.0, rel=1e-5)
        for lam in (peak * 0.8, peak * 1.2):
            assert L.oracle_blackbody_normalised(lam, T) < 1.0
    # Planck ratio closed form between two wavelengths
    T, l1, l2 = 6500.0, 450e-9, 650e-9
    h, c, kb = 6.62606957e-34, 299792458.0, 1.3806488e-23
    planck = lambda l: 1 / (l ** 5 * (math.exp(h * c / (l * kb * T)) - 1))
    r = L.oracle_blackbody_normalised(450.0, T) / L.oracle_blackbody_normalised(650.0, T)
    assert r == pytest.approx(planck(l1) / planck(l2), rel=1e-4)


def test_blackbody_rgb_white_balance():
    out = (C.c_float * 3)()
    L.oracle_blackbody_rgb(6500.0, out)
    r, g, b = out[:]
    assert min(r, g, b) > 0
    L.oracle_blackbody_rgb(2000.0, out)
    assert out[0] > out[2]  # a warm source is red-heavy
    L.oracle_blackbody_rgb(9000.0, out)
    assert out[2] > out[0] * 0.8


def test_uplift_partition_of_unity():
    for lam in np.linspace(360, 829, 200):
        assert L.oracle_uplift(f3(1, 1, 1), float(lam)) == pytest.approx(1.0, abs=1e-6)
        for rgb in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            v = L.oracle_uplift(f3(*rgb), float(lam))
            assert -1e-7 <= v <= 1 + 1e-7
    assert L.oracle_uplift(f3(0.73, 0.73, 0.73), 500.0) == pytest.approx(0.73, rel=1e-6)


# ---------------------------------------------------------------- Piecewise1D (Piecewise.cpp:9-44)
def test_piecewise_discrete_pdf_and_histogram():
    f = np.array([1.0, 0.0, 3.0, 2.0, 4.0], dtype=np.float32)
    pdf = C.c_float()
    counts = np.zeros(5)
    us = (np.arange(20000) + 0.5) / 20000
    for u in us:
        i = L.oracle_piecewise1d_sample(f.ctypes.data_as(C.c_void_p), 5, float(u), C.byref(pdf))
        counts[i] += 1
        assert pdf.value == pytest.approx(f[i] / f.sum(), rel=1e-5)
    np.testing.assert_allclose(counts / len(us), f / f.sum(), atol=2e-4)
    assert counts[1] == 0


# ---------------------------------------------------------------- HG phase (HenyeyGreenstein.cpp:8-12)
@pytest.mark.parametrize("g", [-0.5, 0.0, 0.6, 0.9])
def test_hg_normalised_and_mean_cosine(g):
    n = 20000
    mu = -1 + (np.arange(n) + 0.5) * 2 / n
    p = np.array([L.oracle_hg_p(float(m), g) for m in mu])
    assert np.sum(p) * 2 / n * 2 * math.pi == pytest.approx(1.0, rel=1e-3)
    # the reference's cos convention is cos(wo, wi) with wo reversed: mean of -mu is g
    assert np.sum(p * -mu) * 2 / n * 2 * math.pi == pytest.approx(g, abs=2e-3)


# ---------------------------------------------------------------- BVH vs brute force
def _brute(verts, tris, o, d, tmin, tmax):
    best = (np.inf, 0, 0, 0xFFFFFFFF)
    for k, t in enumerate(tris):
        v0, v1, v2 = verts[t[0]], verts[t[1]], verts[t[2]]
        e1, e2 = v1 - v0, v2 - v0
        pv = np.cross(d, e2)
        det = e1 @ pv
        if det == 0:
            continue
        tv = o - v0
        u = (tv @ pv) / det
        qv = np.cross(tv, e1)
        v = (d @ qv) / det
        t_ = (e2 @ qv) / det
        if u < 0 or v < 0 or u + v > 1 or not (tmin < t_ < tmax):
            continue
        if t_ < best[0]:
            best = (t_, u, v, k)
    return best


def test_oracle_bvh_matches_brute_force():
    sd = scene.cornell_box()
    o, keep = helpers.oracle_scene(sd)
    verts, tris = [], []
    for kind, pl in sd.ops:
        if kind == "mesh":
            base = len(verts)
            verts += list(pl["positions"].astype(np.float64))
            tris += list(pl["indices"].astype(np.int64) + base)
    verts, tris = np.array(verts), np.array(tris)
    rng = np.random.default_rng(7)
    n = 2000
    org = rng.uniform(-0.95, 0.95, size=(n, 3))
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = org, 0.0, dirs, np.inf
    hits = o.intersect(rays)
    ties = 0
    for i in range(n):
        r = rays[i].astype(np.float64)
        bt, bu, bv, bp = _brute(verts, tris, r[0:3], r[4:7], 0.0, np.inf)
        t = hits[i, :1].view(np.float32)[0]
        if bp == 0xFFFFFFFF:
            assert hits[i, 3] == bp
            continue
        assert t == pytest.approx(bt, rel=1e-5)
        if hits[i, 3] != bp:
            # the box bottoms lie in the floor plane (y = -1): coplanar triangles tie in t
            ties += 1
            assert verts[tris[bp]][:, 1].max() == verts[tris[hits[i, 3]]][:, 1].max() == -1.0
    assert ties < n // 100


# ---------------------------------------------------------------- estimator energy balance
@pytest.mark.parametrize("strategy", [abi.LIGHTS_POWER, abi.LIGHTS_TREE])
@pytest.mark.parametrize("spectral", [abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4])
def test_furnace_corrected_is_unbiased(strategy, spectral):
    # Closed box, Lambertian albedo rho, uniform emitted radiance Le: L = Le / (1 - rho).
    rho = 0.5
    sd = helpers.furnace_box(albedo=rho, emit=1.0, subdivide=2)
    o, keep = helpers.oracle_scene(sd, strategy)
    cam = scene.camera_pinhole((0.1, -0.2, 0.05), 32, 32, 0.3, 0.2, 1.2)
    d = helpers.desc(cam, 0, 64, max_bounces=64, min_bounces=3, spectral=spectral, estimator=abi.ESTIMATOR_CORRECTED)
    rgb, cnt = o.render(d, 32, 32, threads=8)
    mean = (rgb / cnt[..., None]).mean(axis=(0, 1))
    expect = 1.0 / (1.0 - rho)
    if spectral == abi.SPECTRAL_HERO4:
        # constant spectrum 1 maps to the sRGB coordinates of illuminant E
        xyz_to_rgb_e = np.array([1.2049, 0.9483, 0.9091]) * expect
        np.testing.assert_allclose(mean, xyz_to_rgb_e, rtol=0.03)
    else:
        np.testing.assert_allclose(mean, expect, rtol=0.01)


def test_faithful_tree_mode_reproduces_reference_bias():
    # TriangleLight::Sample_Li multiplies d^2 * cos * A into the pdf (MeshLight.cpp:145):
    # in the closed box the faithful tree estimate is far from Le / (1 - rho).
    sd = helpers.furnace_box(albedo=0.5, emit=1.0, subdivide=1)
    o, keep = helpers.oracle_scene(sd, abi.LIGHTS_TREE)
    cam = scene.camera_pinhole((0.0, 0.0, 0.0), 16, 16, 0.0, 0.0, 1.2)
    d = helpers.desc(cam, 0, 16, max_bounces=32, estimator=abi.ESTIMATOR_FAITHFUL)
    rgb, cnt = o.render(d, 16, 16, threads=8)
    mean = (rgb / cnt[..., None]).mean()
    assert abs(mean - 2.0) > 0.2


@pytest.mark.parametrize("strategy", [abi.LIGHTS_TREE, abi.LIGHTS_POWER])
@pytest.mark.parametrize("spectral", [abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4])
def test_direct_lighting_furnace(strategy, spectral):
    """Closed Lambertian box of emission Le and albedo rho: the irradiance is pi Le everywhere,
    so direct lighting reflects rho Le and (corrected) the first surface adds Le."""
    rho, emit = 0.5, 1.0
    o, _ = helpers.oracle_scene(helpers.furnace_box(albedo=rho, emit=emit, subdivide=2), strategy)
    W = 16
    cam = scene.camera_pinhole((0.1, -0.2, 0.05), W, W, 0.3, 0.2, 1.2)
    d = helpers.desc(cam, 0, 64, estimator=abi.ESTIMATOR_CORRECTED, spectral=spectral)
    d.integrator = abi.INTEGRATOR_DIRECT
    rgb, cnt = o.render(d, W, W)
    mean = (rgb / cnt[..., None]).mean(axis=(0, 1))
    expect = emit * (1 + rho)
    if spectral == abi.SPECTRAL_HERO4:  # constant spectrum: sRGB coordinates of illuminant E
        np.testing.assert_allclose(mean, np.array([1.2049, 0.9483, 0.9091]) * expect, rtol=0.03)
    else:
        np.testing.assert_allclose(mean, expect, rtol=0.01)


def _many_light_quads(n_quads, seed):
    """Small emissive quads in random positions and orientations (blackbody, varied power)."""
    rng = np.random.default_rng(seed)
    sd = scene.SceneDesc()
    sd.add_material(scene.lambertian((0.5, 0.5, 0.5)))
    for _ in range(n_quads):
        m = sd.add_material(scene.emitter_blackbody(float(rng.uniform(2000, 9000)), float(10 ** rng.uniform(0, 2))))
        c = rng.uniform(-1.5, 1.5, size=3)
        a, b = rng.normal(size=3), rng.normal(size=3)
        a = 0.1 * a / np.linalg.norm(a)
        b = np.cross(a, b)
        b = 0.1 * b / np.linalg.norm(b)
        v = np.array([c - a - b, c + a - b, c + a + b, c - a + b], dtype=np.float32)
        sd.add_mesh(v, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32), m)
    return sd


def test_light_tree_pdfs_sum_to_one_and_match_sampling():
    """Corrected tree: sum over lights of P(pick) = 1 at any point, and pick frequencies follow
    the pdf (brute-force enumeration plus a chi-square-style check)."""
    o, _ = helpers.oracle_scene(_many_light_quads(24, seed=4), abi.LIGHTS_TREE)
    rng = np.random.default_rng(0)
    for _ in range(5):
        p = rng.uniform(-2, 2, size=3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        pdfs, env = o.light_pdfs(p, n, corrected=True)
        assert pdfs.sum() + env == pytest.approx(1.0, abs=1e-5)
        counts = np.zeros(len(pdfs))
        N = 20000
        for k in range(N):
            idx, pk = o.light_sample(p, n, k, 0, corrected=True)
            assert idx >= 0
            assert pk == pytest.approx(pdfs[idx], rel=1e-4)
            counts[idx] += 1
        expected = pdfs * N
        m = expected > 5
        chi2 = ((counts[m] - expected[m]) ** 2 / expected[m]).sum()
        assert chi2 < m.sum() + 6 * np.sqrt(2 * m.sum()) + 10


def test_tree_importance_closed_forms():
    """Node importance (ManyLightSampler.cpp:305-330) against hand-computed values. A node of power E
    one unit above the origin, facing down onto it (axis -y, thetaO 0, thetaE pi/2), point normal +y:
    the corrected orientation measures the emitter cone against the direction from the light to the
    point and gives E / d^2 = E; as written it measures it against the direction towards the light,
    the angle is pi and the node gets nothing (A47). Moving the node to (1, 1, 0) puts both the
    emitter angle and the incidence angle at pi/4: E/2 * cos^2(pi/4) = E/4. A node of diagonal D
    widens both angles by atan(D / 2d). A zero normal (medium point) drops the incidence term. As
    written the incidence cosine is signed, so a node below the shading horizon gets a negative
    importance (A48); corrected takes |cos|."""
    E = 3.0
    down = ((0, 1, 0), 0.0, (0, -1, 0), 0.0, math.pi / 2, E)
    assert oracle.tree_importance(*down, (0, 0, 0), (0, 1, 0), True) == pytest.approx(E, rel=1e-6)
    assert oracle.tree_importance(*down, (0, 0, 0), (0, 1, 0), False) == 0.0
    up = ((0, 1, 0), 0.0, (0, 1, 0), 0.0, math.pi / 2, E)
    assert oracle.tree_importance(*up, (0, 0, 0), (0, 1, 0), True) == 0.0
    assert oracle.tree_importance(*up, (0, 0, 0), (0, 1, 0), False) == pytest.approx(E, rel=1e-6)
    side = ((1, 1, 0), 0.0, (0, -1, 0), 0.0, math.pi / 2, E)
    assert oracle.tree_importance(*side, (0, 0, 0), (0, 1, 0), True) == pytest.approx(E / 4, rel=1e-5)
    # diagonal D at distance sqrt(2): both angles shrink by thetaU = atan(D / (2 sqrt 2))
    D = 0.5
    tu = math.atan(D / (2 * math.sqrt(2)))
    wide = ((1, 1, 0), D, (0, -1, 0), 0.0, math.pi / 2, E)
    want = math.cos(math.pi / 4 - tu) ** 2 * E / 2
    assert oracle.tree_importance(*wide, (0, 0, 0), (0, 1, 0), True) == pytest.approx(want, rel=1e-5)
    # medium point: no incidence term
    assert oracle.tree_importance(*side, (0, 0, 0), (0, 0, 0), True) == pytest.approx(E / 2 * math.cos(math.pi / 4), rel=1e-5)
    # shading normal facing away from the node: |cos| when corrected, cos(pi) = -1 as written
    assert oracle.tree_importance(*down, (0, 0, 0), (0, -1, 0), True) == pytest.approx(E, rel=1e-6)
    assert oracle.tree_importance(*up, (0, 0, 0), (0, -1, 0), False) == pytest.approx(-E, rel=1e-6)
    # outside the emission cone: thetaO 0, thetaE pi/4, emitter angle pi/4 + 0.1
    a = math.pi / 4 + 0.1
    off = ((math.sin(a), math.cos(a), 0), 0.0, (0, -1, 0), 0.0, math.pi / 4, E)
    assert oracle.tree_importance(*off, (0, 0, 0), (0, 1, 0), True) == 0.0


def test_direct_lighting_faithful_runs():
    o, _ = helpers.oracle_scene(scene.cornell_box())
    W = 16
    for est in (abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED):
        d = helpers.desc(scene.cornell_camera(W, W), 0, 4, estimator=est)
        d.integrator = abi.INTEGRATOR_DIRECT
        rgb, cnt = o.render(d, W, W)
        assert np.isfinite(rgb).all() and rgb.sum() > 0


def _wall(z=-1.2, albedo=0.5):
    sd = scene.SceneDesc()
    m = sd.add_material(scene.lambertian((albedo, albedo, albedo)))
    v, i = scene._quad((-1, -1, z), (1, -1, z), (1, 1, z), (-1, 1, z))  # Ng = +z, towards the camera
    sd.add_mesh(v, i, m)
    return sd


@pytest.mark.parametrize("estimator", [abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED])
def test_utility_passes_closed_forms(estimator):
    """DepthPass / NormalPass / AOVPass (UtilityIntegrators.cpp:26-77) on a wall 5 units in front of
    a narrow pinhole: depth = (distance / maxDepth)^power, the background where the wall is missed;
    the wall's normal (0, 0, 1) maps to (0.5, 1, 1) as written (y offset by 1, A50) and (0.5, 0.5, 1)
    corrected, misses to 0; the AOV pass is black."""
    o, _ = helpers.oracle_scene(_wall(), abi.LIGHTS_POWER)
    W = H = 8
    cam = scene.camera_pinhole((0.0, 0.0, 3.8), W, H, math.pi, 0.0, 0.02)  # all rays hit the wall
    wide = scene.camera_pinhole((0.0, 0.0, 3.8), W, H, math.pi, 0.0, 1.2)  # the corners miss it
    for integ, want in [(abi.INTEGRATOR_DEPTH, None), (abi.INTEGRATOR_NORMAL, None), (abi.INTEGRATOR_AOV, 0.0)]:
        d = helpers.desc(cam, 0, 2, estimator=estimator)
        d.integrator = integ
        d.pass_max_depth, d.pass_power = 10.0, 2.0
        rgb, cnt = o.render(d, W, H)
        img = rgb / cnt[..., None]
        if integ == abi.INTEGRATOR_DEPTH:
            np.testing.assert_allclose(img, (5.0 / 10.0) ** 2, rtol=1e-3)
        elif integ == abi.INTEGRATOR_NORMAL:
            y = 1.0 if estimator == abi.ESTIMATOR_FAITHFUL else 0.5
            np.testing.assert_allclose(img.reshape(-1, 3), np.tile([0.5, y, 1.0], (W * H, 1)), atol=1e-6)
        else:
            assert (img == 0).all()
    d = helpers.desc(wide, 0, 1, estimator=estimator)
    d.integrator = abi.INTEGRATOR_DEPTH
    d.pass_max_depth, d.pass_power = 10.0, 1.0
    d.pass_background[:] = [0.25, 0.5, 0.75]
    rgb, cnt = o.render(d, W, H)
    px = (rgb / cnt[..., None]).reshape(-1, 3)
    miss = np.all(px == np.float32([0.25, 0.5, 0.75]), axis=1)
    assert 0 < miss.sum() < W * H
    assert (px[~miss] >= 0.5 - 1e-6).all()


def _ellipsoid_scene(as_instance, xfm, level=2):
    """One smooth icosphere under xfm: as an instance of a proxy, or as a mesh transformed here in
    float64 (normals by the inverse transpose, tangents by the linear part)."""
    sd = scene.SceneDesc()
    m = sd.add_material(scene.lambertian((0.5, 0.5, 0.5)))
    v, f = scene.icosphere(level)
    n = scene._vertex_normals(v, f)
    t = scene._tangents(n)
    if as_instance:
        p = sd.add_proxy(v, f, m, normals=n, tangents=t)
        sd.add_instance(p, xfm)
    else:
        x = np.asarray(xfm, dtype=np.float64)
        M = np.stack([x[0:3], x[3:6], x[6:9]], axis=1)
        vv = v @ M.T + x[9:12]
        nn = n @ np.linalg.inv(M)
        nn /= np.linalg.norm(nn, axis=1, keepdims=True)
        tt = t[:, :3] @ M.T
        tt /= np.linalg.norm(tt, axis=1, keepdims=True)
        tt = np.concatenate([tt, t[:, 3:] * np.sign(np.linalg.det(M))], axis=1)
        sd.add_mesh(vv, f, m, normals=nn, tangents=tt)
    return sd


@pytest.mark.parametrize("mirror", [False, True])
def test_instance_matches_transformed_mesh(mirror):
    """Instance::Commit flattened (Instance.cpp:22-43): an instance of a smooth sphere under rotation,
    non-uniform scale (and a mirror) gives the same shading normals (NormalPass) and depths as the
    same mesh transformed independently in float64 with inverse-transpose normals."""
    sc = (-1.4, 0.7, 1.1) if mirror else (1.4, 0.7, 1.1)
    xfm = scene.affine((0.3, -0.5, 0.9), sc, (0.1, -0.2, 0.0))
    W = H = 24
    cam = scene.cornell_camera(W, H)
    imgs = {}
    for inst in (True, False):
        o, _ = helpers.oracle_scene(_ellipsoid_scene(inst, xfm), abi.LIGHTS_POWER)
        for integ in (abi.INTEGRATOR_NORMAL, abi.INTEGRATOR_DEPTH):
            d = helpers.desc(cam, 0, 1, estimator=abi.ESTIMATOR_CORRECTED)
            d.integrator = integ
            d.pass_max_depth = 10.0
            rgb, cnt = o.render(d, W, H)
            imgs[inst, integ] = rgb / cnt[..., None]
    for integ in (abi.INTEGRATOR_NORMAL, abi.INTEGRATOR_DEPTH):
        a, b = imgs[True, integ], imgs[False, integ]
        hit = np.any(a != (0 if integ == abi.INTEGRATOR_NORMAL else np.float32([1, 0, 1])), axis=2)
        assert 0.1 < hit.mean() < 0.9
        np.testing.assert_allclose(a, b, atol=2e-4)


def test_identity_instance_equals_mesh_and_rejects_singular():
    sd_i, sd_m = _ellipsoid_scene(True, scene.affine()), _ellipsoid_scene(False, scene.affine())
    W = H = 16
    d = helpers.desc(scene.cornell_camera(W, H), 0, 2, estimator=abi.ESTIMATOR_CORRECTED)
    d.integrator = abi.INTEGRATOR_NORMAL
    oi, _ = helpers.oracle_scene(sd_i, abi.LIGHTS_POWER)
    om, _ = helpers.oracle_scene(sd_m, abi.LIGHTS_POWER)
    # the instance renormalises the normals in float32: equal up to rounding
    np.testing.assert_allclose(oi.render(d, W, H)[0], om.render(d, W, H)[0], atol=1e-6)
    bad = scene.SceneDesc()
    m = bad.add_material(scene.lambertian((0.5, 0.5, 0.5)))
    v, f = scene.icosphere(0)
    p = bad.add_proxy(v, f, m)
    bad.add_instance(p, [1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0])
    with pytest.raises(Exception):
        helpers.oracle_scene(bad, abi.LIGHTS_POWER)


def test_emissive_instances_are_lights():
    """An instance with an emissive material is a mesh light (Scene::AddObject adds its light)."""
    o, _ = helpers.oracle_scene(scene.instanced_cornell(n=6), abi.LIGHTS_TREE)
    # 2 (room light) + 2 x 80 (two glowing spheres of 80 triangles) + 4 x 2 (quad instances)
    pdfs, env = o.light_pdfs((0, 0, 0), (0, 1, 0))
    assert len(pdfs) == 2 + 2 * 320 + 8
    assert pdfs.sum() + env == pytest.approx(1.0, abs=1e-5)
