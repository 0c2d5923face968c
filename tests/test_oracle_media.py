"""CPU checks of the oracle's media: homogeneous medium (HomogeneousMedium.cpp:12-92),
the density-grid medium, Henyey-Greenstein (HenyeyGreenstein.cpp:8-36) and the two
volumetric integrators (VolumetricPathIntegrator.cpp:20-162,
MISVolumetricPathIntegrator.cpp:23-215), pinned by closed forms, quadrature and
invariants rather than by the oracle's own formulas."""
import math

import numpy as np
import pytest

import oracle
from lambda_amd import abi, scene
from helpers import desc, oracle_scene

SA = np.array([0.2, 0.5, 0.1], dtype=np.float32)
SS = np.array([0.7, 0.3, 0.9], dtype=np.float32)


def _medium_scene(kind="homogeneous", density=None, bmin=(-1, -1, -1), bmax=(1, 1, 1), g=0.3):
    sd = scene.SceneDesc()
    if kind == "homogeneous":
        scene.homogeneous(sd, SA, SS, g)
    else:
        scene.grid_medium(sd, density, SA, SS, g, bmin, bmax)
    o = oracle.OracleScene()
    keep = None
    # add only the medium: closed-form access needs no geometry and no commit
    for _, pl in sd.ops:
        d = abi.MediumDesc()
        d.type = pl["type"]
        d.sigma_a[:] = [float(x) for x in pl["sigma_a"]]
        d.sigma_s[:] = [float(x) for x in pl["sigma_s"]]
        d.g = float(pl["g"])
        if "density" in pl:
            dens = np.ascontiguousarray(pl["density"], dtype=np.float32)
            keep = dens
            d.nz, d.ny, d.nx = dens.shape
            d.density = dens.ctypes.data_as(abi.FP)
            d.bounds_min[:] = [float(x) for x in pl["bounds_min"]]
            d.bounds_max[:] = [float(x) for x in pl["bounds_max"]]
        o.add_medium(d)
    return o, keep


def _density_integral(dens, bmin, bmax, o, d, t_far, n=200000):
    """Midpoint-rule integral of the nearest-voxel density along o + t d, t in [0, t_far]."""
    nz, ny, nx = dens.shape
    t = (np.arange(n) + 0.5) / n * t_far
    p = o[None, :] + t[:, None] * d[None, :]
    bmin, bmax = np.asarray(bmin, np.float64), np.asarray(bmax, np.float64)
    inside = np.all((p >= bmin) & (p < bmax), axis=1)
    cell = np.floor((p - bmin) / (bmax - bmin) * np.array([nx, ny, nz])).astype(np.int64)
    cell = np.clip(cell, 0, np.array([nx - 1, ny - 1, nz - 1]))
    rho = dens[cell[:, 2], cell[:, 1], cell[:, 0]] * inside
    return rho.sum() * t_far / n


def _rand_dir(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def test_tr_homogeneous_closed_form():
    o, _ = _medium_scene()
    m = oracle.Medium(o, 0)
    for t in (0.0, 0.3, 1.7, 12.0):
        tr = m.tr((0, 0, 0), (0, 0, 1), t)
        np.testing.assert_allclose(tr[:3], np.exp(-(SA + SS) * t), rtol=2e-6)


def test_tr_grid_matches_quadrature():
    rng = np.random.default_rng(5)
    dens = rng.uniform(0, 2, size=(5, 6, 7)).astype(np.float32)
    bmin, bmax = (-1.0, -0.5, -0.8), (1.2, 0.9, 0.7)
    o, _ = _medium_scene("grid", dens, bmin, bmax)
    m = oracle.Medium(o, 0)
    for _ in range(20):
        org = rng.uniform(-2, 2, size=3)
        d = _rand_dir(rng)
        t_far = rng.uniform(0.5, 5.0)
        D = _density_integral(dens, bmin, bmax, org, d, t_far)
        np.testing.assert_allclose(m.tr(org, d, t_far)[:3], np.exp(-(SA + SS) * D), rtol=2e-3, atol=1e-6)


def test_constant_grid_equals_homogeneous():
    dens = np.ones((3, 4, 5), dtype=np.float32)
    og, _ = _medium_scene("grid", dens, (-10, -10, -10), (10, 10, 10))
    oh, _ = _medium_scene()
    mg, mh = oracle.Medium(og, 0), oracle.Medium(oh, 0)
    rng = np.random.default_rng(2)
    for _ in range(50):
        org, d = rng.uniform(-1, 1, size=3), _rand_dir(rng)
        t_far = rng.uniform(0.1, 4.0)
        np.testing.assert_allclose(mg.tr(org, d, t_far), mh.tr(org, d, t_far), rtol=1e-5)
        u0, u1 = rng.uniform(size=2)
        a, b = mg.sample_distance(org, d, t_far, u0, u1), mh.sample_distance(org, d, t_far, u0, u1)
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6)


def test_sample_distance_homogeneous_inversion():
    """Channel chosen by u0, exponential inversion on that channel; pdf is the
    channel average of sigma_t Tr (interaction) or Tr (surface)."""
    o, _ = _medium_scene()
    m = oracle.Medium(o, 0)
    st = (SA + SS).astype(np.float64)
    rng = np.random.default_rng(1)
    for _ in range(200):
        u0, u1 = rng.uniform(size=2)
        t_far = rng.uniform(0.2, 6)
        t, pdf, inter, *w = m.sample_distance((0, 0, 0), (1, 0, 0), t_far, u0, u1)
        ch = min(int(u0 * 3), 2)
        t_exp = min(-math.log(1 - u1) / st[ch], t_far)
        assert t == pytest.approx(t_exp, rel=1e-5)
        assert bool(inter) == (t_exp < t_far)
        tr = np.exp(-st * t)
        if inter:
            assert pdf == pytest.approx(np.mean(st * tr), rel=1e-5)
            np.testing.assert_allclose(w[:3], tr * SS / np.mean(st * tr), rtol=1e-4)
        else:
            assert pdf == pytest.approx(np.mean(tr), rel=1e-5)
            np.testing.assert_allclose(w[:3], tr / np.mean(tr), rtol=1e-4)


def test_sample_distance_grid_inversion():
    """The sampled point's optical depth on the chosen channel equals -log(1-u1)."""
    rng = np.random.default_rng(9)
    dens = rng.uniform(0, 3, size=(6, 6, 6)).astype(np.float32)
    dens[dens < 0.8] = 0
    bmin, bmax = (-1, -1, -1), (1, 1, 1)
    o, _ = _medium_scene("grid", dens, bmin, bmax)
    m = oracle.Medium(o, 0)
    st = (SA + SS).astype(np.float64)
    hits = 0
    for _ in range(60):
        org, d = rng.uniform(-1.5, 1.5, size=3), _rand_dir(rng)
        u0, u1 = rng.uniform(size=2)
        t_far = 4.0
        t, pdf, inter, *w = m.sample_distance(org, d, t_far, u0, u1)
        ch = min(int(u0 * 3), 2)
        if inter:
            hits += 1
            D = _density_integral(dens, bmin, bmax, org, d, t)
            assert st[ch] * D == pytest.approx(-math.log(1 - u1), rel=3e-3, abs=1e-4)
            # corrected PDFDistance reproduces the sampling pdf
            assert m.pdf_distance(org, d, t, corrected=True) == pytest.approx(pdf, rel=1e-4)
        else:
            D = _density_integral(dens, bmin, bmax, org, d, t_far)
            assert st[ch] * D <= -math.log(1 - u1) * (1 + 3e-3) + 1e-4
    assert hits > 10


def test_pdf_distance_faithful_is_zero():
    """A14: `pdf *= 1 / Spectrum::nSamples` is an integer division in the reference."""
    o, _ = _medium_scene()
    m = oracle.Medium(o, 0)
    assert m.pdf_distance((0, 0, 0), (0, 1, 0), 0.7, corrected=False) == 0.0
    assert m.pdf_distance((0, 0, 0), (0, 1, 0), 0.7, corrected=True) > 0.0


def test_equiangular_normalised_and_inverse_cdf():
    """Corrected equiangular density over [0, tFar] integrates to 1 and the
    sample for eps is the point where the CDF reaches 1 - eps (lerp(thetaB, thetaA, eps))."""
    o, _ = _medium_scene()
    m = oracle.Medium(o, 0)
    rng = np.random.default_rng(4)
    for _ in range(10):
        org, d = rng.uniform(-1, 1, size=3), _rand_dir(rng)
        light = rng.uniform(-2, 2, size=3)
        t_far = rng.uniform(0.5, 5)
        n = 20000
        ts = (np.arange(n) + 0.5) / n * t_far
        pdf = np.array([oracle.pdf_equiangular(org, d, light, t_far, t) for t in ts[::20]])
        # analytic check of the density at a few points against D / ((thB - thA)(D^2 + (t - l)^2))
        delta = light - org
        l = np.dot(d, delta)
        D = math.sqrt(np.dot(delta, delta) - l * l)
        thA, thB = -math.atan2(l, D), math.atan2(t_far - l, D)
        ref = D / ((thB - thA) * (D * D + (ts[::20] - l) ** 2))
        np.testing.assert_allclose(pdf, ref, rtol=1e-4)
        assert ref.sum() * (t_far / n) * 20 == pytest.approx(1.0, rel=2e-3)
        for eps in (0.1, 0.5, 0.9):
            t, p, inter, *_ = m.sample_equiangular(org, d, org, t_far, light, eps, corrected=True)
            cdf = (math.atan2(t - l, D) - thA) / (thB - thA)
            assert cdf == pytest.approx(1 - eps, abs=1e-4)
            assert p == pytest.approx(D / ((thB - thA) * (D * D + (t - l) ** 2)), rel=1e-4)


def test_equiangular_faithful_drops_offset():
    """A14 faithful: t = D tan(theta) measured from the foot point is used as the ray
    distance (the `+ l` is missing), and the pdf uses that t."""
    o, _ = _medium_scene()
    m = oracle.Medium(o, 0)
    org, d, light, t_far = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([1.5, 0.5, 0]), 4.0
    eps = 0.3
    tc, *_ = m.sample_equiangular(org, d, org, t_far, light, eps, corrected=True)
    tf, pf, *_ = m.sample_equiangular(org, d, org, t_far, light, eps, corrected=False)
    assert tf == pytest.approx(tc - 1.5, abs=1e-5)
    D = 0.5
    thA, thB = -math.atan2(1.5, D), math.atan2(t_far - 1.5, D)
    assert pf == pytest.approx(D / ((thB - thA) * (D * D + tf * tf)), rel=1e-4)


def _hg(cos_std, g):
    """Henyey-Greenstein with cos between propagation direction and wi."""
    return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos_std) ** 1.5)


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.6])
def test_hg_sampling_corrected(g):
    rng = np.random.default_rng(11)
    wo = _rand_dir(rng)
    d = -wo  # propagation direction
    cosines = []
    for _ in range(4000):
        u0, u1 = rng.uniform(size=2)
        p, wi = oracle.hg_sample(g, wo, u0, u1, corrected=True)
        assert np.linalg.norm(wi) == pytest.approx(1.0, abs=1e-5)
        c = float(np.dot(d, wi))
        assert p == pytest.approx(_hg(c, g), rel=2e-3)
        assert oracle.lib().oracle_hg_p(float(np.dot(wo, wi)), g) == pytest.approx(p, rel=2e-3)
        cosines.append(c)
    # Kolmogorov-Smirnov against the analytic HG CDF in cos
    xs = np.sort(np.array(cosines))
    grid = np.linspace(-1, 1, 4001)
    pdf = np.array([2 * math.pi * _hg(x, g) for x in grid])
    cdf = np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) * 0.5 * np.diff(grid))])
    assert cdf[-1] == pytest.approx(1.0, abs=1e-3)
    emp = np.arange(1, len(xs) + 1) / len(xs)
    ks = np.max(np.abs(np.interp(xs, grid, cdf) - emp))
    assert ks < 1.63 / math.sqrt(len(xs)) + 2e-3  # 99% KS bound


def test_hg_faithful_unnormalised_frame():
    """A21: the frame uses the unnormalised cross(-wo, y), so |wi|^2 =
    sin^2(a) sin^2(theta) + cos^2(theta) with a the angle between wo and y."""
    wo = np.array([0.0, 0.6, 0.8])
    sin_a2 = 1 - 0.6 ** 2
    for u0, u1 in ((0.2, 0.1), (0.7, 0.9)):
        p, wi = oracle.hg_sample(0.4, wo, u0, u1, corrected=False)
        g = 0.4
        sq = (1 - g * g) / (1 - g + 2 * g * u0)
        cos_t = (1 + g * g - sq * sq) / (2 * g)
        assert float(np.dot(wi, wi)) == pytest.approx(sin_a2 * (1 - cos_t ** 2) + cos_t ** 2, rel=1e-4)


# ------------------------------------------------------------------ integrators
def _single_scatter_scene(g=0.4):
    sd = scene.SceneDesc()
    med = scene.homogeneous(sd, SA, SS, g)
    box = sd.add_material(scene.passthrough(interior=med))
    v, i = scene.aabb((-1, -1, -1), (1, 1, 1))
    sd.add_mesh(v, i, box)
    sd.add_point_light((0.2, 0.3, -0.1), 1.0)
    return sd


def _centre_paths(o, d, W, n):
    px = np.full(n, (W // 2) * W + W // 2, dtype=np.uint32)
    return o.trace_paths(d, W, W, px, np.arange(n, dtype=np.uint32), threads=8)[:, :3].astype(np.float64)


def _single_scatter_reference(g, faithful):
    """L = int_0^2 sigma_s Tr(t) p(theta) I Tr_l / d^2 dt along the axis ray through the box.
    faithful: no light-leg transmittance (A42) and the power-heuristic weight
    1 / (1 + p^2) on the delta light's sample (A38)."""
    light = np.array([0.2, 0.3, -0.1])
    st = (SA + SS).astype(np.float64)
    n = 200000
    t = (np.arange(n) + 0.5) / n * 2.0
    x = np.stack([np.zeros(n), np.zeros(n), 1.0 - t], axis=1)
    v = light - x
    dist = np.linalg.norm(v, axis=1)
    cos_std = v[:, 2] * -1.0 / dist  # propagation (0,0,-1) . wi
    p = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos_std) ** 1.5)
    out = []
    for c in range(3):
        tl = 1.0 / (1.0 + p * p) if faithful else np.exp(-st[c] * dist)
        out.append(np.sum(SS[c] * np.exp(-st[c] * t) * p * tl / dist ** 2) * 2.0 / n)
    return np.array(out)


@pytest.mark.parametrize("integrator,estimator", [
    (abi.INTEGRATOR_VOLPATH, abi.ESTIMATOR_CORRECTED),
    (abi.INTEGRATOR_MIS_VOLPATH, abi.ESTIMATOR_CORRECTED),
    (abi.INTEGRATOR_VOLPATH, abi.ESTIMATOR_FAITHFUL),
])
def test_single_scattering_closed_form(integrator, estimator):
    g = 0.4
    o, _ = oracle_scene(_single_scatter_scene(g))
    W = 4096
    d = desc(scene.cornell_camera(W, W), 0, 1, max_bounces=1, min_bounces=8, estimator=estimator)
    d.integrator = integrator
    L = _centre_paths(o, d, W, 200000)
    ref = _single_scatter_reference(g, estimator == abi.ESTIMATOR_FAITHFUL)
    err = L.std(0) / math.sqrt(len(L))
    assert np.all(np.abs(L.mean(0) - ref) < 4.5 * err + 1e-3 * ref), (L.mean(0), ref, err)


def _slab_scene():
    sd = scene.SceneDesc()
    med = scene.homogeneous(sd, [0.2, 0.5, 1.0], [0.0, 0.0, 0.0])
    box = sd.add_material(scene.passthrough(interior=med))
    v, i = scene.aabb((-1, -1, -1), (1, 1, 1))
    sd.add_mesh(v, i, box)
    em = sd.add_material(scene.Material([scene.node(abi.NODE_RGB, p=(1, 1, 1))], emission=0,
                                        intensity=float(np.float32(np.pi))))
    v, i = scene._quad((-3, -3, -2), (3, -3, -2), (3, 3, -2), (-3, 3, -2))
    sd.add_mesh(v, i, em)
    return sd


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_VOLPATH, abi.INTEGRATOR_MIS_VOLPATH])
def test_absorbing_slab(integrator):
    """An emitter (radiance 1) seen through 2 units of an absorbing chromatic medium:
    corrected gives Beer-Lambert per channel; faithful VOLPATH adds the emission
    without the distance-sampling weight, i.e. the channel-averaged transmittance."""
    o, _ = oracle_scene(_slab_scene())
    W = 64
    tr = np.exp(-2 * np.array([0.2, 0.5, 1.0]))
    d = desc(scene.cornell_camera(W, W), 0, 1, max_bounces=4, min_bounces=8, estimator=abi.ESTIMATOR_CORRECTED)
    d.integrator = integrator
    L = _centre_paths(o, d, W, 100000)
    err = L.std(0) / math.sqrt(len(L))
    assert np.all(np.abs(L.mean(0) - tr) < 4.5 * err + 1e-6), (L.mean(0), tr)
    if integrator == abi.INTEGRATOR_VOLPATH:
        d.estimator_mode = abi.ESTIMATOR_FAITHFUL
        L = _centre_paths(o, d, W, 100000)
        err = L.std(0) / math.sqrt(len(L))
        assert np.all(np.abs(L.mean(0) - tr.mean()) < 4.5 * err + 1e-6), (L.mean(0), tr.mean())


@pytest.mark.parametrize("kind", ["homogeneous", "grid"])
def test_corrected_vol_integrators_agree(kind):
    """Both corrected estimators are unbiased, so their image means agree."""
    o, _ = oracle_scene(scene.fog_cornell(kind))
    W, spp = 16, 512
    cam = scene.cornell_camera(W, W)
    px = np.repeat(np.arange(W * W, dtype=np.uint32), spp)
    sm = np.tile(np.arange(spp, dtype=np.uint32), W * W)
    res = []
    for integ in (abi.INTEGRATOR_VOLPATH, abi.INTEGRATOR_MIS_VOLPATH):
        d = desc(cam, 0, spp, estimator=abi.ESTIMATOR_CORRECTED)
        d.integrator = integ
        L = o.trace_paths(d, W, W, px, sm, threads=8)[:, :3].astype(np.float64)
        res.append((L.mean(0), L.std(0) / math.sqrt(len(L))))
    (a, ea), (b, eb) = res
    z = np.abs(a - b) / np.sqrt(ea ** 2 + eb ** 2)
    assert np.all(z < 4.5), (a, b, z)


def test_constant_grid_matches_homogeneous_paths():
    """A unit-density grid over the fog box gives the homogeneous medium's paths."""
    W = 16
    cam = scene.cornell_camera(W, W)
    outs = []
    for kind in ("homogeneous", "grid1"):
        sd = scene.cornell_box()
        bmin, bmax = (-0.95, -0.95, -0.95), (0.95, 0.6, 0.95)
        if kind == "homogeneous":
            med = scene.homogeneous(sd, SA, SS, 0.3)
        else:
            med = scene.grid_medium(sd, np.ones((2, 3, 4), np.float32), SA, SS, 0.3, (-2, -2, -2), (2, 2, 2))
        fog = sd.add_material(scene.passthrough(interior=med))
        v, i = scene.aabb(bmin, bmax)
        sd.add_mesh(v, i, fog)
        sd.add_point_light((0.35, 0.1, 0.3), 0.6)
        o, _ = oracle_scene(sd)
        d = desc(cam, 0, 4, estimator=abi.ESTIMATOR_CORRECTED, spectral=abi.SPECTRAL_HERO4)
        d.integrator = abi.INTEGRATOR_MIS_VOLPATH
        px = np.repeat(np.arange(W * W, dtype=np.uint32), 4)
        sm = np.tile(np.arange(4, dtype=np.uint32), W * W)
        outs.append(o.trace_paths(d, W, W, px, sm, threads=8))
    a, b = outs
    close = np.all(np.abs(a - b) <= 1e-3 * np.maximum(np.abs(a), np.abs(b)) + 1e-6, axis=1)
    assert close.mean() > 0.97, close.mean()


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_VOLPATH, abi.INTEGRATOR_MIS_VOLPATH])
@pytest.mark.parametrize("estimator", [abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED])
@pytest.mark.parametrize("spectral", [abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4])
def test_vol_integrators_finite(integrator, estimator, spectral):
    o, _ = oracle_scene(scene.fog_cornell("grid"))
    W = 16
    d = desc(scene.cornell_camera(W, W), 0, 8, estimator=estimator, spectral=spectral)
    d.integrator = integrator
    rgb, cnt = o.render(d, W, W)
    assert np.isfinite(rgb).all() and (cnt == 8).all() and rgb.sum() > 0
