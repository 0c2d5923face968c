"""Learned path guiding on the GPU (guide.hip, the guided path shade kernel).

Bars:
  - network forward (BF16 MFMA, FP32 accumulation) against the CPU checker's restatement with the
    same BF16 rounding steps: bin probabilities within 2 quantisation steps (2/255 of the largest
    bin) for >= 99.9 % of entries, mixing probability within 1e-3;
  - gradient of the guiding objective against the checker's rounded restatement: cosine > 0.9999
    and |difference| <= 1e-2 of the largest entry (accumulation order and BF16 rounding flips);
  - unbiasedness: with a trained guide the film mean equals the plain estimator's (and the
    furnace's closed form Le / (1 - rho)) within statistical error;
  - gain: after online training the guided estimator's relMSE at equal samples is below the
    plain estimator's on an indirectly lit room.
"""
import numpy as np
import pytest

import oracle
from lambda_amd import abi, gpu, scene
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def test_forward_matches_oracle(dev):
    g = gpu.Guide(dev, seed=3)
    p = g.params()
    p[-144:] += np.random.default_rng(1).normal(scale=0.5, size=144).astype(np.float32)  # non-trivial table
    g.set_params(p)
    x = np.random.default_rng(2).uniform(size=(3000, 5)).astype(np.float32)
    a = g.forward(x)
    c = oracle.guide_forward(p, x)
    step = 2.0 / 255 * c[:, :128].max(axis=1, keepdims=True)
    ok = np.abs(a[:, :128] - c[:, :128]) <= step + 1e-6
    print("forward: entries within 2 steps", ok.mean())
    assert ok.mean() >= 0.999
    np.testing.assert_allclose(a[:, :128].sum(axis=1), 1.0, rtol=1e-5)
    assert np.abs(a[:, 128] - c[:, 128]).max() < 1e-3
    g.close()


def test_forward_full_size(dev):
    """2^20 queries in one inference launch (the size of a bounce's guided-vertex list at 1080p and
    half a sample per pixel): every bin table is a distribution (sum 1, every bin at least 1/255 of
    the largest), the mixing probability stays in [0.05, 0.9], and a spread sample of 4096 rows
    matches the CPU checker within 2 quantisation steps."""
    g = gpu.Guide(dev, seed=3)
    p = g.params()
    p[-144:] += np.random.default_rng(1).normal(scale=0.5, size=144).astype(np.float32)
    g.set_params(p)
    x = np.random.default_rng(5).uniform(size=(1 << 20, 5)).astype(np.float32)
    a = g.forward(x)
    np.testing.assert_allclose(a[:, :128].sum(axis=1), 1.0, rtol=1e-5)
    assert (a[:, :128].min(axis=1) >= a[:, :128].max(axis=1) / 255 * (1 - 1e-6)).all()
    assert a[:, 128].min() >= 0.05 - 1e-6 and a[:, 128].max() <= 0.9 + 1e-6
    rows = np.linspace(0, len(x) - 1, 4096).astype(np.int64)
    c = oracle.guide_forward(p, x[rows])
    step = 2.0 / 255 * c[:, :128].max(axis=1, keepdims=True)
    ok = np.abs(a[rows, :128] - c[:, :128]) <= step + 1e-6
    print("forward 2^20: sample rows within 2 steps", ok.mean())
    assert ok.mean() >= 0.999
    assert np.abs(a[rows, 128] - c[:, 128]).max() < 1e-3
    g.close()


def _records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, :5] = rng.uniform(0, 1, size=(n, 5))
    r[:, 5] = rng.integers(0, 128, size=n)
    r[:, 6] = rng.uniform(0.01, 2.0, size=n)
    r[:, 7] = rng.exponential(1.0, size=n)
    r[::5, 7] = 0.0
    return r


@pytest.mark.parametrize("n", [64, 1000])
def test_gradient_matches_oracle(dev, n):
    g = gpu.Guide(dev, seed=5)
    p = g.params()
    p[-144:] += np.random.default_rng(4).normal(scale=0.3, size=144).astype(np.float32)
    p[64 * 80 + 4 * 64 * 64:64 * 80 + 4 * 64 * 64 + 144 * 64] = np.random.default_rng(6).normal(
        scale=0.1, size=144 * 64).astype(np.float32)
    g.set_params(p)
    rec = _records(n, 7)
    ga, la = g.gradient(rec)
    gc, lc = oracle.guide_gradient(p, rec, rounded=True)
    cos = float(ga @ gc / (np.linalg.norm(ga) * np.linalg.norm(gc)))
    err = np.abs(ga - gc).max() / np.abs(gc).max()
    print("gradient: cos %.7f, max err / max %.2e, loss %.6f vs %.6f" % (cos, err, la, lc))
    assert cos > 0.9999
    assert err < 1e-2
    assert abs(la - lc) < 1e-4 * abs(lc)
    g.close()


def _render(dev, sd, cam, W, H, spp, guide=None, steps=0, passes=1, seed=0, spectral=abi.SPECTRAL_HERO4):
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    film = gpu.Film(dev, W, H)
    per = spp // passes
    stats = []
    for k in range(passes):
        d = gpu.render_desc(s, film, cam, k * per, (k + 1) * per, max_bounces=8, min_bounces=3,
                            spectral=spectral, estimator=abi.ESTIMATOR_CORRECTED, seed=seed,
                            guide=guide, guide_train_steps=steps, capacity=W * H * per)
        stats.append(gpu.render(d))
    rgb, cnt = film.read()
    film.close()
    del keep
    s.close()
    return rgb / cnt[..., None], stats


def test_guided_furnace_is_unbiased(dev):
    """Closed Lambertian box (rho 0.5, unit emission, RGB spectra): L = Le / (1 - rho) = 2 for every
    pixel, also when part of the scattering directions come from a trained guide."""
    sd = helpers.furnace_box(albedo=0.5, emit=1.0, subdivide=2)
    cam = scene.camera_pinhole((0.1, -0.2, 0.05), 48, 48, 0.3, 0.2, 1.2)
    g = gpu.Guide(dev, seed=1, batch=4096)
    try:
        img, st = _render(dev, sd, cam, 48, 48, 256, guide=g, steps=8, passes=8, spectral=abi.SPECTRAL_RGB)
        info = g.stats()
        print("furnace guided mean", img.mean(axis=(0, 1)), "records", sum(x.guide_records for x in st), info)
        assert sum(x.guide_records for x in st) > 0 and info["steps"] == 64
        np.testing.assert_allclose(img.mean(axis=(0, 1)), 2.0, rtol=5e-3)
    finally:
        g.close()


def test_guide_outliving_its_device_is_released():
    """A guide whose device is destroyed first keeps a valid handle: its calls report a state error
    and destroying it frees the handle (no access to the destroyed device)."""
    d = gpu.Device(0)
    g = gpu.Guide(d, seed=1)
    d.close()
    with pytest.raises(gpu.LambdaError) as e:
        g.params()
    assert e.value.status == abi.ERR_STATE
    g.close()


def test_guided_matches_plain_estimator(dev):
    """Cornell box: guided (online-trained) and plain film means agree per channel."""
    sd = scene.cornell_box()
    W = H = 64
    cam = scene.cornell_camera(W, H)
    plain, _ = _render(dev, sd, cam, W, H, 512)
    g = gpu.Guide(dev, seed=2, batch=8192)
    try:
        guided, st = _render(dev, sd, cam, W, H, 512, guide=g, steps=8, passes=8)
        a, b = plain.mean(axis=(0, 1)), guided.mean(axis=(0, 1))
        print("cornell plain", a, "guided", b, "guide", g.stats())
        np.testing.assert_allclose(b, a, rtol=1e-2)
    finally:
        g.close()


def test_guided_traversal_count_keeps_the_film(dev):
    """A guided render with traversal counting (the extension stage that counts node and triangle
    fetches and lists the guided vertices) gives the same film, bit for bit, as the same guided
    render without counting (same guide, no training in between), and counts its fetches."""
    sd = scene.cornell_box()
    W = H = 64
    cam = scene.cornell_camera(W, H)
    s = gpu.Scene(dev)
    keep = sd.push(s, abi.LIGHTS_TREE)
    g = gpu.Guide(dev, seed=8, batch=4096)
    films = [gpu.Film(dev, W, H) for _ in range(2)]
    try:
        st = []
        for f, flags in zip(films, (0, abi.RENDER_COUNT_TRAVERSAL)):
            d = gpu.render_desc(s, f, cam, 0, 16, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_HERO4,
                                estimator=abi.ESTIMATOR_CORRECTED, seed=3, guide=g, guide_train_steps=0,
                                capacity=W * H * 16, flags=flags)
            st.append(gpu.render(d))
        (a, ca), (b, cb) = films[0].read(), films[1].read()
        print("guided counting pass: nodes4", st[1].nodes4_fetched, "rays", st[1].extension_rays)
        assert st[1].nodes4_fetched > 0 and st[1].extension_rays == st[0].extension_rays
        assert np.array_equal(a, b) and np.array_equal(ca, cb)
    finally:
        for f in films:
            f.close()
        g.close()
        del keep
        s.close()


def test_guided_matches_plain_on_glossy_lobes(dev):
    """Room of microfacet BRDFs (Beckmann and GGX, dielectric and conductor Fresnel), a rough
    dielectric and Lambertian walls, beside delta lobes that stay unguided: guided and plain film
    means agree per channel at 4096 spp (the plain mean's standard error is about 0.25 % per
    channel, from eight seeds of the CPU checker's render of the same room)."""
    sd = scene.lobe_zoo()
    W = H = 64
    cam = scene.cornell_camera(W, H)
    plain, _ = _render(dev, sd, cam, W, H, 4096)
    g = gpu.Guide(dev, seed=6, batch=8192)
    try:
        guided, st = _render(dev, sd, cam, W, H, 4096, guide=g, steps=8, passes=8)
        a, b = plain.mean(axis=(0, 1)), guided.mean(axis=(0, 1))
        print("lobe zoo plain", a, "guided", b, "records", sum(x.guide_records for x in st))
        assert sum(x.guide_records for x in st) > 0
        np.testing.assert_allclose(b, a, rtol=2e-2)
    finally:
        g.close()


def test_guiding_reduces_error_on_indirect_light(dev):
    """Indirectly lit room: after online training, the guided estimator's relMSE at 64 spp (its
    last 64 samples, rendered with the trained guide) is lower than the plain estimator's (0.65 to
    0.71 of it over two seeds and two learning rates; training is not bit-reproducible, as the
    gradient is summed with atomics)."""
    sd = scene.cornell_indirect()
    W = H = 64
    cam = scene.cornell_camera(W, H)
    ref, _ = _render(dev, sd, cam, W, H, 16384, seed=11)

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 0.01)))

    plain, _ = _render(dev, sd, cam, W, H, 64, seed=3)
    g = gpu.Guide(dev, seed=4, batch=16384, learning_rate=5e-3)
    try:
        # train online on 16 passes of 32 spp, then render 64 spp with the trained guide
        _render(dev, sd, cam, W, H, 512, guide=g, steps=32, passes=16, seed=5)
        guided, _ = _render(dev, sd, cam, W, H, 64, guide=g, steps=0, seed=3)
        rp, rg = relmse(plain), relmse(guided)
        print("indirect room relMSE: plain %.4e guided %.4e (ratio %.3f), mean plain %s guided %s ref %s"
              % (rp, rg, rg / rp, plain.mean(axis=(0, 1)), guided.mean(axis=(0, 1)), ref.mean(axis=(0, 1))))
        np.testing.assert_allclose(guided.mean(axis=(0, 1)), ref.mean(axis=(0, 1)), rtol=0.03)
        assert rg < 0.9 * rp
    finally:
        g.close()
