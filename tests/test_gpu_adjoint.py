"""Path-replay adjoint (lgpu_render_adjoint) against central finite differences of the forward
render (lgpu_render, volumetric path integrator, corrected estimator, RGB).

The loss is the linear functional sum(w * image) with fixed random weights w, so d_image = w.
Where the forward estimator's sampling does not depend on the parameter, the finite difference of
the forward render with common random numbers is the derivative of the same estimator the
adjoint differentiates, and the two agree to rounding (central difference, h = 1e-3):
  - albedo texels, multi-bounce with a grid medium (Russian roulette off): 256 parameters;
  - roughness and alpha texels under a point light, direct lighting only: 256 parameters;
  - grid densities of a slab that only light paths cross (transmittance only): all 256 voxels;
  - the emission intensity of the area light, multi-bounce.
Where sampling depends on the parameter (roughness and density along multi-bounce paths), the
adjoint (detached sampling) and the finite difference (attached) are different unbiased estimators
of the same derivative: directional derivatives over 8 seeds agree within their statistical error.
"""
import numpy as np
import pytest

from lambda_amd import abi, gpu, scene

pytestmark = pytest.mark.gpu

VOL = abi.INTEGRATOR_VOLPATH


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def _sizes(sd):
    tex = [o[1].size for o in sd.ops if o[0] == "texture"]
    dens = [np.asarray(o[1]["density"]).size for o in sd.ops if o[0] == "medium" and "density" in o[1]]
    mats = sum(1 for o in sd.ops if o[0] == "material")
    return tex, dens, mats


def _desc(s, film, cam, spp, seed, max_b, min_b):
    return gpu.render_desc(s, film, cam, 0, spp, max_bounces=max_b, min_bounces=min_b, spectral=abi.SPECTRAL_RGB,
                           estimator=abi.ESTIMATOR_CORRECTED, seed=seed, integrator=VOL)


class Setup:
    def __init__(self, dev, sd, cam, W, H, spp, max_b, min_b, wseed=1):
        self.dev, self.sd, self.cam, self.W, self.H, self.spp = dev, sd, cam, W, H, spp
        self.max_b, self.min_b = max_b, min_b
        self.s = gpu.Scene(dev)
        self.keep = sd.push(self.s)
        self.tex, self.dens, self.nmat = _sizes(sd)
        self.tex_off = np.concatenate([[0], np.cumsum(self.tex)]).astype(int)
        self.dens_off = np.concatenate([[0], np.cumsum(self.dens)]).astype(int)
        self.w = np.random.default_rng(wseed).normal(size=(H, W, 3)).astype(np.float32)
        self.tex_data = [o[1].reshape(-1).copy() for o in sd.ops if o[0] == "texture"]
        self.dens_data = [np.asarray(o[1]["density"], dtype=np.float32).reshape(-1).copy()
                          for o in sd.ops if o[0] == "medium" and "density" in o[1]]

    def close(self):
        self.s.close()

    def loss(self, seed, scene_=None):
        film = gpu.Film(self.dev, self.W, self.H)
        gpu.render(_desc(scene_ or self.s, film, self.cam, self.spp, seed, self.max_b, self.min_b))
        rgb, cnt = film.read()
        film.close()
        return float((rgb / cnt[..., None] * self.w.astype(np.float64)).sum())

    def adjoint(self, seed):
        film = gpu.Film(self.dev, self.W, self.H)
        gt, gd, ge, st = gpu.render_adjoint(_desc(self.s, film, self.cam, self.spp, seed, self.max_b, self.min_b), self.w,
                                            sum(self.tex), sum(self.dens), self.nmat)
        film.close()
        return gt, gd, ge

    def fd_texel(self, t, k, h, seed):
        base = self.tex_data[t]
        v = base.copy()
        v[k] = base[k] + h
        self.s.update_texture(t, v)
        lp = self.loss(seed)
        v[k] = base[k] - h
        self.s.update_texture(t, v)
        lm = self.loss(seed)
        self.s.update_texture(t, base)
        return (lp - lm) / (2 * h)

    def fd_voxel(self, m, k, h, seed):
        base = self.dens_data[m]
        v = base.copy()
        v[k] = base[k] + h
        self.s.update_density(m, v)
        lp = self.loss(seed)
        v[k] = base[k] - h
        self.s.update_density(m, v)
        lm = self.loss(seed)
        self.s.update_density(m, base)
        return (lp - lm) / (2 * h)


def _check(fd, ad, what):
    fd, ad = np.asarray(fd), np.asarray(ad)
    scale = np.abs(fd).max()
    err = np.abs(fd - ad)
    print("%s: %d parameters, max |fd| %.4g, max |fd - adjoint| %.3g (%.2e of max), rel L2 %.2e"
          % (what, len(fd), scale, err.max(), err.max() / scale, np.linalg.norm(fd - ad) / np.linalg.norm(fd)))
    assert scale > 0
    assert np.all(err <= 2e-3 * scale + 2e-3 * np.abs(fd)), np.argsort(-err)[:5]
    assert np.linalg.norm(fd - ad) <= 2e-3 * np.linalg.norm(fd)


def _pick(g, n, rng, floor=1e-3):
    live = np.flatnonzero(np.abs(g) > floor * np.abs(g).max())
    return rng.choice(live, size=min(n, len(live)), replace=False)


def test_adjoint_primal_equals_render(dev):
    """The primal pass of the adjoint render accumulates the same film as lgpu_render."""
    sd, ids = scene.inverse_room(tex_res=8, medium=True)
    cam = scene.cornell_camera(32, 32)
    S = Setup(dev, sd, cam, 32, 32, 8, 6, 2)
    try:
        f1 = gpu.Film(dev, 32, 32)
        gpu.render(_desc(S.s, f1, cam, 8, 9, 6, 2))
        f2 = gpu.Film(dev, 32, 32)
        gpu.render_adjoint(_desc(S.s, f2, cam, 8, 9, 6, 2), S.w, sum(S.tex), sum(S.dens), S.nmat)
        a, b = f1.read()[0], f2.read()[0]
        f1.close()
        f2.close()
        np.testing.assert_array_equal(a, b)
    finally:
        S.close()


def test_albedo_gradients_match_finite_differences(dev):
    sd, ids = scene.inverse_room(tex_res=16, medium=True)
    S = Setup(dev, sd, scene.cornell_camera(48, 48), 48, 48, 16, 4, 4)  # min = max bounces: no roulette
    try:
        gt, _, _ = S.adjoint(seed=7)
        t = ids["albedo"]
        ga = gt[S.tex_off[t]:S.tex_off[t + 1]]
        picks = _pick(ga, 256, np.random.default_rng(2))
        fd = [S.fd_texel(t, int(k), 1e-3, 7) for k in picks]
        assert len(picks) == 256
        _check(fd, ga[picks], "albedo (multi-bounce, grid medium)")
    finally:
        S.close()


def test_roughness_gradients_match_finite_differences(dev):
    """Point light, direct lighting only: light sampling does not depend on roughness."""
    sd, ids = scene.inverse_room(tex_res=16, medium=False, point_light=True)
    S = Setup(dev, sd, scene.cornell_camera(64, 64), 64, 64, 8, 1, 1)
    try:
        gt, _, _ = S.adjoint(seed=3)
        rng = np.random.default_rng(4)
        # up to 208 roughness and alpha texels the light reaches, albedo texels for the rest
        pool = []
        for name in ("roughness", "alpha"):
            t = ids[name]
            pool += [(t, int(k)) for k in _pick(gt[S.tex_off[t]:S.tex_off[t + 1]], 10 ** 6, rng)]
        picks = [pool[j] for j in rng.choice(len(pool), size=min(208, len(pool)), replace=False)]
        t = ids["albedo"]
        picks += [(t, int(k)) for k in _pick(gt[S.tex_off[t]:S.tex_off[t + 1]], 256 - len(picks), rng)]
        assert len(picks) == 256 and len(pool) >= 128
        fd = [S.fd_texel(t, k, 1e-3, 3) for t, k in picks]
        ad = [gt[S.tex_off[t] + k] for t, k in picks]
        _check(fd, ad, "roughness / alpha / albedo (direct, point light)")
    finally:
        S.close()


def test_density_gradients_match_finite_differences(dev):
    """Light reaches the floor through the slab only (camera rays stay outside it): the density
    enters through transmittance alone."""
    sd, ids = scene.slab_scene()
    S = Setup(dev, sd, scene.slab_camera(48, 48), 48, 48, 8, 1, 1)
    try:
        _, gd, _ = S.adjoint(seed=5)
        m = 0
        fd = [S.fd_voxel(m, k, 1e-3, 5) for k in range(S.dens[m])]
        _check(fd, gd[:S.dens[m]], "density (transmittance, 256 voxels)")
    finally:
        S.close()


def test_emission_gradient_matches_finite_differences(dev):
    sd, ids = scene.inverse_room(tex_res=8, medium=True)
    cam = scene.cornell_camera(32, 32)
    S = Setup(dev, sd, cam, 32, 32, 16, 4, 4)
    try:
        _, _, ge = S.adjoint(seed=11)
        h = 1e-2
        ls = []
        for I in (10.0 + h, 10.0 - h):
            sd2, _ = scene.inverse_room(tex_res=8, medium=True, light_intensity=I)
            s2 = gpu.Scene(dev)
            keep = sd2.push(s2)
            ls.append(S.loss(11, s2))
            del keep
            s2.close()
        fd = (ls[0] - ls[1]) / (2 * h)
        print("emission: fd %.6g adjoint %.6g" % (fd, ge[ids["light"]]))
        assert abs(fd - ge[ids["light"]]) <= 2e-3 * abs(fd)
        assert np.count_nonzero(ge) == 1
    finally:
        S.close()


def _directional(S, seeds, direction, fd_fn):
    a, f = [], []
    for sd_ in seeds:
        a.append(direction(S.adjoint(sd_)))
        f.append(fd_fn(sd_))
    a, f = np.array(a), np.array(f)
    se = np.sqrt(a.var(ddof=1) / len(a) + f.var(ddof=1) / len(f))
    return a.mean(), f.mean(), se


def test_multibounce_roughness_and_density_are_unbiased(dev):
    """Directional derivatives of the image sum (every roughness and alpha texel + h; every
    density scaled by 1 + h; h = 0.05) at 2048 spp over 8 seeds: the adjoint's mean equals the
    finite difference's within 4 standard errors and 8 % (measured: roughness -2.33 +- 0.06 vs
    -2.39 +- 0.04, density 7.66 +- 0.01 vs 7.72 +- 0.06 at 4096 spp)."""
    sd, ids = scene.inverse_room(tex_res=8, medium=True)
    S = Setup(dev, sd, scene.cornell_camera(32, 32), 32, 32, 2048, 6, 2)
    S.w[:] = 1.0
    seeds = list(range(20, 28))
    try:
        tr, ta = ids["roughness"], ids["alpha"]

        def rough_dir(g):
            return float(g[0][S.tex_off[tr]:S.tex_off[ta + 1]].sum())

        def rough_fd(seed, h=0.05):
            base = [S.tex_data[tr].copy(), S.tex_data[ta].copy()]
            out = []
            for sgn in (1, -1):
                S.s.update_texture(tr, base[0] + sgn * h)
                S.s.update_texture(ta, base[1] + sgn * h)
                out.append(S.loss(seed))
            S.s.update_texture(tr, base[0])
            S.s.update_texture(ta, base[1])
            return (out[0] - out[1]) / (2 * h)

        am, fm, se = _directional(S, seeds, rough_dir, rough_fd)
        print("roughness direction: adjoint %.5g fd %.5g se %.3g" % (am, fm, se))
        assert abs(am - fm) <= 4 * se and abs(am - fm) <= 0.08 * abs(fm)

        dens = S.dens_data[0]

        def dens_dir(g):
            return float((g[1][:S.dens[0]] * dens).sum())

        def dens_fd(seed, h=0.05):
            out = []
            for sgn in (1, -1):
                S.s.update_density(0, dens * (1 + sgn * h))
                out.append(S.loss(seed))
            S.s.update_density(0, dens)
            return (out[0] - out[1]) / (2 * h)

        am, fm, se = _directional(S, seeds, dens_dir, dens_fd)
        print("density direction: adjoint %.5g fd %.5g se %.3g" % (am, fm, se))
        assert abs(am - fm) <= 4 * se and abs(am - fm) <= 0.08 * abs(fm)
    finally:
        S.close()


def test_adjoint_rejects_unsupported_configurations(dev):
    sd, _ = scene.inverse_room(tex_res=4, medium=False)
    S = Setup(dev, sd, scene.cornell_camera(8, 8), 8, 8, 1, 2, 1)
    try:
        film = gpu.Film(dev, 8, 8)
        for kw in (dict(integrator=abi.INTEGRATOR_PATH), dict(spectral=abi.SPECTRAL_HERO4),
                   dict(estimator=abi.ESTIMATOR_FAITHFUL)):
            args = dict(spectral=abi.SPECTRAL_RGB, estimator=abi.ESTIMATOR_CORRECTED, integrator=VOL)
            args.update(kw)
            d = gpu.render_desc(S.s, film, S.cam, 0, 1, **args)
            with pytest.raises(gpu.LambdaError) as e:
                gpu.render_adjoint(d, S.w, sum(S.tex), 0, S.nmat)
            assert e.value.status == abi.ERR_UNSUPPORTED
        film.close()
    finally:
        S.close()


def test_torch_harness_matches_adjoint(dev):
    """The autograd harness's gradients are the adjoint render's (same seeds) in the parameters'
    shapes; forward returns the film mean."""
    import torch
    from lambda_amd import diff
    sd, ids = scene.inverse_room(tex_res=8, medium=True)
    cam = scene.cornell_camera(24, 24)
    r = diff.DiffRenderer(dev, sd, cam, 24, 24, 16, [("texture", ids["albedo"]), ("texture", ids["roughness"]),
                                                    ("density", ids["medium"])])
    S = Setup(dev, sd, cam, 24, 24, 16, 6, 2)
    try:
        tex = [o[1] for o in sd.ops if o[0] == "texture"]
        dens = [o[1]["density"] for o in sd.ops if o[0] == "medium"][0]
        pa = torch.tensor(tex[ids["albedo"]], requires_grad=True)
        pr = torch.tensor(tex[ids["roughness"]][..., 0], requires_grad=True)
        pd = torch.tensor(dens, requires_grad=True)
        img = r.render(pa, pr, pd, seed=5)
        assert img.shape == (24, 24, 3)
        w = torch.tensor(S.w)
        (img * w).sum().backward()
        gt, gd, _ = S.adjoint(5 + 0x9E3779B9)
        np.testing.assert_allclose(pa.grad.numpy().reshape(-1), gt[S.tex_off[0]:S.tex_off[1]], rtol=1e-4, atol=1e-7)
        np.testing.assert_allclose(pr.grad.numpy().reshape(-1), gt[S.tex_off[1]:S.tex_off[2]], rtol=1e-4, atol=1e-7)
        np.testing.assert_allclose(pd.grad.numpy().reshape(-1), gd, rtol=1e-4, atol=1e-7)
        assert pr.grad.shape == pr.shape and pd.grad.shape == pd.shape
    finally:
        r.close()
        S.close()


def test_inverse_rendering_recovers_albedo(dev):
    """Adam on an L2 image loss through the harness (256 spp per iteration, 60 iterations) pulls a
    perturbed albedo texture (+U[-0.2, 0.2]) back towards the one that rendered the target
    (4096 spp): the error of the texels the camera sees well and the image error (the target's
    seed and spp) both drop."""
    import torch
    from lambda_amd import diff
    sd, ids = scene.inverse_room(tex_res=8, medium=False)
    cam = scene.cornell_camera(64, 64)
    true = [o[1] for o in sd.ops if o[0] == "texture"][ids["albedo"]].copy()
    r = diff.DiffRenderer(dev, sd, cam, 64, 64, 4096, [("texture", ids["albedo"])], max_bounces=4, min_bounces=2)
    try:
        target = r.forward_image(1000)
        start = np.clip(true + np.random.default_rng(3).uniform(-0.2, 0.2, size=true.shape), 0.05, 0.95)
        r.upload([torch.tensor(start.astype(np.float32))])
        img0 = r.forward_image(1000)  # the target's seed: the image difference is the parameters'
        g = np.abs(r.adjoint(np.ones((64, 64, 3), np.float32), 7)[0])
        seen = g > 0.05 * g.max()
        err0 = np.abs(start - true)[seen].mean()
        r.spp = 256
        p = torch.tensor(start.astype(np.float32), requires_grad=True)
        opt = torch.optim.Adam([p], lr=0.02)
        tgt = torch.from_numpy(target)
        for it in range(60):
            opt.zero_grad()
            loss = ((r.render(p, seed=it) - tgt) ** 2).mean()
            loss.backward()
            opt.step()
            with torch.no_grad():
                p.clamp_(0.02, 0.98)
        err1 = np.abs(p.detach().numpy() - true)[seen].mean()
        r.spp = 4096
        r.upload([p])
        img1 = r.forward_image(1000)
        m0, m1 = float(((img0 - target) ** 2).mean()), float(((img1 - target) ** 2).mean())
        print("albedo error on %d well-seen entries %.4f -> %.4f, image MSE %.3e -> %.3e"
              % (seen.sum(), err0, err1, m0, m1))
        assert err1 < 0.6 * err0
        assert m1 < 0.3 * m0
    finally:
        r.close()


def test_config4_scale_albedo_gradients_match_finite_differences(dev):
    """The differentiable inverse-rendering workload at full size (50k triangles, 512^2 albedo and
    roughness textures, 64^3 density grid, 512x512 image): 256 random albedo texels whose adjoint
    gradient is non-negligible, central differences at h = 1e-3 of 4-spp renders with Russian
    roulette off (3 bounces)."""
    sd, ids = scene.inverse_scene()
    S = Setup(dev, sd, scene.inverse_camera(512, 512), 512, 512, 4, 3, 3)
    try:
        gt, gd, ge = S.adjoint(seed=21)
        t = ids["albedo"]
        ga = gt[S.tex_off[t]:S.tex_off[t + 1]]
        assert np.count_nonzero(gd) > 0 and np.count_nonzero(gt[S.tex_off[1]:S.tex_off[2]]) > 0
        picks = _pick(ga, 256, np.random.default_rng(8), floor=1e-2)
        assert len(picks) == 256
        fd = [S.fd_texel(t, int(k), 1e-3, 21) for k in picks]
        _check(fd, ga[picks], "config-4 albedo (512^2 textures, 64^3 grid, 512x512)")
    finally:
        S.close()
