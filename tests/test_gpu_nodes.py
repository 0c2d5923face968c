"""GPU parity of the shader-graph nodes (procedural textures, converters, spectral inputs,
surface information, Schlick Fresnel) through the C ABI against the CPU checker: a Cornell
room whose every surface and both emitters use them, per path, under the path and
direct-lighting integrators, and inside fog (3-D texture variants at medium interactions)
under both volumetric integrators. Bar as for the other parity tests: per-path radiance within
1e-4 relative for at least 99.9 % of paths."""
import numpy as np
import pytest

from lambda_amd import abi, gpu, scene
from tests import helpers

pytestmark = pytest.mark.gpu

REL, ABS = 1e-4, 1e-6


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def _fog(sd):
    med = scene.homogeneous(sd, [0.1, 0.15, 0.25], [0.5, 0.4, 0.3], g=0.2)
    fog = sd.add_material(scene.passthrough(interior=med))
    v, i = scene.aabb((-0.95, -0.95, -0.95), (0.95, 0.9, 0.95))
    sd.add_mesh(v, i, fog)
    return sd


def _parity(dev, sd, integrator, spectral, estimator, W=48, H=48, spp=2, max_bounces=8):
    s = gpu.Scene(dev)
    k1 = sd.push(s, abi.LIGHTS_TREE)
    o, k2 = helpers.oracle_scene(sd, abi.LIGHTS_TREE)
    cam = scene.cornell_camera(W, H)
    film = gpu.Film(dev, W, H)
    pixels = np.tile(np.arange(W * H, dtype=np.uint32), spp)
    samples = np.repeat(np.arange(spp, dtype=np.uint32), W * H)
    d = gpu.render_desc(s, film, cam, 0, spp, max_bounces=max_bounces, min_bounces=3, spectral=spectral,
                        estimator=estimator, integrator=integrator)
    g = gpu.trace_paths(d, pixels, samples)
    cd = helpers.desc(cam, 0, spp, max_bounces, 3, spectral, estimator)
    cd.integrator = integrator
    c = o.trace_paths(cd, W, H, pixels, samples)
    ok = helpers.rel_close(g[:, :4], c[:, :4], REL, ABS).all(axis=1) & (g[:, 4] == c[:, 4])
    return g, c, ok


SURFACE = [(i, sp, e) for i in (abi.INTEGRATOR_PATH, abi.INTEGRATOR_DIRECT)
           for sp in (abi.SPECTRAL_RGB, abi.SPECTRAL_HERO4)
           for e in (abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED)]


@pytest.mark.parametrize("integrator,spectral,estimator", SURFACE)
def test_procedural_cornell_per_path_parity(dev, integrator, spectral, estimator):
    g, c, ok = _parity(dev, scene.procedural_cornell(), integrator, spectral, estimator)
    print(integrator, spectral, estimator, "matching", ok.mean(), "nonzero",
          (np.abs(c[:, :3]).sum(axis=1) > 0).mean())
    assert np.isfinite(g).all()
    assert (np.abs(c[:, :3]).sum(axis=1) > 0).mean() > 0.5
    assert ok.mean() >= 0.999


VOLUME = [(i, e) for i in (abi.INTEGRATOR_VOLPATH, abi.INTEGRATOR_MIS_VOLPATH)
          for e in (abi.ESTIMATOR_FAITHFUL, abi.ESTIMATOR_CORRECTED)]


@pytest.mark.parametrize("integrator,estimator", VOLUME)
def test_procedural_fog_per_path_parity(dev, integrator, estimator):
    g, c, ok = _parity(dev, _fog(scene.procedural_cornell()), integrator, abi.SPECTRAL_HERO4, estimator)
    print(integrator, estimator, "matching", ok.mean(), "nonzero", (np.abs(c[:, :3]).sum(axis=1) > 0).mean())
    assert np.isfinite(g).all()
    assert ok.mean() >= 0.999
