"""Batched renders (lgpu_render_batch): many scenes from one path pool, one hipGraph per chunk.

Bar: each job's film equals lgpu_render of the same job (per-path radiance is identical, the FP64
film sums group the samples differently: rtol 1e-9), sample counts exactly, for every pool size
(one chunk, many chunks), for 1 to 8 graph branches and with the autotuner's choice.
"""
import numpy as np
import pytest

from lambda_amd import abi, gpu, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = gpu.Device(0)
    yield d
    d.close()


def _jobs(dev):
    """Uneven jobs: surfaces (RGB and hero, with and without shader programs) and media, each a
    configuration the single-render parity and media tests cover."""
    specs = [
        (scene.cornell_box(), scene.cornell_camera(40, 32), 40, 32, 12, abi.INTEGRATOR_PATH, abi.SPECTRAL_HERO4, 8),
        (scene.cornell_box(), scene.cornell_camera(64, 64), 64, 64, 4, abi.INTEGRATOR_PATH, abi.SPECTRAL_RGB, 6),
        (scene.procedural_cornell(), scene.cornell_camera(48, 48), 48, 48, 8, abi.INTEGRATOR_PATH, abi.SPECTRAL_HERO4, 8),
        (scene.fog_cornell("grid"), scene.cornell_camera(32, 32), 32, 32, 16, abi.INTEGRATOR_VOLPATH,
         abi.SPECTRAL_HERO4, 8),
        (scene.fog_cornell("homogeneous"), scene.cornell_camera(24, 40), 24, 40, 10, abi.INTEGRATOR_MIS_VOLPATH,
         abi.SPECTRAL_RGB, 8),
        (scene.inverse_room(tex_res=8, medium=True)[0], scene.cornell_camera(32, 24), 32, 24, 20, abi.INTEGRATOR_VOLPATH,
         abi.SPECTRAL_RGB, 6),
        (scene.lobe_zoo(), scene.cornell_camera(16, 16), 16, 16, 64, abi.INTEGRATOR_PATH, abi.SPECTRAL_HERO4, 8),
    ]
    out = []
    for k, (sd, cam, W, H, spp, integ, spectral, maxb) in enumerate(specs):
        s = gpu.Scene(dev)
        keep = sd.push(s)
        out.append(dict(scene=s, keep=keep, cam=cam, W=W, H=H, spp=spp, integ=integ, spectral=spectral, maxb=maxb,
                        seed=17 + k))
    return out


def _desc(j, film):
    return gpu.render_desc(j["scene"], film, j["cam"], 3, 3 + j["spp"], max_bounces=j["maxb"], min_bounces=2,
                           spectral=j["spectral"], estimator=abi.ESTIMATOR_CORRECTED, seed=j["seed"],
                           integrator=j["integ"])


@pytest.fixture(scope="module")
def jobs(dev):
    js = _jobs(dev)
    for j in js:  # the reference: each job on its own
        f = gpu.Film(dev, j["W"], j["H"])
        gpu.render(_desc(j, f))
        j["ref"] = f.read()
        f.close()
    yield js
    for j in js:
        j["scene"].close()


def _batch(dev, jobs, **kw):
    films = [gpu.Film(dev, j["W"], j["H"]) for j in jobs]
    st = gpu.render_batch([_desc(j, f) for j, f in zip(jobs, films)], **kw)
    out = [f.read() for f in films]
    for f in films:
        f.close()
    return out, st


@pytest.mark.parametrize("capacity,branches", [(0, 0), (4096, 4), (1 << 16, 1)])
def test_batch_equals_individual_renders(dev, jobs, capacity, branches):
    out, st = _batch(dev, jobs, path_capacity=capacity, branches=branches)
    print("batch: %d jobs, %d samples, %d chunks, %d graph nodes, capacity %d, %d branches, %.1f ms (capture %.1f)"
          % (st.jobs, st.samples, st.chunks, st.graph_nodes, st.path_capacity, st.branches, st.ms_render,
             st.ms_capture))
    assert st.jobs == len(jobs) and st.samples == sum(j["W"] * j["H"] * j["spp"] for j in jobs)
    if capacity == 4096:
        assert st.chunks > 1
    for j, (rgb, cnt) in zip(jobs, out):
        np.testing.assert_array_equal(cnt, j["ref"][1])
        np.testing.assert_allclose(rgb, j["ref"][0], rtol=1e-9, atol=1e-12)


def test_batch_autotune_picks_a_configuration(dev, jobs):
    out, st = _batch(dev, jobs, autotune=True)
    print("autotuned: capacity %d branches %d, trials %.1f ms, render %.1f ms"
          % (st.path_capacity, st.branches, st.ms_autotune, st.ms_render))
    assert st.ms_autotune > 0
    for j, (rgb, cnt) in zip(jobs, out):
        np.testing.assert_array_equal(cnt, j["ref"][1])
        np.testing.assert_allclose(rgb, j["ref"][0], rtol=1e-9, atol=1e-12)
    _, st2 = _batch(dev, jobs, autotune=True)  # cached choice: no new trials
    assert st2.ms_autotune == 0 and st2.path_capacity == st.path_capacity


def test_batch_reports_paths_beyond_the_captured_iterations(dev, jobs):
    with pytest.raises(gpu.LambdaError) as e:
        _batch(dev, jobs[:2], max_iterations=2)
    assert e.value.status == abi.ERR_STATE


def test_batch_rejects_shared_films_and_unsupported_jobs(dev, jobs):
    f = gpu.Film(dev, jobs[0]["W"], jobs[0]["H"])
    try:
        with pytest.raises(gpu.LambdaError) as e:
            gpu.render_batch([_desc(jobs[0], f), _desc(jobs[0], f)])
        assert e.value.status == abi.ERR_INVALID_ARGUMENT
        d = _desc(jobs[0], f)
        d.estimator_mode = abi.ESTIMATOR_FAITHFUL
        with pytest.raises(gpu.LambdaError) as e:
            gpu.render_batch([d])
        assert e.value.status == abi.ERR_UNSUPPORTED
    finally:
        f.close()
