"""ctypes binding of liblambda_gpu.so (the C ABI in include/lambda_gpu.h).

The classes mirror the reference's render surface: a Device (lambdaCreateDevice),
a Scene that takes materials/meshes/lights and is committed with a light
strategy (Scene::Commit + LightSampler::Commit), a Film (Film.h) and render()
/ trace_paths() for the render call (MosaicRenderer::Render / TileRenderer).
There is no CPU fallback: loading fails loudly when the HIP library is absent.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblambda_gpu.so")
_LIB = None

EXPORTS = [
    "lgpu_device_create", "lgpu_device_destroy", "lgpu_last_error", "lgpu_abi_version", "lgpu_scene_create",
    "lgpu_scene_destroy", "lgpu_scene_add_texture", "lgpu_scene_add_medium", "lgpu_scene_add_material",
    "lgpu_scene_add_mesh", "lgpu_scene_add_light", "lgpu_scene_commit", "lgpu_scene_info", "lgpu_scene_light_tree",
    "lgpu_camera_init", "lgpu_camera_set_rotation", "lgpu_camera_set_forwards", "lgpu_camera_set_fov",
    "lgpu_film_create", "lgpu_film_destroy", "lgpu_film_clear", "lgpu_film_read", "lgpu_film_read_rgba",
    "lgpu_render", "lgpu_trace_paths", "lgpu_intersect", "lgpu_occluded", "lgpu_philox", "lgpu_copy_bench",
    "lgpu_scene_add_proxy", "lgpu_scene_add_instance", "lgpu_progressive_create", "lgpu_progressive_destroy",
    "lgpu_progressive_set_callback", "lgpu_progressive_start", "lgpu_progressive_stop", "lgpu_progressive_clear",
    "lgpu_progressive_read", "lgpu_progressive_running", "lgpu_obj_load", "lgpu_obj_counts", "lgpu_obj_mesh",
    "lgpu_obj_material_info", "lgpu_obj_destroy", "lgpu_film_post", "lgpu_write_png", "lgpu_film_write", "lgpu_film_denoise",
    "lgpu_render_noalloc", "lgpu_guide_create", "lgpu_guide_destroy", "lgpu_guide_num_params",
    "lgpu_guide_get_params", "lgpu_guide_set_params", "lgpu_guide_forward", "lgpu_guide_gradient",
    "lgpu_guide_train", "lgpu_guide_stats", "lgpu_render_adjoint", "lgpu_scene_update_texture",
    "lgpu_scene_update_density", "lgpu_render_batch",
]


class LambdaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("lgpu status %d: %s" % (status, msg))
        self.status = status


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liblambda_gpu.so is not built (run __graft_entry__.build() or lambda_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        sigs = {
            "lgpu_device_create": [i32, vp], "lgpu_device_destroy": [vp], "lgpu_scene_create": [vp, vp],
            "lgpu_scene_destroy": [vp], "lgpu_scene_add_texture": [vp, vp, vp], "lgpu_scene_add_medium": [vp, vp, vp],
            "lgpu_scene_add_material": [vp, vp, vp], "lgpu_scene_add_mesh": [vp, vp, vp],
            "lgpu_scene_add_light": [vp, vp, vp], "lgpu_scene_commit": [vp, u32],
            "lgpu_scene_info": [vp, vp, vp, vp, vp], "lgpu_scene_light_tree": [vp, vp, u32],
            "lgpu_film_create": [vp, u32, u32, vp], "lgpu_film_destroy": [vp], "lgpu_film_clear": [vp],
            "lgpu_film_read": [vp, vp, vp], "lgpu_film_read_rgba": [vp, vp], "lgpu_render": [vp, vp],
            "lgpu_render_noalloc": [vp, vp], "lgpu_render_adjoint": [vp, vp, vp], "lgpu_render_batch": [vp, u32, vp, vp],
            "lgpu_scene_update_texture": [vp, u32, vp], "lgpu_scene_update_density": [vp, u32, vp],
            "lgpu_guide_create": [vp, vp, vp], "lgpu_guide_destroy": [vp], "lgpu_guide_get_params": [vp, vp],
            "lgpu_guide_set_params": [vp, vp], "lgpu_guide_forward": [vp, u32, vp, vp],
            "lgpu_guide_gradient": [vp, u32, vp, vp, vp], "lgpu_guide_train": [vp, u32, vp],
            "lgpu_guide_stats": [vp, vp, vp, vp],
            "lgpu_trace_paths": [vp, u32, vp, vp, vp], "lgpu_intersect": [vp, u32, vp, vp],
            "lgpu_occluded": [vp, u32, vp, vp], "lgpu_philox": [vp, u32, vp, vp, vp],
            "lgpu_copy_bench": [vp, C.c_uint64, u32, vp],
            "lgpu_scene_add_proxy": [vp, vp, vp], "lgpu_scene_add_instance": [vp, u32, vp, u32, vp],
            "lgpu_progressive_create": [vp, vp], "lgpu_progressive_destroy": [vp],
            "lgpu_progressive_set_callback": [vp, vp, vp], "lgpu_progressive_start": [vp],
            "lgpu_progressive_stop": [vp], "lgpu_progressive_clear": [vp], "lgpu_progressive_read": [vp, vp, vp],
            "lgpu_progressive_running": [vp],
            "lgpu_obj_load": [C.c_char_p, vp], "lgpu_obj_counts": [vp, vp, vp], "lgpu_obj_mesh": [vp, u32, vp, vp],
            "lgpu_obj_material_info": [vp, u32, vp], "lgpu_obj_destroy": [vp],
            "lgpu_film_post": [vp, vp, vp, vp], "lgpu_write_png": [C.c_char_p, vp, u32, u32],
            "lgpu_film_write": [vp, vp, vp], "lgpu_film_denoise": [vp, vp, vp, vp, vp],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.lgpu_last_error.restype = C.c_char_p
        L.lgpu_guide_num_params.restype = u32
        L.lgpu_guide_num_params.argtypes = []
        L.lgpu_abi_version.restype = u32
        L.lgpu_camera_init.argtypes = [vp, u32, vp, f32, f32]
        L.lgpu_camera_set_rotation.argtypes = [vp, f32, f32]
        L.lgpu_camera_set_forwards.argtypes = [vp, vp]
        L.lgpu_camera_set_fov.argtypes = [vp, f32]
        if L.lgpu_abi_version() != abi.ABI_VERSION:
            raise ImportError("liblambda_gpu.so has ABI %d, this binding expects %d: rebuild it"
                              % (L.lgpu_abi_version(), abi.ABI_VERSION))
        _LIB = L
    return _LIB


def _check(status):
    if status != abi.OK:
        raise LambdaError(status, lib().lgpu_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Device:
    """lambdaCreateDevice (Lambda.cpp:148-152): one HIP device and its stream."""

    def __init__(self, index=0):
        h = C.c_void_p()
        _check(lib().lgpu_device_create(index, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lgpu_device_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def copy_bench(self, nbytes, iters=10):
        """Mean time (ms) of a plain device copy of nbytes (read + write)."""
        ms = C.c_double()
        _check(lib().lgpu_copy_bench(self.h, nbytes, iters, C.byref(ms)))
        return ms.value

    def philox(self, counters, key):
        counters = np.ascontiguousarray(counters, dtype=np.uint32).reshape(-1, 4)
        key = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros_like(counters)
        _check(lib().lgpu_philox(self.h, len(counters), _ptr(counters), _ptr(key), _ptr(out)))
        return out


class Scene:
    """Scene (Scene.h:25-108) flattened at commit into BVH, light sampler and bytecode."""

    def __init__(self, device):
        self.device = device
        h = C.c_void_p()
        _check(lib().lgpu_scene_create(device.h, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lgpu_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _add(self, fn, d):
        out = C.c_uint32()
        _check(fn(self.h, C.byref(d), C.byref(out)))
        return out.value

    def add_texture(self, d):
        return self._add(lib().lgpu_scene_add_texture, d)

    def add_medium(self, d):
        return self._add(lib().lgpu_scene_add_medium, d)

    def add_material(self, d):
        return self._add(lib().lgpu_scene_add_material, d)

    def add_mesh(self, d):
        return self._add(lib().lgpu_scene_add_mesh, d)

    def add_light(self, d):
        return self._add(lib().lgpu_scene_add_light, d)

    def add_proxy(self, d):
        return self._add(lib().lgpu_scene_add_proxy, d)

    def add_instance(self, proxy, xfm, material=abi.NONE):
        x = (C.c_float * 12)(*[float(v) for v in xfm])
        out = C.c_uint32()
        _check(lib().lgpu_scene_add_instance(self.h, proxy, x, material, C.byref(out)))
        return out.value

    def commit(self, strategy=abi.LIGHTS_TREE):
        _check(lib().lgpu_scene_commit(self.h, strategy))

    def update_texture(self, tex_id, texels):
        """Replaces a committed texture's texels (w * h * channels floats) in place."""
        t = np.ascontiguousarray(texels, dtype=np.float32)
        _check(lib().lgpu_scene_update_texture(self.h, tex_id, _ptr(t)))

    def update_density(self, medium_id, density):
        """Replaces a committed grid medium's densities (nx * ny * nz floats) in place."""
        g = np.ascontiguousarray(density, dtype=np.float32)
        _check(lib().lgpu_scene_update_density(self.h, medium_id, _ptr(g)))

    def info(self):
        v = [C.c_uint32() for _ in range(4)]
        _check(lib().lgpu_scene_info(self.h, *[C.byref(x) for x in v]))
        return dict(triangles=v[0].value, bvh_nodes=v[1].value, lights=v[2].value, tree_nodes=v[3].value)

    def light_tree(self):
        n = self.info()["tree_nodes"]
        out = np.zeros((max(n, 1), 15), dtype=np.float32)
        _check(lib().lgpu_scene_light_tree(self.h, _ptr(out), n))
        return out[:n]

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        hits = np.zeros((len(rays), 4), dtype=np.uint32)
        _check(lib().lgpu_intersect(self.h, len(rays), _ptr(rays), _ptr(hits)))
        return hits

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.zeros(len(rays), dtype=np.uint32)
        _check(lib().lgpu_occluded(self.h, len(rays), _ptr(rays), _ptr(occ)))
        return occ


class Film:
    """Film (Film.h:19-45) with FP64 per-pixel sums."""

    def __init__(self, device, width, height):
        self.device, self.width, self.height = device, width, height
        h = C.c_void_p()
        _check(lib().lgpu_film_create(device.h, width, height, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lgpu_film_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def clear(self):
        _check(lib().lgpu_film_clear(self.h))

    def read(self):
        rgb = np.zeros((self.height, self.width, 3), dtype=np.float64)
        cnt = np.zeros((self.height, self.width), dtype=np.uint32)
        _check(lib().lgpu_film_read(self.h, _ptr(rgb), _ptr(cnt)))
        return rgb, cnt

    def write(self, rgb, counts=None):
        """Sets the film's sums (H x W x 3 float64) and counts (H x W, default 1) (lgpu_film_write)."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float64)
        if rgb.shape != (self.height, self.width, 3):
            raise ValueError("film write: expected shape %s" % ((self.height, self.width, 3),))
        cnt = None if counts is None else np.ascontiguousarray(counts, dtype=np.uint32)
        if cnt is not None and cnt.shape != (self.height, self.width):
            raise ValueError("film write: counts must be H x W")
        _check(lib().lgpu_film_write(self.h, _ptr(rgb), _ptr(cnt) if cnt is not None else None))

    def denoise(self, desc, albedo=None, normal=None, out=None):
        """Edge-avoiding A-Trous denoise of this film into `out` (default: this film) (lgpu_film_denoise)."""
        o = out if out is not None else self
        _check(lib().lgpu_film_denoise(self.h, albedo.h if albedo is not None else None,
                                       normal.h if normal is not None else None, C.byref(desc), o.h))
        return o

    def post(self, desc):
        """Post-processed image: (RGBA32F, RGBA8) of the film (lgpu_film_post)."""
        f32 = np.zeros((self.height, self.width, 4), dtype=np.float32)
        u8 = np.zeros((self.height, self.width, 4), dtype=np.uint8)
        _check(lib().lgpu_film_post(self.h, C.byref(desc), _ptr(f32), _ptr(u8)))
        return f32, u8

    def read_rgba(self):
        out = np.zeros((self.height, self.width, 4), dtype=np.float32)
        _check(lib().lgpu_film_read_rgba(self.h, _ptr(out)))
        return out


def write_png(path, rgba8):
    rgba8 = np.ascontiguousarray(rgba8, dtype=np.uint8)
    h, w = rgba8.shape[:2]
    _check(lib().lgpu_write_png(str(path).encode(), _ptr(rgba8), w, h))


def load_obj(path):
    """OBJ ingestion through the library (host only): ([mesh dict], [material dict]). Mesh arrays
    are copied out; 'material' indexes the material list."""
    L = lib()
    h = C.c_void_p()
    _check(L.lgpu_obj_load(str(path).encode(), C.byref(h)))
    try:
        nm, nmat = C.c_uint32(), C.c_uint32()
        _check(L.lgpu_obj_counts(h, C.byref(nm), C.byref(nmat)))
        meshes = []
        for i in range(nm.value):
            d = abi.MeshDesc()
            name = C.c_char_p()
            _check(L.lgpu_obj_mesh(h, i, C.byref(d), C.byref(name)))
            nv, nt = d.num_vertices, d.num_triangles

            def arr(p, n):
                return None if not p else np.ctypeslib.as_array(p, shape=(n,)).copy()
            m = dict(name=name.value.decode(), material=d.material,
                     positions=arr(d.positions, 3 * nv).reshape(-1, 3), indices=arr(d.indices, 3 * nt).reshape(-1, 3))
            for k, w in (("normals", 3), ("tangents", 4), ("uvs", 2)):
                a = arr(getattr(d, k), w * nv)
                m[k] = None if a is None else a.reshape(-1, w)
            meshes.append(m)
        mats = []
        for i in range(nmat.value):
            om = abi.ObjMaterial()
            _check(L.lgpu_obj_material_info(h, i, C.byref(om)))
            mats.append(dict(name=om.name.decode(), kd=tuple(om.kd) if om.has_kd else None,
                             ke=tuple(om.ke) if om.has_ke else None))
        return meshes, mats
    finally:
        L.lgpu_obj_destroy(h)


_CALLBACK = C.CFUNCTYPE(None, C.c_void_p)


class Progressive:
    """Progressive renderer (ProgressiveRender.cpp:13-62): a native worker thread renders passes of
    desc's spp-range size into desc's film; read() copies the latest completed RGBA32F image."""

    def __init__(self, desc):
        self.desc = desc  # keeps the film and scene objects referenced by the caller alive
        self.width, self.height = None, None
        h = C.c_void_p()
        _check(lib().lgpu_progressive_create(C.byref(desc), C.byref(h)))
        self.h = h
        self._cb = None

    def set_callback(self, fn):
        """fn() runs on the worker thread after each pass."""
        self._cb = _CALLBACK(lambda user: fn())
        _check(lib().lgpu_progressive_set_callback(self.h, C.cast(self._cb, C.c_void_p), None))

    def start(self):
        _check(lib().lgpu_progressive_start(self.h))

    def stop(self):
        _check(lib().lgpu_progressive_stop(self.h))

    def clear(self):
        _check(lib().lgpu_progressive_clear(self.h))

    def running(self):
        return bool(lib().lgpu_progressive_running(self.h))

    def read(self, width, height):
        out = np.zeros((height, width, 4), dtype=np.float32)
        n = C.c_uint32()
        _check(lib().lgpu_progressive_read(self.h, _ptr(out), C.byref(n)))
        return out, n.value

    def close(self):
        if getattr(self, "h", None):
            lib().lgpu_progressive_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Guide:
    """Learned path guiding network (lgpu_guide_*): created on a device, passed to render_desc."""

    def __init__(self, dev, seed=0, record_capacity=0, batch=0, learning_rate=0.0, max_depth=0):
        d = abi.GuideDesc(seed, record_capacity, batch, learning_rate, max_depth)
        h = C.c_void_p()
        _check(lib().lgpu_guide_create(dev.h, C.byref(d), C.byref(h)))
        self.h = h
        self.dev = dev

    def close(self):
        if getattr(self, "h", None):
            lib().lgpu_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def num_params():
        return int(lib().lgpu_guide_num_params())

    def params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().lgpu_guide_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params()
        _check(lib().lgpu_guide_set_params(self.h, _ptr(p)))

    def forward(self, inputs):
        """(n, 5) inputs in [0, 1] -> (n, 129): bin probabilities and the mixing probability."""
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 5)
        out = np.zeros((len(x), 129), dtype=np.float32)
        _check(lib().lgpu_guide_forward(self.h, len(x), _ptr(x), _ptr(out)))
        return out

    def gradient(self, records):
        """(n, 8) records (5 inputs, bin, BSDF pdf, weight) -> (gradient, loss)."""
        r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 8)
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = C.c_float()
        _check(lib().lgpu_guide_gradient(self.h, len(r), _ptr(r), _ptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, steps):
        loss = C.c_float()
        _check(lib().lgpu_guide_train(self.h, steps, C.byref(loss)))
        return loss.value

    def stats(self):
        rec, steps, loss = C.c_uint32(), C.c_uint32(), C.c_float()
        _check(lib().lgpu_guide_stats(self.h, C.byref(rec), C.byref(steps), C.byref(loss)))
        return {"records": rec.value, "steps": steps.value, "loss": loss.value}


def render_desc(scene, film, camera, spp_begin=0, spp_end=1, max_bounces=8, min_bounces=3,
                spectral=abi.SPECTRAL_RGB, estimator=abi.ESTIMATOR_CORRECTED, seed=0, capacity=0, flags=0,
                integrator=abi.INTEGRATOR_PATH, max_depth=1.0, depth_power=1.0, background=(1.0, 0.0, 1.0),
                geometric_normal=False, guide=None, guide_train_steps=0):
    """lgpu_render_desc. The utility-pass fields default to the reference's DepthPass/NormalPass
    defaults (UtilityIntegrators.h:18-30)."""
    d = abi.RenderDesc()
    d.scene = scene.h.value if scene is not None and hasattr(scene, "h") else None
    d.film = film.h.value if film is not None else None
    d.camera = camera
    d.integrator = integrator
    d.spp_begin, d.spp_end = spp_begin, spp_end
    d.max_bounces, d.min_bounces = max_bounces, min_bounces
    d.spectral_mode, d.estimator_mode = spectral, estimator
    d.seed, d.path_capacity, d.flags = seed, capacity, flags
    d.pass_max_depth, d.pass_power = max_depth, depth_power
    d.pass_background[:] = [float(x) for x in background]
    d.pass_geometric_normal = 1 if geometric_normal else 0
    d.guide = guide.h.value if guide is not None else None
    d.guide_train_steps = guide_train_steps
    return d


def render(desc):
    """The render call (MosaicRenderer::Render): accumulates spp [begin, end) into the film."""
    st = abi.RenderStats()
    _check(lib().lgpu_render(C.byref(desc), C.byref(st)))
    return st


def render_batch(descs, path_capacity=0, autotune=False, branches=0, max_iterations=0):
    """Renders every job (render_desc) together from one path pool, one hipGraph per chunk."""
    arr = (abi.RenderDesc * len(descs))(*descs)
    o = abi.BatchOptions()
    o.path_capacity = path_capacity
    o.flags = abi.BATCH_AUTOTUNE if autotune else 0
    o.max_iterations = max_iterations
    o.branches = branches
    st = abi.BatchStats()
    _check(lib().lgpu_render_batch(arr, len(descs), C.byref(o), C.byref(st)))
    return st


def render_adjoint(desc, d_image, n_texels=0, n_density=0, n_materials=0):
    """Path-replay adjoint render: accumulates the primal samples into desc's film and returns the
    gradients of sum(d_image * film mean of this call's samples) as float32 arrays
    (textures in id order, grid densities in medium id order, per-material emission intensity;
    None where the count is 0), and the render stats."""
    d_image = np.ascontiguousarray(d_image, dtype=np.float32)
    gt = np.zeros(n_texels, dtype=np.float32) if n_texels else None
    gd = np.zeros(n_density, dtype=np.float32) if n_density else None
    ge = np.zeros(n_materials, dtype=np.float32) if n_materials else None
    a = abi.AdjointDesc()
    a.d_image = d_image.ctypes.data
    a.grad_textures = gt.ctypes.data if gt is not None else None
    a.grad_density = gd.ctypes.data if gd is not None else None
    a.grad_emission = ge.ctypes.data if ge is not None else None
    st = abi.RenderStats()
    _check(lib().lgpu_render_adjoint(C.byref(desc), C.byref(a), C.byref(st)))
    return gt, gd, ge, st


def trace_paths(desc, pixels, samples):
    """Per-path radiance (5 floats per path) for explicit (pixel, sample) pairs."""
    pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
    samples = np.ascontiguousarray(samples, dtype=np.uint32)
    out = np.zeros((len(pixels), 5), dtype=np.float32)
    _check(lib().lgpu_trace_paths(C.byref(desc), len(pixels), _ptr(pixels), _ptr(samples), _ptr(out)))
    return out
