"""ctypes binding of the CPU checker (oracle/oracle.cpp).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.
"""
import ctypes as C
import os

import numpy as np

from lambda_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        L.oracle_scene_create.restype = vp
        L.oracle_scene_destroy.argtypes = [vp]
        for n in ("texture", "medium", "material", "mesh", "light"):
            getattr(L, "oracle_scene_add_" + n).argtypes = [vp, vp]
        L.oracle_scene_commit.argtypes = [vp, u32]
        L.oracle_light_tree.argtypes = [vp, vp, u32]
        L.oracle_trace_paths.argtypes = [vp, vp, u32, u32, u32, vp, vp, vp, C.c_int]
        L.oracle_render.argtypes = [vp, vp, u32, u32, u32, u32, u32, u32, C.c_int, vp, vp]
        L.oracle_intersect.argtypes = [vp, u32, vp, vp]
        L.oracle_occluded.argtypes = [vp, u32, vp, vp]
        L.oracle_philox.argtypes = [vp, vp, vp]
        L.oracle_fr_dielectric.argtypes = [f32, f32, f32]
        L.oracle_fr_dielectric.restype = f32
        L.oracle_fr_conductor.argtypes = [f32, vp, vp, vp, vp]
        L.oracle_microfacet_d.argtypes = [C.c_int, vp, f32, f32]
        L.oracle_microfacet_d.restype = f32
        L.oracle_microfacet_lambda.argtypes = [C.c_int, vp, f32, f32]
        L.oracle_microfacet_lambda.restype = f32
        L.oracle_blackbody_rgb.argtypes = [f32, vp]
        L.oracle_blackbody_normalised.argtypes = [f32, f32]
        L.oracle_blackbody_normalised.restype = f32
        L.oracle_cmf.argtypes = [f32, vp]
        L.oracle_uplift.argtypes = [vp, f32]
        L.oracle_uplift.restype = f32
        L.oracle_hg_p.argtypes = [f32, f32]
        L.oracle_hg_p.restype = f32
        L.oracle_piecewise1d_sample.argtypes = [vp, u32, f32, vp]
        L.oracle_piecewise1d_sample.restype = u32
        L.oracle_medium_tr.argtypes = [vp, u32, vp, vp, f32, C.c_int, vp, vp]
        L.oracle_medium_sample_distance.argtypes = [vp, u32, vp, vp, f32, f32, f32, C.c_int, vp, vp]
        L.oracle_medium_pdf_distance.argtypes = [vp, u32, vp, vp, f32, C.c_int, vp, C.c_int]
        L.oracle_medium_pdf_distance.restype = f32
        L.oracle_medium_sample_equiangular.argtypes = [vp, u32, vp, vp, vp, f32, vp, f32, C.c_int, vp, C.c_int, vp]
        L.oracle_medium_pdf_equiangular.argtypes = [vp, vp, vp, f32, f32, C.c_int]
        L.oracle_medium_pdf_equiangular.restype = f32
        L.oracle_light_pdfs.argtypes = [vp, vp, vp, C.c_int, vp]
        L.oracle_light_sample.argtypes = [vp, vp, vp, C.c_int, u32, u32, vp]
        L.oracle_hg_sample.argtypes = [f32, vp, f32, f32, C.c_int, vp]
        L.oracle_hg_sample.restype = f32
        L.oracle_eval_node.argtypes = [vp, u32, u32, vp, C.c_int, vp, vp]
        L.oracle_post.argtypes = [vp, u32, u32, vp, vp, vp]
        L.oracle_post.restype = None
        L.oracle_denoise.argtypes = [u32, u32, vp, vp, vp, vp, vp]
        L.oracle_denoise.restype = None
        L.oracle_scene_add_proxy.argtypes = [vp, vp]
        L.oracle_scene_add_instance.argtypes = [vp, u32, vp, u32]
        L.oracle_bxdf_sample_n.argtypes = [vp, u32, vp, C.c_int, vp, C.c_int, u32, u32, f32, vp]
        L.oracle_bxdf_eval_n.argtypes = [vp, u32, u32, vp, vp, C.c_int, vp, C.c_int, vp]
        L.oracle_generate_ray_n.argtypes = [vp, u32, vp, u32, C.c_int, vp]
        L.oracle_light_sample_li_n.argtypes = [vp, u32, vp, vp, u32, C.c_int, vp, C.c_int, u32, vp]
        L.oracle_env_pdf_n.argtypes = [vp, u32, u32, vp, vp]
        L.oracle_guide_num_params.restype = C.c_int
        L.oracle_guide_forward.argtypes = [vp, u32, vp, vp]
        L.oracle_guide_forward.restype = None
        L.oracle_guide_gradient.argtypes = [vp, u32, vp, C.c_int, vp]
        L.oracle_guide_gradient.restype = f32
        L.oracle_guide_loss64.argtypes = [vp, u32, vp]
        L.oracle_guide_loss64.restype = C.c_double
        L.oracle_tree_importance.argtypes = [vp, vp, vp, C.c_int]
        L.oracle_tree_importance.restype = f32
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleScene:
    """Same add_* / commit surface as the GPU scene so SceneDesc.push() drives both."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.oracle_scene_create()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_scene_destroy(self.h)
            self.h = None

    def _chk(self, r, what):
        if r < 0:
            raise RuntimeError("oracle %s failed (%d)" % (what, r))
        return r

    def add_texture(self, d):
        return self._chk(self.L.oracle_scene_add_texture(self.h, C.byref(d)), "add_texture")

    def add_medium(self, d):
        return self._chk(self.L.oracle_scene_add_medium(self.h, C.byref(d)), "add_medium")

    def add_material(self, d):
        return self._chk(self.L.oracle_scene_add_material(self.h, C.byref(d)), "add_material")

    def add_mesh(self, d):
        return self._chk(self.L.oracle_scene_add_mesh(self.h, C.byref(d)), "add_mesh")

    def add_proxy(self, d):
        return self._chk(self.L.oracle_scene_add_proxy(self.h, C.byref(d)), "add_proxy")

    def add_instance(self, proxy, xfm, material=0xFFFFFFFF):
        x = (C.c_float * 12)(*[float(v) for v in xfm])
        return self._chk(self.L.oracle_scene_add_instance(self.h, proxy, x, material), "add_instance")

    def add_light(self, d):
        return self._chk(self.L.oracle_scene_add_light(self.h, C.byref(d)), "add_light")

    def commit(self, strategy):
        return self._chk(self.L.oracle_scene_commit(self.h, strategy), "commit")

    def light_tree(self, max_nodes=1 << 16):
        out = np.zeros((max_nodes, 15), dtype=np.float32)
        n = self.L.oracle_light_tree(self.h, _ptr(out), max_nodes)
        return out[:n]

    def trace_paths(self, desc, width, height, pixels, samples, threads=8):
        pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
        samples = np.ascontiguousarray(samples, dtype=np.uint32)
        out = np.zeros((len(pixels), 5), dtype=np.float32)
        self._chk(self.L.oracle_trace_paths(self.h, C.byref(desc), width, height, len(pixels), _ptr(pixels),
                                            _ptr(samples), _ptr(out), threads), "trace_paths")
        return out

    def render(self, desc, width, height, window=None, threads=8):
        x0, y0, w, h = window if window else (0, 0, width, height)
        rgb = np.zeros((h, w, 3), dtype=np.float64)
        cnt = np.zeros((h, w), dtype=np.uint32)
        self._chk(self.L.oracle_render(self.h, C.byref(desc), width, height, x0, y0, w, h, threads, _ptr(rgb),
                                       _ptr(cnt)), "render")
        return rgb, cnt

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        hits = np.zeros((len(rays), 4), dtype=np.uint32)
        self._chk(self.L.oracle_intersect(self.h, len(rays), _ptr(rays), _ptr(hits)), "intersect")
        return hits

    def light_pdfs(self, p, n, corrected=True, max_lights=1 << 16):
        out = np.zeros(max_lights + 1, dtype=np.float32)
        p, n = _v3(p), _v3(n)
        k = self._chk(self.L.oracle_light_pdfs(self.h, _ptr(p), _ptr(n), int(corrected), _ptr(out)), "light_pdfs")
        return out[:k], float(out[k])

    def light_sample(self, p, n, pixel, sample, corrected=True):
        pdf = np.zeros(1, dtype=np.float32)
        p, n = _v3(p), _v3(n)
        idx = self.L.oracle_light_sample(self.h, _ptr(p), _ptr(n), int(corrected), pixel, sample, _ptr(pdf))
        return idx, float(pdf[0])

    def eval_node(self, material, node, uv=(0, 0), point=(0, 0, 0), normal=(0, 1, 0), eta=1.0, wo_y=1.0,
                  medium=False, mode=0, lambdas=None):
        """Value of one shader node at an event: (type, 4 floats)."""
        ev = np.array([*uv, *point, *normal, eta, wo_y, 1.0 if medium else 0.0], dtype=np.float32)
        out = np.zeros(4, dtype=np.float32)
        lam = np.asarray(lambdas if lambdas is not None else (0, 0, 0, 0), dtype=np.float32)
        t = self.L.oracle_eval_node(self.h, material, node, _ptr(ev), mode, _ptr(lam), _ptr(out))
        if t < 0:
            raise ValueError("eval_node: bad material or node")
        return t, out

    def bxdf_sample(self, material, wo, n, pixel=0, mode=0, corrected=True, lambdas=None, eta=1.001):
        """n samples of a material's BxDF in the world frame (normal +y) for local direction wo:
        dict of wi (n, 3), pdf, f (n, 4), spec, side (spawn offset sign along the normal), eta."""
        out = np.zeros((n, 11), dtype=np.float32)
        w = _v3(wo)
        lam = np.asarray(lambdas if lambdas is not None else (0, 0, 0, 0), dtype=np.float32)
        self._chk(self.L.oracle_bxdf_sample_n(self.h, material, _ptr(w), mode, _ptr(lam), int(corrected), pixel, n,
                                              eta, _ptr(out)), "bxdf_sample")
        return dict(wi=out[:, 0:3], pdf=out[:, 3], f=out[:, 4:8], spec=out[:, 8] != 0, side=out[:, 9], eta=out[:, 10])

    def bxdf_eval(self, material, wo, wi, mode=0, corrected=True, lambdas=None):
        """f (n, 4) and pdf (n,) of a material's BxDF for pairs of local directions."""
        wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, dtype=np.float32), np.shape(wi)))
        wi = np.ascontiguousarray(np.asarray(wi, dtype=np.float32))
        n = len(wi)
        out = np.zeros((n, 5), dtype=np.float32)
        lam = np.asarray(lambdas if lambdas is not None else (0, 0, 0, 0), dtype=np.float32)
        self._chk(self.L.oracle_bxdf_eval_n(self.h, material, n, _ptr(wo), _ptr(wi), mode, _ptr(lam), int(corrected),
                                            _ptr(out)), "bxdf_eval")
        return out[:, :4], out[:, 4]

    def light_sample_li(self, light, p, n, count, pixel=0, mode=0, corrected=True, lambdas=None):
        """count draws of Light::Sample_Li of scene light `light` at (p, n): wi (count, 3), pdf, Li (count, 4)."""
        out = np.zeros((count, 8), dtype=np.float32)
        lam = np.asarray(lambdas if lambdas is not None else (0, 0, 0, 0), dtype=np.float32)
        p, n = _v3(p), _v3(n)
        self._chk(self.L.oracle_light_sample_li_n(self.h, light, _ptr(p), _ptr(n), count, mode, _ptr(lam),
                                                  int(corrected), pixel, _ptr(out)), "light_sample_li")
        return out[:, 0:3], out[:, 3], out[:, 4:8]

    def env_pdf(self, light, wi):
        wi = np.ascontiguousarray(np.asarray(wi, dtype=np.float32).reshape(-1, 3))
        out = np.zeros(len(wi), dtype=np.float32)
        self._chk(self.L.oracle_env_pdf_n(self.h, light, len(wi), _ptr(wi), _ptr(out)), "env_pdf")
        return out

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.zeros(len(rays), dtype=np.uint32)
        self._chk(self.L.oracle_occluded(self.h, len(rays), _ptr(rays), _ptr(occ)), "occluded")
        return occ


def _v3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(3))


class Medium:
    """Closed-form access to one medium of an OracleScene (mode 0 RGB, 1 HERO4)."""

    def __init__(self, scene, med, mode=0, lambdas=(0, 0, 0, 0)):
        self.s, self.med, self.mode = scene, med, mode
        self.lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float32))

    def tr(self, o, d, t_far):
        out = np.zeros(4, dtype=np.float32)
        o, d = _v3(o), _v3(d)
        lib().oracle_medium_tr(self.s.h, self.med, _ptr(o), _ptr(d), t_far, self.mode, _ptr(self.lam), _ptr(out))
        return out

    def sample_distance(self, o, d, t_far, u0, u1):
        out = np.zeros(7, dtype=np.float32)
        o, d = _v3(o), _v3(d)
        lib().oracle_medium_sample_distance(self.s.h, self.med, _ptr(o), _ptr(d), t_far, u0, u1, self.mode,
                                            _ptr(self.lam), _ptr(out))
        return out

    def pdf_distance(self, o, d, t, corrected=True):
        o, d = _v3(o), _v3(d)
        return lib().oracle_medium_pdf_distance(self.s.h, self.med, _ptr(o), _ptr(d), t, self.mode, _ptr(self.lam),
                                                int(corrected))

    def sample_equiangular(self, o, d, hit_point, t_far, light, eps, corrected=True):
        out = np.zeros(7, dtype=np.float32)
        o, d, hp, lp = _v3(o), _v3(d), _v3(hit_point), _v3(light)
        lib().oracle_medium_sample_equiangular(self.s.h, self.med, _ptr(o), _ptr(d), _ptr(hp), t_far, _ptr(lp), eps,
                                               self.mode, _ptr(self.lam), int(corrected), _ptr(out))
        return out


def pdf_equiangular(o, d, light, t_far, t, corrected=True):
    o, d, lp = _v3(o), _v3(d), _v3(light)
    return lib().oracle_medium_pdf_equiangular(_ptr(o), _ptr(d), _ptr(lp), t_far, t, int(corrected))


def hg_sample(g, wo, u0, u1, corrected=True):
    wi = np.zeros(3, dtype=np.float32)
    wo = _v3(wo)
    p = lib().oracle_hg_sample(g, _ptr(wo), u0, u1, int(corrected), _ptr(wi))
    return p, wi


def post(rgba, desc):
    """Post-processing of an RGBA32F image (h, w, 4): (RGBA32F, RGBA8)."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[:2]
    o32 = np.zeros((h, w, 4), dtype=np.float32)
    o8 = np.zeros((h, w, 4), dtype=np.uint8)
    lib().oracle_post(_ptr(rgba), w, h, C.byref(desc), _ptr(o32), _ptr(o8))
    return o32, o8


def denoise(colour, desc, albedo=None, normal=None):
    """Edge-avoiding A-Trous filter of float32 (h, w, 3) film means; returns (h, w, 3)."""
    c = np.ascontiguousarray(colour, dtype=np.float32)
    h, w = c.shape[:2]
    a = None if albedo is None else np.ascontiguousarray(albedo, dtype=np.float32)
    n = None if normal is None else np.ascontiguousarray(normal, dtype=np.float32)
    out = np.zeros((h, w, 3), dtype=np.float32)
    lib().oracle_denoise(w, h, _ptr(c), _ptr(a) if a is not None else None, _ptr(n) if n is not None else None,
                         C.byref(desc), _ptr(out))
    return out


def tree_importance(center, diag, axis, theta_o, theta_e, power, p, n, corrected=True):
    """Light-tree node importance at point p with shading normal n (zero normal: medium point)."""
    node = np.array([*center, diag, *axis, theta_o, theta_e, power], dtype=np.float32)
    return float(lib().oracle_tree_importance(_ptr(node), _ptr(_v3(p)), _ptr(_v3(n)), int(corrected)))


def generate_rays(cam, uv, pixel=0, corrected=True):
    """Camera::GenerateRay at image points uv (n, 2): origins (n, 3), directions (n, 3)."""
    uv = np.ascontiguousarray(np.asarray(uv, dtype=np.float32).reshape(-1, 2))
    out = np.zeros((len(uv), 6), dtype=np.float32)
    r = lib().oracle_generate_ray_n(C.byref(cam), len(uv), _ptr(uv), pixel, int(corrected), _ptr(out))
    if r < 0:
        raise RuntimeError("generate_rays failed")
    return out[:, :3], out[:, 3:]


def guide_forward(params, inputs):
    """Guiding network (BF16 steps as the device): (n, 5) inputs -> (n, 129) bin probabilities + c."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 5)
    out = np.zeros((len(x), 129), dtype=np.float32)
    lib().oracle_guide_forward(_ptr(p), len(x), _ptr(x), _ptr(out))
    return out


def guide_gradient(params, records, rounded=True):
    """Gradient of the guiding objective over (n, 8) records and the loss."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 8)
    g = np.zeros(lib().oracle_guide_num_params(), dtype=np.float32)
    loss = lib().oracle_guide_gradient(_ptr(p), len(r), _ptr(r), int(rounded), _ptr(g))
    return g, float(loss)


def guide_loss64(params, records):
    p = np.ascontiguousarray(params, dtype=np.float64)
    r = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 8)
    return float(lib().oracle_guide_loss64(_ptr(p), len(r), _ptr(r)))


def guide_num_params():
    return int(lib().oracle_guide_num_params())


def philox(ctr, key):
    out = np.zeros(4, dtype=np.uint32)
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    lib().oracle_philox(_ptr(c), _ptr(k), _ptr(out))
    return out
