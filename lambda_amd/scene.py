"""Backend-neutral scene description and the procedural workloads.

A SceneDesc holds numpy arrays and node lists; `push(backend)` replays them in
order into any object exposing add_texture/add_medium/add_material/add_mesh/
add_light/commit (the GPU binding in lambda_amd.gpu, or the CPU checker used
by the tests). Light order is the call order, as Scene::AddObject/AddLight
(Scene.cpp:109-121) define it.
"""
import ctypes as C
import math

import numpy as np

from . import abi


def node(kind, inputs=(), p=()):
    """One shader-graph node: (type, inputs, params)."""
    return (int(kind), tuple(inputs), tuple(float(x) for x in p))


class Material:
    def __init__(self, nodes, bxdf=abi.NONE, emission=abi.NONE, intensity=1.0,
                 medium_interior=abi.NONE, medium_exterior=abi.NONE):
        self.nodes = list(nodes)
        self.bxdf = bxdf
        self.emission = emission
        self.intensity = float(intensity)
        self.medium_interior = medium_interior
        self.medium_exterior = medium_exterior

    def to_c(self):
        arr = (abi.Node * max(1, len(self.nodes)))()
        for i, (t, ins, p) in enumerate(self.nodes):
            arr[i].type = t
            for k in range(4):
                arr[i].in_[k] = ins[k] if k < len(ins) else abi.NONE
            for k in range(12):
                arr[i].p[k] = p[k] if k < len(p) else 0.0
        d = abi.MaterialDesc(len(self.nodes), arr, self.bxdf, self.emission, self.intensity,
                             self.medium_interior, self.medium_exterior)
        return d, arr


def lambertian(rgb):
    return Material([node(abi.NODE_RGB, p=rgb), node(abi.NODE_LAMBERTIAN, [0])], bxdf=1)


def emitter_blackbody(temperature, intensity, rgb_albedo=None):
    """MeshLight material: blackbody emission (GraphInputs.cpp:82-116); optional
    Lambertian BxDF (rgb_albedo) so the emitter also reflects."""
    nodes = [node(abi.NODE_SCALAR, p=[temperature]), node(abi.NODE_BLACKBODY, [0])]
    bxdf = abi.NONE
    if rgb_albedo is not None:
        nodes += [node(abi.NODE_RGB, p=rgb_albedo), node(abi.NODE_LAMBERTIAN, [2])]
        bxdf = 3
    return Material(nodes, bxdf=bxdf, emission=1, intensity=intensity)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class SceneDesc:
    def __init__(self):
        self.ops = []  # ("texture"|"medium"|"material"|"mesh"|"light", payload)
        self.n_tris = 0

    def add_texture(self, data):
        data = _f32(data)
        ch = 1 if data.ndim == 2 else data.shape[2]
        self.ops.append(("texture", data.reshape(data.shape[0], data.shape[1], ch)))
        return sum(1 for o in self.ops if o[0] == "texture") - 1

    def add_medium(self, **kw):
        self.ops.append(("medium", kw))
        return sum(1 for o in self.ops if o[0] == "medium") - 1

    def add_material(self, mat):
        self.ops.append(("material", mat))
        return sum(1 for o in self.ops if o[0] == "material") - 1

    def add_mesh(self, positions, indices, material, normals=None, tangents=None, uvs=None):
        positions = _f32(positions).reshape(-1, 3)
        indices = _u32(indices).reshape(-1, 3)
        self.ops.append(("mesh", dict(positions=positions, indices=indices, material=int(material),
                                      normals=None if normals is None else _f32(normals).reshape(-1, 3),
                                      tangents=None if tangents is None else _f32(tangents).reshape(-1, 4),
                                      uvs=None if uvs is None else _f32(uvs).reshape(-1, 2))))
        self.n_tris += len(indices)
        return sum(1 for o in self.ops if o[0] == "mesh") - 1

    def add_point_light(self, position, intensity, rgb=(1, 1, 1), temperature=0.0):
        self.ops.append(("light", dict(type=abi.LIGHT_POINT, position=position, intensity=intensity, rgb=rgb,
                                       temperature=temperature)))

    def add_spot_light(self, position, axis, cone_angle, falloff_start, intensity, rgb=(1, 1, 1), temperature=0.0):
        self.ops.append(("light", dict(type=abi.LIGHT_SPOT, position=position, axis=axis, cone_angle=cone_angle,
                                       falloff_start=falloff_start, intensity=intensity, rgb=rgb,
                                       temperature=temperature)))

    def add_env_light(self, rgb_image, intensity=1.0, offset=(0.0, 0.0)):
        img = _f32(rgb_image)
        self.ops.append(("light", dict(type=abi.LIGHT_ENVIRONMENT, env=img, intensity=intensity, offset=offset)))

    def push(self, backend, strategy=abi.LIGHTS_TREE):
        keep = []
        for kind, pl in self.ops:
            if kind == "texture":
                h, w, ch = pl.shape
                d = abi.TextureDesc(w, h, ch, pl.ctypes.data_as(abi.FP))
                keep.append(pl)
                backend.add_texture(d)
            elif kind == "medium":
                d = abi.MediumDesc()
                d.type = pl["type"]
                for k in ("sigma_a", "sigma_s", "bounds_min", "bounds_max"):
                    if k in pl:
                        getattr(d, k)[:] = [float(x) for x in pl[k]]
                d.g = float(pl.get("g", 0.0))
                if "density" in pl:
                    dens = _f32(pl["density"])
                    keep.append(dens)
                    d.nz, d.ny, d.nx = dens.shape
                    d.density = dens.ctypes.data_as(abi.FP)
                backend.add_medium(d)
            elif kind == "material":
                d, arr = pl.to_c()
                keep.append(arr)
                backend.add_material(d)
            elif kind == "mesh":
                d = abi.MeshDesc()
                d.num_vertices = len(pl["positions"])
                d.num_triangles = len(pl["indices"])
                d.positions = pl["positions"].ctypes.data_as(abi.FP)
                d.indices = pl["indices"].ctypes.data_as(abi.UP)
                for k in ("normals", "tangents", "uvs"):
                    if pl[k] is not None:
                        setattr(d, k, pl[k].ctypes.data_as(abi.FP))
                d.material = pl["material"]
                backend.add_mesh(d)
            else:
                d = abi.LightDesc()
                d.type = pl["type"]
                d.intensity = float(pl["intensity"])
                if "position" in pl:
                    d.position[:] = [float(x) for x in pl["position"]]
                if "axis" in pl:
                    d.axis[:] = [float(x) for x in pl["axis"]]
                d.cone_angle = float(pl.get("cone_angle", 0.0))
                d.falloff_start = float(pl.get("falloff_start", 0.0))
                d.rgb[:] = [float(x) for x in pl.get("rgb", (1, 1, 1))]
                d.temperature = float(pl.get("temperature", 0.0))
                if "env" in pl:
                    img = pl["env"]
                    keep.append(img)
                    d.env_height, d.env_width = img.shape[0], img.shape[1]
                    d.env_rgb = img.ctypes.data_as(abi.FP)
                    d.env_offset[:] = [float(x) for x in pl["offset"]]
                backend.add_light(d)
        backend.commit(strategy)
        return keep


# --------------------------------------------------------------------- cameras
def camera_pinhole(origin, width, height, phi, theta, fov):
    """PinholeCamera with Camera::SetRotation/SetFov (Camera.cpp:7-35)."""
    return _camera(abi.CAMERA_PINHOLE, origin, width, height, phi, theta, fov)


def camera_thin_lens(origin, width, height, phi, theta, fov, focal_length, aperture_size):
    cam = _camera(abi.CAMERA_THIN_LENS, origin, width, height, phi, theta, fov)
    cam.focal_length = focal_length
    cam.aperture = abi.APERTURE_CIRCULAR
    cam.aperture_size = aperture_size
    return cam


def _camera(kind, origin, width, height, phi, theta, fov):
    f = np.float32
    cam = abi.Camera()
    cam.type = kind
    cam.origin[:] = [float(x) for x in origin]
    cam.aspect = float(f(height) / f(width))
    cam.tan_fov = float(f(math.tan(float(f(0.5) * f(fov)))))
    ct, st = f(math.cos(f(theta))), f(math.sin(f(theta)))
    cp, sp = f(math.cos(f(phi))), f(math.sin(f(phi)))
    z = np.array([ct * sp, st, ct * cp], dtype=np.float32)
    mag = f(np.sqrt(f(z[0] * z[0] + z[1] * z[1]) + z[2] * z[2]))
    inv = f(f(1.0) / mag)
    z = (z * inv).astype(np.float32)
    up = np.array([0, 1, 0], dtype=np.float32)
    x = _cross(z, up)
    y = _cross(x, z)
    cam.z_hat[:] = [float(v) for v in z]
    cam.x_hat[:] = [float(v) for v in x]
    cam.y_hat[:] = [float(v) for v in y]
    return cam


def _cross(a, b):
    f = np.float32
    return np.array([f(a[1] * b[2]) - f(a[2] * b[1]), f(a[2] * b[0]) - f(a[0] * b[2]),
                     f(a[0] * b[1]) - f(a[1] * b[0])], dtype=np.float32)


# --------------------------------------------------------------------- workloads
def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2),(p0,p2,p3); Ng = (p1-p0)x(p2-p0)."""
    return np.array([p0, p1, p2, p3], dtype=np.float32), np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)


def _box(center, size, rot_y_deg):
    sx, sy, sz = [s * 0.5 for s in size]
    c = np.cos(np.radians(rot_y_deg))
    s = np.sin(np.radians(rot_y_deg))
    corners = np.array([[x, y, z] for x in (-sx, sx) for y in (-sy, sy) for z in (-sz, sz)], dtype=np.float64)
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    v = corners @ rot.T + np.asarray(center)
    # corner index = 4*ix + 2*iy + iz; outward-facing CCW faces
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in faces:
        tris += [[a, b, cc], [a, cc, d]]
    return v.astype(np.float32), np.array(tris, dtype=np.uint32)


def cornell_box(temperature=6500.0, intensity=10.0):
    """The small CPU-runnable workload: room [-1,1]^3 open towards +z (5 quads,
    10 triangles), a tall and a short box rotated +15 and -18 degrees about y
    (24 triangles), one 0.5 x 0.5 emissive quad at y = 0.999 facing -y with
    6500 K blackbody emission (2 triangles). 36 triangles, all Lambertian."""
    sd = SceneDesc()
    white = sd.add_material(lambertian((0.73, 0.73, 0.73)))
    red = sd.add_material(lambertian((0.65, 0.05, 0.05)))
    green = sd.add_material(lambertian((0.12, 0.45, 0.15)))
    light = sd.add_material(emitter_blackbody(temperature, intensity))
    # floor (normal +y), ceiling (-y), back wall (+z), left wall red (+x), right wall green (-x)
    v, i = _quad((-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1))
    sd.add_mesh(v, i, white)
    v, i = _quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1))
    sd.add_mesh(v, i, white)
    v, i = _quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1))
    sd.add_mesh(v, i, white)
    v, i = _quad((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1))
    sd.add_mesh(v, i, red)
    v, i = _quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1))
    sd.add_mesh(v, i, green)
    v, i = _box((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0)
    sd.add_mesh(v, i, white)
    v, i = _box((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0)
    sd.add_mesh(v, i, white)
    y = 0.999
    v, i = _quad((-0.25, y, -0.25), (0.25, y, -0.25), (0.25, y, 0.25), (-0.25, y, 0.25))
    sd.add_mesh(v, i, light)
    return sd


def cornell_camera(width=256, height=256):
    """Pinhole at (0,0,3.8), SetRotation(pi, 0) (looking down -z), SetFov(0.6911)."""
    return camera_pinhole((0.0, 0.0, 3.8), width, height, math.pi, 0.0, 0.6911)
