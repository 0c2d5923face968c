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
        return sum(1 for o in self.ops if o[0] in ("mesh", "instance")) - 1

    def add_proxy(self, positions, indices, material, normals=None, tangents=None, uvs=None):
        """InstanceProxy: a mesh kept for instancing (not rendered by itself)."""
        self.add_mesh(positions, indices, material, normals, tangents, uvs)
        op = self.ops[-1]
        self.n_tris -= len(op[1]["indices"])
        self.ops[-1] = ("proxy", op[1])
        return sum(1 for o in self.ops if o[0] == "proxy") - 1

    def add_instance(self, proxy, xfm, material=abi.NONE):
        """Instance of a proxy under a 3x4 column-major affine transform (columns x, y, z, translation)."""
        xfm = [float(v) for v in np.asarray(xfm, dtype=np.float32).reshape(12)]
        ops = [o for o in self.ops if o[0] == "proxy"]
        self.ops.append(("instance", dict(proxy=int(proxy), xfm=xfm, material=int(material))))
        self.n_tris += len(ops[proxy][1]["indices"])
        return sum(1 for o in self.ops if o[0] in ("mesh", "instance")) - 1

    def add_obj(self, path, materials=None, transform=None):
        """Meshes of an OBJ file (library loader). materials: {OBJ material name: material id};
        others follow MaterialImport::BuildMaterial (MaterialImport.cpp:233-281): a Kd colour gives
        Oren-Nayar (albedo Kd, sigma socket 2), anything else the magenta Lambertian of a missing
        material. transform: optional 3x4 column-major affine (the meshes become instances).
        Returns the object ids."""
        from . import gpu
        meshes, mats = gpu.load_obj(path)
        ids = {}
        for i, m in enumerate(mats):
            if materials and m["name"] in materials:
                ids[i] = materials[m["name"]]
            elif m["kd"] is not None:
                ids[i] = self.add_material(Material([node(abi.NODE_RGB, p=m["kd"]), node(abi.NODE_SCALAR, p=[2.0]),
                                                     node(abi.NODE_OREN_NAYAR_BRDF, [0, 1])], bxdf=2))
            else:
                ids[i] = self.add_material(lambertian((1.0, 0.0, 1.0)))
        out = []
        for m in meshes:
            args = (m["positions"], m["indices"], ids[m["material"]])
            kw = dict(normals=m["normals"], tangents=m["tangents"], uvs=m["uvs"])
            if transform is None:
                out.append(self.add_mesh(*args, **kw))
            else:
                out.append(self.add_instance(self.add_proxy(*args, **kw), transform))
        return out

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
            elif kind == "instance":
                backend.add_instance(pl["proxy"], pl["xfm"], pl["material"])
            elif kind in ("mesh", "proxy"):
                d = abi.MeshDesc()
                d.num_vertices = len(pl["positions"])
                d.num_triangles = len(pl["indices"])
                d.positions = pl["positions"].ctypes.data_as(abi.FP)
                d.indices = pl["indices"].ctypes.data_as(abi.UP)
                for k in ("normals", "tangents", "uvs"):
                    if pl[k] is not None:
                        setattr(d, k, pl[k].ctypes.data_as(abi.FP))
                d.material = pl["material"]
                (backend.add_mesh if kind == "mesh" else backend.add_proxy)(d)
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


def cornell_indirect(temperature=6500.0, intensity=30.0):
    """The Cornell room lit only indirectly: the emitter is a 0.4 x 0.4 quad at y = 0.3 above the
    tall box facing up (+y, one-sided in the corrected estimator), so most of the room sees light
    only after a bounce off the ceiling and walls (a case for guided BSDF sampling)."""
    sd = cornell_box(temperature, intensity)
    sd.ops = [op for op in sd.ops if not (op[0] == "mesh" and op[1]["material"] == 3)]
    sd.n_tris -= 2
    y = 0.3
    v, i = _quad((-0.55, y, -0.5), (-0.55, y, -0.1), (-0.15, y, -0.1), (-0.15, y, -0.5))
    sd.add_mesh(v, i, 3)
    return sd


def passthrough(interior=abi.NONE, exterior=abi.NONE):
    """Material with neither BxDF nor emission: a media boundary that rays cross
    (MediaBoundary, Media.h:31-57; skipped by Scene::IntersectTr, Scene.cpp:45-59)."""
    return Material([], medium_interior=interior, medium_exterior=exterior)


def aabb(bmin, bmax):
    """Axis-aligned box, 12 triangles, Ng = (v1-v0)x(v2-v0) facing outwards, so
    that MediaBoundary::GetMedium gives the interior for entering rays."""
    c = [(a + b) * 0.5 for a, b in zip(bmin, bmax)]
    size = [b - a for a, b in zip(bmin, bmax)]
    return _box(c, size, 0.0)


def homogeneous(sd, sigma_a, sigma_s, g=0.0):
    return sd.add_medium(type=abi.MEDIUM_HOMOGENEOUS, sigma_a=sigma_a, sigma_s=sigma_s, g=g)


def grid_medium(sd, density, sigma_a, sigma_s, g, bmin, bmax):
    """Piecewise-constant density grid (z-major, shape (nz, ny, nx)) over [bmin, bmax];
    sigma = density * sigma_{a,s}."""
    return sd.add_medium(type=abi.MEDIUM_GRID, sigma_a=sigma_a, sigma_s=sigma_s, g=g, density=density,
                         bounds_min=bmin, bounds_max=bmax)


def fbm_density(res=128, octaves=5, freq=2.0, seed=7):
    """d = max(0, fBm) of 3-D Perlin noise sampled at voxel centres of [-1,1]^3
    (lacunarity 2, gain 0.5); z-major float32 (res, res, res)."""
    c = (np.arange(res, dtype=np.float64) + 0.5) / res * 2.0 - 1.0
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    acc = np.zeros(len(p))
    amp, f = 1.0, freq
    for o in range(octaves):
        acc += amp * perlin3(p * f + 17.31 * o, seed=seed)
        amp *= 0.5
        f *= 2.0
    return np.maximum(acc, 0.0).reshape(res, res, res).astype(np.float32)


def fog_cornell(kind="homogeneous", res=16, point_light=True, sigma_scale=1.0):
    """Cornell room (walls, ceiling quad light) with a pass-through fog box
    [-0.95,0.95] x [-0.95,0.6] x [-0.95,0.95] holding a chromatic medium
    (homogeneous, or a res^3 random-density grid), an inner Lambertian box whose
    exterior medium is the fog, and optionally a point light inside the fog."""
    sd = cornell_box()
    bmin, bmax = (-0.95, -0.95, -0.95), (0.95, 0.6, 0.95)
    sa = [0.05 * sigma_scale, 0.1 * sigma_scale, 0.2 * sigma_scale]
    ss = [0.6 * sigma_scale, 0.45 * sigma_scale, 0.3 * sigma_scale]
    if kind == "homogeneous":
        med = homogeneous(sd, sa, ss, g=0.3)
    else:
        dens = np.random.default_rng(3).uniform(0.0, 2.0, size=(res, res, res)).astype(np.float32)
        dens[dens < 0.4] = 0.0
        med = grid_medium(sd, dens, sa, ss, 0.3, bmin, bmax)
    fog = sd.add_material(passthrough(interior=med))
    v, i = aabb(bmin, bmax)
    sd.add_mesh(v, i, fog)
    inner = sd.add_material(Material([node(abi.NODE_RGB, p=(0.6, 0.6, 0.5)), node(abi.NODE_LAMBERTIAN, [0])],
                                     bxdf=1, medium_exterior=med))
    v, i = _box((-0.3, -0.6, 0.2), (0.4, 0.4, 0.4), 20.0)
    sd.add_mesh(v, i, inner)
    if point_light:
        sd.add_point_light((0.35, 0.1, 0.3), 0.6, rgb=(1.0, 0.8, 0.6))
    return sd


def media_scene(res=128, sigma_t=4.0, albedo=0.9, g=0.6, n_points=64, octaves=5):
    """Heterogeneous-media workload: a pass-through box [-1,1]^3 whose interior is a
    res^3 FP32 density grid d = max(0, 5-octave Perlin fBm) with sigma_t = 4 d,
    sigma_s = 0.9 sigma_t, HG g = 0.6; one spotlight at (0,3,0) aiming down
    (cone 30 deg, falloff from 20 deg) and 64 point lights on a radius-3 sphere
    (Fibonacci layout)."""
    sd = SceneDesc()
    dens = fbm_density(res, octaves)
    ss = sigma_t * albedo
    sa = sigma_t - ss
    med = grid_medium(sd, dens, [sa] * 3, [ss] * 3, g, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    box = sd.add_material(passthrough(interior=med))
    v, i = aabb((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    sd.add_mesh(v, i, box)
    sd.add_spot_light((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), math.radians(30.0), math.radians(20.0), 30.0,
                      temperature=5000.0)
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for k in range(n_points):
        y = 1.0 - 2.0 * (k + 0.5) / n_points
        r = math.sqrt(max(0.0, 1.0 - y * y))
        phi = golden * k
        pos = (3.0 * r * math.cos(phi), 3.0 * y, 3.0 * r * math.sin(phi))
        sd.add_point_light(pos, 1.5, temperature=3000.0 + 6000.0 * (k % 8) / 7.0)
    return sd


def media_camera(width=1024, height=1024):
    """Pinhole at (0, 0.4, 3.8) looking down -z, SetFov(0.6911)."""
    return camera_pinhole((0.0, 0.4, 3.8), width, height, math.pi, 0.0, 0.6911)


def _quad_uv(p0, p1, p2, p3):
    v, i = _quad(p0, p1, p2, p3)
    return v, i, np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32)


def procedural_textures(seed=7):
    """Texture set of the procedural-material scene: 0 RGBA (alpha ramp), 1 RGB spectra,
    2 a 41-sample SPD (380-780 nm)."""
    rng = np.random.default_rng(seed)
    rgba = rng.uniform(0.1, 0.9, size=(16, 16, 4)).astype(np.float32)
    rgba[..., 3] = np.linspace(0.05, 0.95, 16)[None, :]
    rgb = rng.uniform(0.05, 0.8, size=(8, 8, 3)).astype(np.float32)
    lam = np.linspace(380, 780, 41)
    spd = (0.6 + 0.4 * np.sin(lam / 37.0)).astype(np.float32).reshape(1, 41)
    return rgba, rgb, spd


def procedural_materials(sd):
    """Materials built from every procedural, converter and input node (GraphTexture.cpp,
    GraphConverters.cpp, GraphMaths.cpp, GraphInputs.cpp). Returns the material ids
    (floor, back, left, right, tall, short, emitter_spd, emitter_voronoi)."""
    rgba, rgb, spd = procedural_textures()
    t_rgba, t_rgb, t_spd = sd.add_texture(rgba), sd.add_texture(rgb), sd.add_texture(spd)
    N = abi
    floor = sd.add_material(Material([
        node(N.NODE_VORONOI, p=[6]), node(N.NODE_OCTAVE_NOISE, p=[2, 3, N.NODE_VORONOI, 3]),
        node(N.NODE_PERLIN, p=[5]), node(N.NODE_VALUE_NOISE, p=[7]),
        node(N.NODE_MERGE_RGB, [2, 3, 1]), node(N.NODE_LAMBERTIAN, [4])], bxdf=5))
    back = sd.add_material(Material([
        node(N.NODE_SCALAR, p=[2.333]), node(N.NODE_FRESNEL, [0]), node(N.NODE_RGB, p=[.7, .7, .7]),
        node(N.NODE_LAMBERTIAN, [2]), node(N.NODE_SPECULAR_BRDF, [2]), node(N.NODE_MIX_BSDF, [3, 4, 1])], bxdf=5))
    left = sd.add_material(Material([
        node(N.NODE_NORMAL), node(N.NODE_VEC3, p=[.25, .25, .25]), node(N.NODE_VECTOR_MATHS, [0, 1], p=[N.OP_MULTIPLY]),
        node(N.NODE_VEC3, p=[.5, .5, .5]), node(N.NODE_VECTOR_MATHS, [2, 3], p=[N.OP_ADD]),
        node(N.NODE_VEC3, p=[0, 0, 1]), node(N.NODE_CROSS, [0, 5]), node(N.NODE_LENGTH, [6]),
        node(N.NODE_SEPARATE_XYZ, [4], p=[0]), node(N.NODE_SCALAR, p=[.5]),
        node(N.NODE_SCALAR_MATHS, [7, 9], p=[N.OP_MULTIPLY]), node(N.NODE_MERGE_RGB, [8, 10, 9, 9]),
        node(N.NODE_LAMBERTIAN, [11])], bxdf=12))
    right = sd.add_material(Material([
        node(N.NODE_IMAGE_CHANNEL, p=[t_rgba, 3]), node(N.NODE_SCALAR_TO_COLOUR, [0]),
        node(N.NODE_IMAGE_COLOUR, p=[t_rgba]), node(N.NODE_SEPARATE_R, [2], p=[3]), node(N.NODE_SCALAR, p=[.35]),
        node(N.NODE_OREN_NAYAR_BRDF, [1, 4]), node(N.NODE_LAMBERTIAN, [2]), node(N.NODE_MIX_BSDF, [5, 6, 3])], bxdf=7))
    tall = sd.add_material(Material([node(N.NODE_SPECTRAL_TEXTURE, p=[t_rgb]), node(N.NODE_LAMBERTIAN, [0])], bxdf=1))
    short = sd.add_material(Material([
        node(N.NODE_CHECKER, p=[8]), node(N.NODE_SCALAR_TO_COLOUR, [0]), node(N.NODE_LAMBERTIAN, [1])], bxdf=2))
    e_spd = sd.add_material(Material([node(N.NODE_SPECTRAL, p=[t_spd, 380, 780])], emission=0, intensity=8.0))
    e_vor = sd.add_material(Material([node(N.NODE_VORONOI, p=[4]), node(N.NODE_CHECKER, p=[3]),
                                      node(N.NODE_SCALAR_MATHS, [0, 1], p=[N.OP_ADD])], emission=2, intensity=4.0))
    return floor, back, left, right, tall, short, e_spd, e_vor


def procedural_cornell():
    """Cornell room whose surfaces (with uvs) use the procedural, converter and input nodes: a
    ceiling quad light with a sampled-SPD emission and a small back-wall emitter with a
    Voronoi + checker emission."""
    sd = SceneDesc()
    floor, back, left, right, tall, short, e_spd, e_vor = procedural_materials(sd)
    for quad, mat in [(((-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1)), floor),
                      (((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)), floor),
                      (((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)), back),
                      (((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1)), left),
                      (((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1)), right)]:
        v, i, uv = _quad_uv(*quad)
        sd.add_mesh(v, i, mat, uvs=uv)
    for (c, size, rot), mat in [(((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0), tall),
                                (((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0), short)]:
        v, i = _box(c, size, rot)
        uv = np.stack([v[:, 0] + v[:, 2], v[:, 1]], axis=1).astype(np.float32)
        sd.add_mesh(v, i, mat, uvs=uv)
    v, i, uv = _quad_uv((-0.25, 0.999, -0.25), (0.25, 0.999, -0.25), (0.25, 0.999, 0.25), (-0.25, 0.999, 0.25))
    sd.add_mesh(v, i, e_spd, uvs=uv)
    v, i, uv = _quad_uv((-0.6, 0.2, -0.999), (-0.2, 0.2, -0.999), (-0.2, 0.6, -0.999), (-0.6, 0.6, -0.999))
    sd.add_mesh(v, i, e_vor, uvs=uv)
    return sd


def lobe_zoo(level=2):
    """Cornell room holding one object per reference lobe that the other workloads leave out:
    FresnelBSDF and SpecularBTDF spheres, a SpecularBRDF box, a GhostBTDF card, an Oren-Nayar BTDF
    sheet (mixed with Lambertian), Beckmann and Trowbridge-Reitz microfacet BRDFs with the
    dielectric Fresnel (one through RoughnessToAlpha), a Beckmann conductor and a rough dielectric
    (Specular.cpp:18-95, Ghost.cpp:13-28, OrenNayar.cpp:52-59, Microfacet.cpp:12-69,
    MicrofacetDistribution.cpp:11-71)."""
    sd = cornell_box()
    N = abi
    white = (1.0, 1.0, 1.0)

    def mat(nodes, root):
        return sd.add_material(Material(nodes, bxdf=root))

    fresnel = mat([node(N.NODE_RGB, p=white), node(N.NODE_SCALAR, p=[1.5]), node(N.NODE_FRESNEL_BSDF, [0, 1])], 2)
    spec_t = mat([node(N.NODE_RGB, p=(0.9, 0.95, 1.0)), node(N.NODE_SPECULAR_BTDF, [0], p=[1.33])], 1)
    mirror = mat([node(N.NODE_RGB, p=(0.8, 0.75, 0.7)), node(N.NODE_SPECULAR_BRDF, [0])], 1)
    ghost = mat([node(N.NODE_SCALAR, p=[0.6]), node(N.NODE_GHOST_BTDF, [0])], 1)
    on_t = mat([node(N.NODE_RGB, p=(0.7, 0.5, 0.3)), node(N.NODE_SCALAR, p=[0.4]),
                node(N.NODE_OREN_NAYAR_BTDF, [0, 1]), node(N.NODE_LAMBERTIAN, [0]),
                node(N.NODE_SCALAR, p=[0.35]), node(N.NODE_MIX_BSDF, [2, 3, 4])], 5)

    def micro(ggx, alpha, conductor, direct_alpha=True):
        p = [1.0 if ggx else 0.0, 1.0 if conductor else 0.0, 1.45, *COPPER_ETA, *COPPER_K, 1.0 if direct_alpha else 0.0]
        return mat([node(N.NODE_RGB, p=(0.9, 0.8, 0.7)), node(N.NODE_SCALAR, p=[alpha]),
                     node(N.NODE_MICROFACET_BRDF, [0, 1], p=p)], 2)

    beck_d = micro(False, 0.25, False)
    ggx_d = micro(True, 0.4, False, direct_alpha=False)  # roughness 0.4 through RoughnessToAlpha
    beck_c = micro(False, 0.15, True)
    rough = sd.add_material(rough_dielectric(0.2, 1.5))
    v, f = icosphere(level)
    for c, r, m in [((-0.55, -0.72, 0.35), 0.27, fresnel), ((0.0, -0.75, 0.55), 0.24, spec_t),
                    ((0.55, -0.72, 0.3), 0.27, beck_d), ((-0.5, 0.05, -0.45), 0.25, ggx_d),
                    ((0.5, 0.0, -0.5), 0.28, beck_c), ((0.05, -0.35, -0.25), 0.3, rough)]:
        p = v * r + np.array(c)
        n = _vertex_normals(p, f)
        sd.add_mesh(p, f, m, normals=n, tangents=_tangents(n))
    bv, bi = _box((0.1, 0.55, -0.7), (0.5, 0.3, 0.2), 25.0)
    sd.add_mesh(bv, bi, mirror)
    gv, gi = _quad((-0.35, -0.95, 0.75), (0.35, -0.95, 0.75), (0.35, -0.3, 0.75), (-0.35, -0.3, 0.75))
    sd.add_mesh(gv, gi, ghost)
    tv, ti = _quad((0.7, -0.2, -0.9), (0.7, -0.2, 0.1), (0.7, 0.5, 0.1), (0.7, 0.5, -0.9))
    sd.add_mesh(tv, ti, on_t)
    return sd


def _planar_uv(v):
    """uv of box vertices: (x, y) and z folded into [0, 1] so every face varies over the texture."""
    v = np.asarray(v, dtype=np.float32)
    return np.stack([(v[:, 0] + 1.0) * 0.4 + (v[:, 2] + 1.0) * 0.1, (v[:, 1] + 1.0) * 0.45], axis=1).astype(np.float32)


def inverse_room(tex_res=16, medium=True, point_light=False, light_intensity=10.0, seed=3):
    """Differentiable-rendering workload (config-4 style): the Cornell room whose floor, back wall
    and tall box read their albedo from one RGB texture and their roughness from two scalar
    textures; optionally a grid medium (6^3 random densities) in a pass-through box
    [-0.8,0.8] x [0.3,0.75] x [-0.8,0.8] under the ceiling light, or a point light instead of the
    emissive quad.
      texture 0: albedo, tex_res^2 RGB in [0.2, 0.8]
      texture 1: roughness, tex_res^2 in [0.25, 0.7] (GGX conductor, through RoughnessToAlpha)
      texture 2: alpha, tex_res^2 in [0.1, 0.4] (Beckmann conductor, alpha given directly)
      floor:     Lambertian(albedo)
      back wall: Mix(Lambertian(albedo), Microfacet GGX conductor(roughness), 0.5)
      tall box:  Microfacet Beckmann conductor(R = albedo, alpha)
    Returns the SceneDesc and a dict of ids (textures, medium, light material)."""
    rng = np.random.default_rng(seed)
    R = tex_res
    sd = SceneDesc()
    N = abi
    t_alb = sd.add_texture(rng.uniform(0.2, 0.8, size=(R, R, 3)).astype(np.float32))
    t_rough = sd.add_texture(rng.uniform(0.25, 0.7, size=(R, R)).astype(np.float32))
    t_alpha = sd.add_texture(rng.uniform(0.1, 0.4, size=(R, R)).astype(np.float32))
    med = abi.NONE
    if medium:
        dens = rng.uniform(0.0, 1.5, size=(6, 6, 6)).astype(np.float32)
        med = grid_medium(sd, dens, (0.15, 0.15, 0.15), (0.9, 0.75, 0.6), 0.3, (-0.8, 0.3, -0.8), (0.8, 0.75, 0.8))
    white = sd.add_material(lambertian((0.73, 0.73, 0.73)))
    red = sd.add_material(lambertian((0.65, 0.05, 0.05)))
    green = sd.add_material(lambertian((0.12, 0.45, 0.15)))
    floor = sd.add_material(Material([node(N.NODE_IMAGE_COLOUR, p=[t_alb]), node(N.NODE_LAMBERTIAN, [0])], bxdf=1))
    ggx = [1.0, 1.0, 1.45, *COPPER_ETA, *COPPER_K, 0.0]
    back = sd.add_material(Material([
        node(N.NODE_IMAGE_COLOUR, p=[t_alb]), node(N.NODE_LAMBERTIAN, [0]), node(N.NODE_RGB, p=(0.9, 0.9, 0.9)),
        node(N.NODE_IMAGE_SCALAR, p=[t_rough]), node(N.NODE_MICROFACET_BRDF, [2, 3], p=ggx),
        node(N.NODE_SCALAR, p=[0.5]), node(N.NODE_MIX_BSDF, [1, 4, 5])], bxdf=6))
    beck = [0.0, 1.0, 1.45, *COPPER_ETA, *COPPER_K, 1.0]
    tall = sd.add_material(Material([
        node(N.NODE_IMAGE_COLOUR, p=[t_alb]), node(N.NODE_IMAGE_CHANNEL, p=[t_alpha, 0]),
        node(N.NODE_MICROFACET_BRDF, [0, 1], p=beck)], bxdf=2))
    light = abi.NONE
    if not point_light:
        light = sd.add_material(emitter_blackbody(6500.0, light_intensity))
    v, i, uv = _quad_uv((-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1))
    sd.add_mesh(v, i, floor, uvs=uv)
    v, i = _quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1))
    sd.add_mesh(v, i, white)
    v, i, uv = _quad_uv((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1))
    sd.add_mesh(v, i, back, uvs=uv)
    v, i = _quad((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1))
    sd.add_mesh(v, i, red)
    v, i = _quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1))
    sd.add_mesh(v, i, green)
    v, i = _box((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 15.0)
    sd.add_mesh(v, i, tall, uvs=_planar_uv(v))
    v, i = _box((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -18.0)
    sd.add_mesh(v, i, white)
    if point_light:
        sd.add_point_light((0.1, 0.85, 0.2), 1.5)
    else:
        y = 0.999
        v, i = _quad((-0.25, y, -0.25), (0.25, y, -0.25), (0.25, y, 0.25), (-0.25, y, 0.25))
        sd.add_mesh(v, i, light)
    if medium:
        fog = sd.add_material(passthrough(interior=med))
        v, i = aabb((-0.8, 0.3, -0.8), (0.8, 0.75, 0.8))
        sd.add_mesh(v, i, fog)
    return sd, dict(albedo=t_alb, roughness=t_rough, alpha=t_alpha, medium=med, light=light)


def slab_scene(tex_res=8, grid=(8, 4, 8), seed=5):
    """A textured floor lit by an emissive quad through a grid-medium slab
    [-0.8,0.8] x [0.1,0.5] x [-0.8,0.8], seen by slab_camera from below the slab: camera rays never
    enter the medium, every light path to the floor crosses it (transmittance only)."""
    rng = np.random.default_rng(seed)
    sd = SceneDesc()
    N = abi
    t_alb = sd.add_texture(rng.uniform(0.3, 0.9, size=(tex_res, tex_res, 3)).astype(np.float32))
    nx, ny, nz = grid
    dens = rng.uniform(0.2, 1.8, size=(nz, ny, nx)).astype(np.float32)
    med = grid_medium(sd, dens, (0.6, 0.8, 1.0), (0.4, 0.3, 0.2), 0.0, (-0.8, 0.1, -0.8), (0.8, 0.5, 0.8))
    floor = sd.add_material(Material([node(N.NODE_IMAGE_COLOUR, p=[t_alb]), node(N.NODE_LAMBERTIAN, [0])], bxdf=1))
    light = sd.add_material(emitter_blackbody(6500.0, 12.0))
    v, i, uv = _quad_uv((-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1))
    sd.add_mesh(v, i, floor, uvs=uv)
    y = 0.95
    v, i = _quad((-0.4, y, -0.4), (0.4, y, -0.4), (0.4, y, 0.4), (-0.4, y, 0.4))
    sd.add_mesh(v, i, light)
    fog = sd.add_material(passthrough(interior=med))
    v, i = aabb((-0.8, 0.1, -0.8), (0.8, 0.5, 0.8))
    sd.add_mesh(v, i, fog)
    return sd, dict(albedo=t_alb, medium=med, light=light)


def slab_camera(width=64, height=64):
    """Pinhole at (0, -0.2, 2.2) pitched 0.45 rad down, SetFov(0.6): every ray descends."""
    return camera_pinhole((0.0, -0.2, 2.2), width, height, math.pi, -0.45, 0.6)


def camera_spherical(origin, width, height, phi=math.pi, theta=0.0, offset_phi=0.0, offset_theta=0.0):
    """SphericalCamera (Camera.cpp:66-71): u -> phi = 2 pi u + offset, v -> theta = pi v + offset."""
    cam = _camera(abi.CAMERA_SPHERICAL, origin, width, height, phi, theta, 1.0)
    cam.offset_phi, cam.offset_theta = offset_phi, offset_theta
    return cam


def camera_blades(origin, width, height, phi, theta, fov, focal_length, aperture_size, blades):
    """Thin lens with a BladeAperture of `blades` blades (Aperture.cpp:10-19)."""
    cam = camera_thin_lens(origin, width, height, phi, theta, fov, focal_length, aperture_size)
    cam.aperture = abi.APERTURE_BLADE
    cam.blades = blades
    return cam


def cornell_camera(width=256, height=256):
    """Pinhole at (0,0,3.8), SetRotation(pi, 0) (looking down -z), SetFov(0.6911)."""
    return camera_pinhole((0.0, 0.0, 3.8), width, height, math.pi, 0.0, 0.6911)


# --------------------------------------------------------------------- procedural helpers
def _perm(seed):
    p = np.random.default_rng(seed).permutation(256)
    return np.concatenate([p, p]).astype(np.int64)


def _fade(t):
    return t * t * t * (t * (t * 6 - 15) + 10)


_GRAD = np.array([[1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
                  [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1], [1, 1, 0], [-1, 1, 0], [0, -1, 1], [0, -1, -1]],
                 dtype=np.float64)


def perlin3(p, seed=0):
    """Improved Perlin noise (Perlin 2002) at points p (N x 3), values in about [-1, 1]."""
    perm = _perm(seed)
    p = np.asarray(p, dtype=np.float64)
    pi = np.floor(p).astype(np.int64)
    pf = p - pi
    pi &= 255
    u, v, w = _fade(pf[:, 0]), _fade(pf[:, 1]), _fade(pf[:, 2])

    def g(ix, iy, iz, dx, dy, dz):
        h = perm[perm[perm[ix] + iy] + iz] & 15
        gr = _GRAD[h]
        return gr[:, 0] * dx + gr[:, 1] * dy + gr[:, 2] * dz

    x, y, z = pi[:, 0], pi[:, 1], pi[:, 2]
    fx, fy, fz = pf[:, 0], pf[:, 1], pf[:, 2]
    n000 = g(x, y, z, fx, fy, fz)
    n100 = g(x + 1, y, z, fx - 1, fy, fz)
    n010 = g(x, y + 1, z, fx, fy - 1, fz)
    n110 = g(x + 1, y + 1, z, fx - 1, fy - 1, fz)
    n001 = g(x, y, z + 1, fx, fy, fz - 1)
    n101 = g(x + 1, y, z + 1, fx - 1, fy, fz - 1)
    n011 = g(x, y + 1, z + 1, fx, fy - 1, fz - 1)
    n111 = g(x + 1, y + 1, z + 1, fx - 1, fy - 1, fz - 1)
    x00 = n000 + u * (n100 - n000)
    x10 = n010 + u * (n110 - n010)
    x01 = n001 + u * (n101 - n001)
    x11 = n011 + u * (n111 - n011)
    y0 = x00 + v * (x10 - x00)
    y1 = x01 + v * (x11 - x01)
    return y0 + w * (y1 - y0)


def icosphere(level):
    """Unit icosphere by repeated midpoint subdivision (20 * 4^level faces)."""
    t = (1 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = [np.array(x, dtype=np.float64) / np.linalg.norm(x) for x in v]
    faces = f
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.uint32)


def _vertex_normals(v, f):
    n = np.zeros_like(v)
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def _tangents(n):
    up = np.where(np.abs(n[:, 1:2]) < 0.99, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(up, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return np.concatenate([t, np.ones((len(n), 1))], axis=1)


def affine(rot=(0.0, 0.0, 0.0), scale=(1.0, 1.0, 1.0), translate=(0.0, 0.0, 0.0)):
    """3x4 column-major affine transform (affine3.h layout): R_x R_y R_z diag(scale), then translate."""
    cx, cy, cz = np.cos(rot)
    sx, sy, sz = np.sin(rot)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    m = rx @ ry @ rz @ np.diag(scale)
    return np.concatenate([m[:, 0], m[:, 1], m[:, 2], np.asarray(translate, dtype=np.float64)]).astype(np.float32)


def instanced_cornell(n=24, seed=2, level=2):
    """Cornell room holding `n` instances of one smooth icosphere proxy (uvs, tangents) under random
    rotations, non-uniform scales and translations (every fourth one mirrored), with the proxy's
    material, a Lambertian override or an emissive override, plus instances of a quad light proxy."""
    sd = cornell_box()
    rng = np.random.default_rng(seed)
    v, f = icosphere(level)
    nrm = _vertex_normals(v, f)
    uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi), 0.5 + v[:, 1] * 0.5], axis=1)
    glossy = sd.add_material(ggx_conductor(0.2))
    blue = sd.add_material(lambertian((0.2, 0.3, 0.7)))
    glow = sd.add_material(emitter_blackbody(3500.0, 3.0, rgb_albedo=(0.5, 0.5, 0.5)))
    sphere = sd.add_proxy(v, f, glossy, normals=nrm, tangents=_tangents(nrm), uvs=uv)
    for k in range(n):
        sc = rng.uniform(0.06, 0.16, size=3)
        if k % 4 == 3:
            sc[0] = -sc[0]
        x = affine(rng.uniform(0, 2 * np.pi, size=3), sc, rng.uniform([-0.8, -0.85, -0.8], [0.8, 0.6, 0.6]))
        sd.add_instance(sphere, x, [abi.NONE, blue, glow][k % 3])
    lq = sd.add_material(emitter_blackbody(6000.0, 6.0))
    qv, qi = _quad((-0.05, 0, -0.05), (0.05, 0, -0.05), (0.05, 0, 0.05), (-0.05, 0, 0.05))  # Ng = -y
    quad = sd.add_proxy(qv, qi, lq)
    for k in range(4):
        sd.add_instance(quad, affine((0.0, 0.5 * k, 0.0), (1, 1, 1), (-0.6 + 0.4 * k, 0.95, -0.5)))
    return sd


COPPER_ETA = (0.200438, 0.924033, 1.102212)  # RGB-fitted copper optical constants
COPPER_K = (3.912949, 2.452848, 2.142188)


def ggx_conductor(alpha, eta=COPPER_ETA, k=COPPER_K):
    """MicrofacetBRDF + TrowbridgeReitz + FresnelConductor with alpha given directly."""
    p = [1.0, 1.0, 1.0, eta[0], eta[1], eta[2], k[0], k[1], k[2], 1.0]
    return Material([node(abi.NODE_RGB, p=(1, 1, 1)), node(abi.NODE_SCALAR, p=[alpha]),
                     node(abi.NODE_MICROFACET_BRDF, [0, 1], p=p)], bxdf=2)


def rough_dielectric(alpha, ior=1.5):
    return Material([node(abi.NODE_RGB, p=(1, 1, 1)), node(abi.NODE_SCALAR, p=[alpha]),
                     node(abi.NODE_ROUGH_DIELECTRIC, [0, 1], p=[ior])], bxdf=2)


def oren_nayar(rgb, sigma):
    return Material([node(abi.NODE_RGB, p=rgb), node(abi.NODE_SCALAR, p=[sigma]),
                     node(abi.NODE_OREN_NAYAR_BRDF, [0, 1])], bxdf=2)


def sky(width=1024, height=512):
    """Procedural FP32 sky: vertical gradient plus a sun disk (equirectangular)."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * np.pi
    phi = u * 2 * np.pi
    th, ph = np.meshgrid(theta, phi, indexing="ij")
    up = np.cos(th)
    horizon = np.array([0.9, 0.85, 0.8])
    zenith = np.array([0.25, 0.45, 0.9])
    t = np.clip(up, 0, 1)[..., None]
    img = horizon * (1 - t) + zenith * t
    img = np.where(up[..., None] < 0, np.array([0.12, 0.1, 0.08]), img)
    d = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], axis=-1)
    sun = np.array([0.4, 0.75, -0.5])
    sun /= np.linalg.norm(sun)
    img = img + (d @ sun > np.cos(0.03))[..., None] * np.array([400.0, 380.0, 340.0])
    return img.astype(np.float32)


def many_light_scene(n_spheres=200, sphere_level=4, ground_res=297, n_light_quads=2048, env=True, seed=0):
    """The many-light glossy mesh workload: displaced icospheres on a 40 x 40 ground,
    emissive quads sampled through the light tree, an HDR sky.

    Level-4 icospheres (5,120 faces) make 200 spheres 1,024,000 triangles; with the
    297 x 297-quad ground (176,418 triangles) the scene has 1,200,418 triangles
    (level 6 would give 16.4M, inconsistent with the 1.2M figure; the count is
    what is honoured)."""
    rng = np.random.default_rng(seed)
    sd = SceneDesc()
    # ground: [-20, 20]^2 at y = 0, normal +y
    g = ground_res
    xs = np.linspace(-20, 20, g + 1)
    gx, gz = np.meshgrid(xs, xs, indexing="ij")
    gv = np.stack([gx.ravel(), np.zeros(gx.size), gz.ravel()], axis=1)
    idx = np.arange((g + 1) * (g + 1)).reshape(g + 1, g + 1)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    gf = np.concatenate([np.stack([a, d, c], 1), np.stack([a, c, b], 1)])  # (d-a)x(c-a) points +y
    ground = sd.add_material(oren_nayar((0.5, 0.5, 0.5), 0.3491))
    sd.add_mesh(gv, gf, ground)
    # spheres on the integer grid points of [-20, 19]^2
    base_v, base_f = icosphere(sphere_level)
    cells = rng.choice(1600, size=n_spheres, replace=False)
    for k, cell in enumerate(cells):
        cx, cz = -20 + (cell % 40) + 0.5, -20 + (cell // 40) + 0.5
        r = rng.uniform(0.3, 0.8)
        centre = np.array([cx, r, cz])
        p = base_v * r + centre
        disp = perlin3(p * 4.0, seed=seed + 1) * 0.05
        p = p + base_v * disp[:, None]
        n = _vertex_normals(p, base_f)
        kind = rng.integers(0, 3)
        if kind == 0:
            mat = ggx_conductor(float(rng.uniform(0.05, 0.5)))
        elif kind == 1:
            mat = rough_dielectric(float(rng.uniform(0.05, 0.5)))
        else:
            mat = oren_nayar(tuple(rng.uniform(0.2, 0.9, size=3)), 0.3491)
        mi = sd.add_material(mat)
        sd.add_mesh(p, base_f, mi, normals=n, tangents=_tangents(n))
    # emissive quads: side 0.05 at heights U[1, 3], facing down
    for q in range(n_light_quads):
        T = float(rng.uniform(2000, 9000))
        inten = float(10 ** rng.uniform(0, 3))
        c = np.array([rng.uniform(-20, 20), rng.uniform(1, 3), rng.uniform(-20, 20)])
        s = 0.025
        v = np.array([[c[0] - s, c[1], c[2] - s], [c[0] + s, c[1], c[2] - s], [c[0] + s, c[1], c[2] + s],
                      [c[0] - s, c[1], c[2] + s]])
        m = sd.add_material(emitter_blackbody(T, inten))
        sd.add_mesh(v, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32), m)  # (v1-v0)x(v2-v0) = -y
    if env:
        sd.add_env_light(sky(), intensity=1.0)
    return sd


def _smooth_field(rng, res, coarse, lo, hi, channels=1):
    """A res x res texture: a coarse random grid upsampled bilinearly, in [lo, hi]."""
    g = rng.uniform(lo, hi, size=(coarse + 1, coarse + 1, channels))
    t = (np.arange(res) + 0.5) / res * coarse
    i = np.minimum(t.astype(int), coarse - 1)
    f = (t - i)[:, None]
    rows = g[i] * (1 - f)[:, :, None] + g[i + 1] * f[:, :, None]          # (res, coarse + 1, ch)
    out = rows[:, i] * (1 - f.T)[:, :, None] + rows[:, i + 1] * f.T[:, :, None]
    out = out.astype(np.float32)
    return out[..., 0] if channels == 1 else out


def inverse_scene(tex_res=512, grid_res=64, n_spheres=9, ground_res=48, n_light_quads=64, seed=0, albedo=None,
                  roughness=None, density=None):
    """The differentiable inverse-rendering workload: a 12 x 12 patch of the many-light field
    (ground_res^2-quad ground and n_spheres level-4 displaced icospheres: 50,688 triangles),
    tex_res^2 FP32 albedo (RGB) and roughness textures on every surface (ground: Lambertian with
    the albedo; spheres: Mix(Lambertian(albedo), GGX conductor(roughness through
    RoughnessToAlpha)) with spherical uvs), a grid_res^3 fBm density grid (sigma_t 2 d, albedo
    0.8, HG g 0.3) in a pass-through box [-4,4] x [1.2,2.7] x [-4,4] (a cloud layer between the
    spheres and the lights: no surface lies inside it), n_light_quads emissive quads at heights
    2.8 to 4 through the light tree, and the HDR sky.
    albedo / roughness / density: parameter arrays replacing the generated ones (same shapes).
    Returns the SceneDesc and the ids {albedo, roughness (texture ids), medium}."""
    rng = np.random.default_rng(seed)
    sd = SceneDesc()
    N = abi
    alb = _smooth_field(rng, tex_res, 16, 0.25, 0.85, 3) if albedo is None else albedo
    rough = _smooth_field(rng, tex_res, 16, 0.2, 0.7) if roughness is None else roughness
    t_alb = sd.add_texture(alb)
    t_rough = sd.add_texture(rough)
    dens = fbm_density(grid_res, octaves=4, seed=seed + 11) * 1.5 if density is None else density
    med = grid_medium(sd, dens, (0.4, 0.4, 0.4), (1.6, 1.6, 1.6), 0.3, (-4.0, 1.2, -4.0), (4.0, 2.7, 4.0))
    ground_m = sd.add_material(Material([node(N.NODE_IMAGE_COLOUR, p=[t_alb]), node(N.NODE_LAMBERTIAN, [0])], bxdf=1))
    ggx = [1.0, 1.0, 1.45, *COPPER_ETA, *COPPER_K, 0.0]
    sphere_m = sd.add_material(Material([
        node(N.NODE_IMAGE_COLOUR, p=[t_alb]), node(N.NODE_LAMBERTIAN, [0]), node(N.NODE_RGB, p=(0.95, 0.95, 0.95)),
        node(N.NODE_IMAGE_SCALAR, p=[t_rough]), node(N.NODE_MICROFACET_BRDF, [2, 3], p=ggx),
        node(N.NODE_SCALAR, p=[0.5]), node(N.NODE_MIX_BSDF, [1, 4, 5])], bxdf=6))
    g = ground_res
    xs = np.linspace(-6, 6, g + 1)
    gx, gz = np.meshgrid(xs, xs, indexing="ij")
    gv = np.stack([gx.ravel(), np.zeros(gx.size), gz.ravel()], axis=1)
    guv = np.stack([(gx.ravel() + 6) / 12, (gz.ravel() + 6) / 12], axis=1)
    idx = np.arange((g + 1) * (g + 1)).reshape(g + 1, g + 1)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    gf = np.concatenate([np.stack([a, d, c], 1), np.stack([a, c, b], 1)])
    sd.add_mesh(gv, gf, ground_m, uvs=guv)
    base_v, base_f = icosphere(4)
    cells = rng.choice(144, size=n_spheres, replace=False)
    suv = np.stack([np.arctan2(base_v[:, 2], base_v[:, 0]) / (2 * np.pi) + 0.5,
                    np.arccos(np.clip(base_v[:, 1], -1, 1)) / np.pi], axis=1)
    for cell in cells:
        cx, cz = -6 + (cell % 12) + 0.5, -6 + (cell // 12) + 0.5
        r = rng.uniform(0.3, 0.5)
        centre = np.array([cx, r, cz])
        p = base_v * r + centre
        p = p + base_v * (perlin3(p * 4.0, seed=seed + 1) * 0.05)[:, None]
        n = _vertex_normals(p, base_f)
        sd.add_mesh(p, base_f, sphere_m, normals=n, tangents=_tangents(n), uvs=suv)
    for q in range(n_light_quads):
        T = float(rng.uniform(3000, 9000))
        inten = float(10 ** rng.uniform(1, 2.5))
        c = np.array([rng.uniform(-6, 6), rng.uniform(2.8, 4), rng.uniform(-6, 6)])
        s_ = 0.05
        v = np.array([[c[0] - s_, c[1], c[2] - s_], [c[0] + s_, c[1], c[2] - s_], [c[0] + s_, c[1], c[2] + s_],
                      [c[0] - s_, c[1], c[2] + s_]])
        m = sd.add_material(emitter_blackbody(T, inten))
        sd.add_mesh(v, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32), m)
    fog = sd.add_material(passthrough(interior=med))
    v, i = aabb((-4.0, 1.2, -4.0), (4.0, 2.7, 4.0))
    sd.add_mesh(v, i, fog)
    sd.add_env_light(sky(256, 128), intensity=0.5)
    return sd, dict(albedo=t_alb, roughness=t_rough, medium=med)


def inverse_camera(width=512, height=512):
    """Pinhole at (0, 5, 11) looking at the patch centre, SetFov(0.9)."""
    return camera_pinhole((0.0, 5.0, 11.0), width, height, math.pi, -math.atan2(5.0, 11.0), 0.9)


def batch_job(k, seed=0):
    """Job k of the 64-scene batch workload: a patch of the many-light field of log-uniform size
    (about 10k to 2M triangles: icosphere level 3 or 4 spheres plus a ground grid), 8 to 2048
    emissive quads (16 to 4096 emissive triangles) through the light tree, a square resolution
    log-uniform in 64..1024; one job in four adds a grid medium cloud (32^3 or 64^3 fBm) over the
    field and renders with the volumetric path integrator, the others with the path integrator and
    the HDR sky. Returns (SceneDesc, camera, width, height, integrator, max_bounces)."""
    rng = np.random.default_rng(1000 + 7919 * seed + k)
    tris = 10000.0 * 200.0 ** rng.uniform()
    level = 4 if tris > 60000 else 3
    per = 5120 if level == 4 else 1280
    ground = int(max(8, min(160, math.sqrt(tris * 0.1 / 2))))
    n_spheres = int(max(1, min(1600, (tris - 2 * ground * ground) / per)))
    n_quads = int(8 * 256 ** rng.uniform())
    media = k % 4 == 3
    sd = many_light_scene(n_spheres=n_spheres, sphere_level=level, ground_res=ground, n_light_quads=n_quads,
                          env=not media, seed=1000 + k)
    if media:
        res = 64 if rng.uniform() < 0.5 else 32
        dens = fbm_density(res, octaves=4, seed=k) * 1.2
        med = grid_medium(sd, dens, (0.3, 0.3, 0.3), (1.2, 1.1, 1.0), 0.4, (-20.0, 1.5, -20.0), (20.0, 4.0, 20.0))
        fog = sd.add_material(passthrough(interior=med))
        v, i = aabb((-20.0, 1.5, -20.0), (20.0, 4.0, 20.0))
        sd.add_mesh(v, i, fog)
    W = int(64 * 16 ** rng.uniform())
    W = max(64, min(1024, W - W % 8))
    cam = many_light_camera(W, W)
    integ = abi.INTEGRATOR_VOLPATH if media else abi.INTEGRATOR_PATH
    return sd, cam, W, W, integ, 8


def many_light_camera(width=1920, height=1080):
    """Thin lens at (0, 6, 24) looking at the field centre, aperture 0.05, focused on the centre."""
    origin = (0.0, 6.0, 24.0)
    theta = -math.atan2(6.0, 24.0)
    return camera_thin_lens(origin, width, height, math.pi, theta, 0.9, math.hypot(6.0, 24.0), 0.05)
