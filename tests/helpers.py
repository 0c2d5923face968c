"""Shared test helpers: scene construction on both backends, comparisons."""
import numpy as np

import oracle
from lambda_amd import abi, scene


def oracle_scene(sd, strategy=abi.LIGHTS_TREE):
    o = oracle.OracleScene()
    keep = sd.push(o, strategy)
    return o, keep


def desc(camera, spp_begin=0, spp_end=1, max_bounces=8, min_bounces=3, spectral=abi.SPECTRAL_RGB,
         estimator=abi.ESTIMATOR_CORRECTED, seed=0):
    d = abi.RenderDesc()
    d.camera = camera
    d.integrator = abi.INTEGRATOR_PATH
    d.spp_begin, d.spp_end = spp_begin, spp_end
    d.max_bounces, d.min_bounces = max_bounces, min_bounces
    d.spectral_mode, d.estimator_mode, d.seed = spectral, estimator, seed
    d.pass_max_depth, d.pass_power = 1.0, 1.0
    d.pass_background[:] = [1.0, 0.0, 1.0]
    return d


def furnace_box(albedo=0.5, emit=1.0, subdivide=1):
    """Closed cube [-1,1]^3, faces facing inwards, Lambertian albedo `albedo`
    and uniform emission radiance `emit` (intensity = pi * emit, MeshLight.cpp:95).
    Every face is split into 2*subdivide^2 triangles (many triangle lights)."""
    sd = scene.SceneDesc()
    mat = scene.Material([scene.node(abi.NODE_RGB, p=(1, 1, 1)),
                          scene.node(abi.NODE_RGB, p=(albedo, albedo, albedo)),
                          scene.node(abi.NODE_LAMBERTIAN, [1])], bxdf=2, emission=0,
                         intensity=float(np.float32(np.pi) * np.float32(emit)))
    sd.add_material(mat)
    faces = []
    for ax in range(3):
        for s in (-1.0, 1.0):
            faces.append((ax, s))
    for ax, s in faces:
        u_ax, v_ax = [a for a in range(3) if a != ax]
        verts, tris = [], []
        n = subdivide
        for i in range(n + 1):
            for j in range(n + 1):
                p = [0.0, 0.0, 0.0]
                p[ax] = s
                p[u_ax] = -1 + 2 * i / n
                p[v_ax] = -1 + 2 * j / n
                verts.append(p)
        for i in range(n):
            for j in range(n):
                a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
                tris += [[a, b, c], [a, c, d]]
        verts = np.array(verts, dtype=np.float32)
        tris = np.array(tris, dtype=np.uint32)
        # orient so that Ng = (v1-v0)x(v2-v0) points into the box (-s along ax)
        v0, v1, v2 = verts[tris[0, 0]], verts[tris[0, 1]], verts[tris[0, 2]]
        ng = np.cross(v1 - v0, v2 - v0)
        if ng[ax] * s > 0:
            tris = tris[:, [0, 2, 1]]
        sd.add_mesh(verts, tris, 0)
    return sd


def rel_close(a, b, rel=1e-4, abs_=1e-7):
    return np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)) + abs_
