"""OBJ ingestion (AssetImporter.cpp:5-15 with Assimp's Triangulate and CalcTangentSpace,
MeshImport.cpp:11-92) through the library's host loader: corner vertices, quad splitting, uv flip,
tangent frames and handedness, smoothing across corners, MTL colours, errors. No device needed."""
import math

import numpy as np
import pytest

from lambda_amd import abi, gpu, scene

PLANE = """# two coplanar quads in y = 0 sharing the edge x = 1, uv = (x / 2, z)
mtllib plane.mtl
o Plane
v 0 0 0
v 1 0 0
v 1 0 1
v 0 0 1
v 2 0 0
v 2 0 1
vt 0 0
vt 0.5 0
vt 0.5 1
vt 0 1
vt 1 0
vt 1 1
vn 0 1 0
usemtl Grey
f 1/1/1 4/4/1 3/3/1 2/2/1
f 2/2/1 3/3/1 6/6/1 5/5/1
o Tri
usemtl Glow
f 1/1/1 2/2/1 3/3/1
o Bare
f 1 2 3 4 5
"""
MTL = """newmtl Grey
Kd 0.4 0.5 0.6
newmtl Glow
Kd 0.1 0.1 0.1
Ke 3 2 1
"""


@pytest.fixture()
def plane(tmp_path):
    (tmp_path / "plane.mtl").write_text(MTL)
    p = tmp_path / "plane.obj"
    p.write_text(PLANE)
    return p


def test_meshes_corners_and_uv_flip(plane):
    meshes, mats = gpu.load_obj(plane)
    assert [m["name"] for m in meshes] == ["Plane", "Tri", "Bare"]
    grey, tri, bare = meshes
    assert grey["positions"].shape == (8, 3) and grey["indices"].shape == (4, 3)  # one vertex per corner
    assert tri["positions"].shape == (3, 3) and tri["indices"].shape == (1, 3)
    assert bare["indices"].shape == (3, 3)  # pentagon fan
    assert bare["normals"] is None and bare["tangents"] is None and bare["uvs"] is None
    # uv (u, v) -> (1 - u, 1 - v); the first corner of the first face is vt 1 = (0, 0)
    np.testing.assert_allclose(grey["uvs"][0], [1.0, 1.0])
    np.testing.assert_allclose(grey["uvs"][2], [0.5, 0.0])
    names = [m["name"] for m in mats]
    assert names[:2] == ["Grey", "Glow"]
    assert mats[0]["kd"] == pytest.approx((0.4, 0.5, 0.6)) and mats[0]["ke"] is None
    assert mats[1]["ke"] == pytest.approx((3, 2, 1))
    # the material stays current across `o` statements until the next `usemtl`
    assert grey["material"] == 0 and tri["material"] == 1 and bare["material"] == 1


def test_tangent_frame_and_handedness(plane, tmp_path):
    """The tangent follows dp/du (u grows with x: +x). The stored handedness is the sign of
    dot(cross(n, t), b) for the CalcTangentSpace bitangent (MeshImport.cpp:39-49): mirroring the uv
    layout in v flips it and leaves the tangent alone; mirroring in u flips the tangent. (Assimp
    is absent: its bitangent sign convention is restated from its source, parity unpinned.)"""
    grey = gpu.load_obj(plane)[0][0]
    t = grey["tangents"]
    np.testing.assert_allclose(t[:, :3], np.tile([1.0, 0.0, 0.0], (8, 1)), atol=1e-6)
    assert np.abs(t[:, 3]).min() == 1.0 and len(set(t[:, 3])) == 1
    np.testing.assert_allclose(grey["normals"], np.tile([0.0, 1.0, 0.0], (8, 1)))
    text = PLANE.replace("mtllib plane.mtl\n", "")
    for flip_u, flip_v in ((False, True), (True, False)):
        lines = []
        for ln in text.splitlines():
            if ln.startswith("vt "):
                u, v = map(float, ln.split()[1:3])
                ln = "vt %g %g" % (-u if flip_u else u, -v if flip_v else v)
            lines.append(ln)
        q = tmp_path / ("flip%d%d.obj" % (flip_u, flip_v))
        q.write_text("\n".join(lines) + "\n")
        m = gpu.load_obj(q)[0][0]["tangents"]
        if flip_v:
            np.testing.assert_allclose(m[:, :3], t[:, :3], atol=1e-6)
            assert (m[:, 3] == -t[:, 3]).all()
        else:
            np.testing.assert_allclose(m[:, :3], -t[:, :3], atol=1e-6)


def test_concave_quad_is_split_from_its_concave_corner(tmp_path):
    # a dart: corner 2 (index 1-based 3) is concave
    p = tmp_path / "dart.obj"
    p.write_text("v 0 0 0\nv 2 0 0\nv 1 0 0.5\nv 1 0 2\nf 1 2 3 4\n")
    m = gpu.load_obj(p)[0][0]
    assert m["indices"].tolist() == [[2, 3, 0], [2, 0, 1]]
    # a convex quad starts at corner 0
    p.write_text("v 0 0 0\nv 2 0 0\nv 2 0 2\nv 0 0 2\nf 1 2 3 4\n")
    assert gpu.load_obj(p)[0][0]["indices"].tolist() == [[0, 1, 2], [0, 2, 3]]


def test_tangents_are_smoothed_only_across_matching_corners(tmp_path):
    """Two quads folded at 90 degrees along x = 1 share positions but not normals: their corners keep
    their own tangents; a coplanar pair with a rotated uv layout is merged where tangents agree."""
    fold = """v 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\nv 1 1 0\nv 1 1 1
vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1
vn 0 1 0\nvn -1 0 0
f 1/1/1 4/4/1 3/3/1 2/2/1
f 2/1/2 3/4/2 6/3/2 5/2/2
"""
    p = tmp_path / "fold.obj"
    p.write_text(fold)
    m = gpu.load_obj(p)[0][0]
    t, n = m["tangents"][:, :3], m["normals"]
    assert np.abs((t * n).sum(axis=1)).max() < 1e-6  # tangents lie in the surface
    assert np.allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-6)
    np.testing.assert_allclose(t[:4], np.tile([1, 0, 0], (4, 1)), atol=1e-6)
    np.testing.assert_allclose(t[4:], np.tile([0, 1, 0], (4, 1)), atol=1e-6)


def test_bad_files_are_errors(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 7\n")
    with pytest.raises(gpu.LambdaError):
        gpu.load_obj(p)
    with pytest.raises(gpu.LambdaError):
        gpu.load_obj(tmp_path / "missing.obj")
    p.write_text("v 0 0 0\n")
    with pytest.raises(gpu.LambdaError):
        gpu.load_obj(p)


def test_scene_from_obj_uses_reference_material_mapping(plane):
    sd = scene.SceneDesc()
    ids = sd.add_obj(plane)
    assert len(ids) == 3
    mats = [pl for k, pl in sd.ops if k == "material"]
    # Oren-Nayar with albedo Kd and sigma socket 2 (Glow's Ke is not turned into emission, as
    # BuildMaterial sends emissive models to the missing-material branch)
    assert mats[0].nodes[2][0] == abi.NODE_OREN_NAYAR_BRDF and mats[0].nodes[1][2] == (2.0,)
    assert mats[0].nodes[0][2] == pytest.approx((0.4, 0.5, 0.6))


def test_obj_without_materials_gets_the_missing_material(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 0 1\nf 1 2 3\n")
    sd = scene.SceneDesc()
    sd.add_obj(p)
    mats = [pl for k, pl in sd.ops if k == "material"]
    assert mats[0].nodes[0][2] == (1.0, 0.0, 1.0)
