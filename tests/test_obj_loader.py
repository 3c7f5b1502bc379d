"""OBJ meshes (raytracer/obj_loader.py, the reference's ObjLoader.Load
without subdivision) and mesh entities in the YAML front-end: the parse
rules on small known-answer files, and the reference's own model files
(tests/golden/models) against what their geometry fixes — a unit sphere
whose normals are its positions and whose UVs are its longitude/latitude, a
torus with R = 1.5 and r = 0.5, a closed unit cube of area 6 (no GPU needed)."""
import os

import numpy as np
import pytest

from raytracer import abi, obj_loader, scene_loader

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODELS = os.path.join(HERE, "models")


def tris_of(m):
    return m.tris.reshape(-1, 3, 3)


def test_fan_triangulation_and_index_forms():
    text = """# quad, pentagon, negative indices
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v -1 0.5 0
o ignored
g ignored
s 1
usemtl ignored
f 1 2 3 4
f 1 2 3 4 5
f -5 -4 -3
"""
    m = obj_loader.parse(text)
    assert len(m) == 6 and m.normals is None and m.uvs is None and m.vertex_count == 5
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [-1, 0.5, 0]], np.float32)
    want = [(0, 1, 2), (0, 2, 3), (0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 1, 2)]
    assert np.array_equal(tris_of(m), P[np.array(want)])


def test_normal_and_uv_channels():
    base = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvn 0 0 2\nvn 0 1 1\nvn 1 0 1\nvt 0 0\nvt 1 0\nvt 0 1 0.5\n"
    m = obj_loader.parse(base + "f 1/1/1 2/2/2 3/3/3\nf 2/2/2 4/1/1 3/3/3\n")
    assert m.normals.shape == (2, 9) and m.uvs.shape == (2, 6)
    assert np.array_equal(m.normals[0], [0, 0, 2, 0, 1, 1, 1, 0, 1])        # raw: SmoothTriangle normalises
    assert np.array_equal(m.uvs[1], [1, 0, 0, 0, 0, 1])
    # v//vn: smooth without UVs (barycentric)
    m = obj_loader.parse(base + "f 1//1 2//2 3//3\n")
    assert m.normals is not None and m.uvs is None
    # v/vt only: a flat Triangle (its UVs are dropped)
    m = obj_loader.parse(base + "f 1/1 2/2 3/3\n")
    assert m.normals is None and m.uvs is None
    # a channel some faces lack is dropped for the whole file
    m = obj_loader.parse(base + "f 1/1/1 2/2/2 3/3/3\nf 2 4 3\n")
    assert m.normals is None and m.uvs is None and len(m) == 2
    m = obj_loader.parse(base + "f 1/1/1 2/2/2 3/3/3\nf 2//2 4//1 3//3\n")
    assert m.normals is not None and m.uvs is None
    # negative attribute indices count from the lists so far
    m = obj_loader.parse(base + "f 1/-3/-3 2/-2/-2 3/-1/-1\n")
    assert np.array_equal(m.normals[0], [0, 0, 2, 0, 1, 1, 1, 0, 1]) and np.array_equal(m.uvs[0], [0, 0, 1, 0, 0, 1])
    # an out-of-range normal index on a used face: the reference throws
    with pytest.raises(obj_loader.ObjLoadError):
        obj_loader.parse(base + "f 1/1/1 2/2/2 3/3/9\n")
    with pytest.raises(obj_loader.ObjLoadError, match="zero-length"):
        obj_loader.parse(base + "vn 0 0 0\nf 1//4 2//2 3//3\n")


def test_degenerate_triangles_are_skipped_in_float32():
    # cross product squared length: (a * b)^2 against 1e-12 in float32
    def tri(a, b):
        return f"v 0 0 0\nv {a!r} 0 0\nv 0 {b!r} 0\nf 1 2 3\n"

    assert obj_loader.parse("v 0 0 0\nv 1 1 1\nv 2 2 2\nf 1 2 3\n") is None       # collinear
    assert obj_loader.parse(tri(1e-3, 0.9e-3)) is None                            # 8.1e-13 < 1e-12
    assert len(obj_loader.parse(tri(1e-3, 1.1e-3))) == 1                          # 1.21e-12
    m = obj_loader.parse("v 0 0 0\nv 1 0 0\nv 2 0 0\nv 1 1 0\nf 1 2 3 4\n")       # first fan triangle degenerate
    assert len(m) == 1 and np.array_equal(tris_of(m)[0], [[0, 0, 0], [2, 0, 0], [1, 1, 0]])


def test_malformed_lines_and_warnings():
    w = []
    text = ("﻿v 0 0 0\r\nv 1 2\r\nv\t5\t5\t5\nv 1 0 0\rv 0 1 0\n"
            "vn 1 0\nvt 0.5\nf 1 2\nf 1 0 2 3\nf 1 2 3 9\nf 1/x/1 2 3\n")
    m = obj_loader.parse(text.lstrip("﻿"), w)
    assert m.vertex_count == 3                     # 'v 1 2' is malformed; 'v<TAB>' is not a vertex line
    joined = "\n".join(w)
    for s in ("line 2: malformed vertex position", "malformed vertex normal", "malformed texture coordinate",
              "fewer than 3 vertices", "malformed face vertex '0'", "malformed face vertex '9'"):
        assert s in joined, s
    assert len(m) == 3                             # 'f 1 0 2 3' and 'f 1 2 3 9' keep three vertices each
    assert obj_loader.parse("# nothing\n\n", w) is None and "zero faces" in w[-1]


@pytest.mark.parametrize("tok,want", [("1e-3", np.float32(1e-3)), ("+.5", 0.5), ("1.", 1.0), ("-0", -0.0),
                                      ("  2.5\t", 2.5), ("Infinity", np.inf), ("-infinity", -np.inf),
                                      ("1E+02", 100.0), ("16777217", 16777216.0),
                                      ("16777217.000000001", 16777218.0), ("16777218.999999999", 16777218.0),
                                      ("1_0", None), ("0x10", None), ("inf", None), ("1e", None), ("", None),
                                      ("1,5", None)])
def test_float_grammar_and_single_rounding(tok, want):
    got = obj_loader.parse_float32(tok)
    if want is None:
        assert got is None
    else:
        assert got is not None and np.float32(got).tobytes() == np.float32(want).tobytes()


def test_reference_models_geometry():
    cube = obj_loader.load(os.path.join(MODELS, "test-cube.obj"))
    t = tris_of(cube).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    assert len(cube) == 12 and cube.normals is None and abs(area.sum() - 6.0) < 1e-9
    assert np.abs(t).max() == 0.5

    sph = obj_loader.load(os.path.join(MODELS, "test-sphere.obj"))
    p = tris_of(sph).astype(np.float64)
    n = sph.normals.reshape(-1, 3, 3).astype(np.float64)
    uv = sph.uvs.reshape(-1, 3, 2).astype(np.float64)
    assert len(sph) == 960 and np.abs(np.linalg.norm(p, axis=2) - 1).max() < 1e-5
    assert np.abs(n - p).max() < 1e-5                      # normals are the unit positions
    lat = np.arccos(np.clip(p[..., 1], -1, 1)) / np.pi
    assert np.abs(uv[..., 1] - lat).max() < 1e-5
    lon = np.arctan2(p[..., 2], p[..., 0]) / (2 * np.pi) % 1.0
    d = np.abs(uv[..., 0] - lon)
    pole = np.abs(p[..., 1]) > 0.9999                         # longitude is free at the poles
    assert np.minimum(d, 1 - d)[~pole].max() < 1e-5

    tor = obj_loader.load(os.path.join(MODELS, "test-torus-mesh.obj"))
    p = tris_of(tor).astype(np.float64)
    ring = np.linalg.norm(p[..., [0, 2]], axis=2)
    assert len(tor) == 1536 and np.abs((ring - 1.5) ** 2 + p[..., 1] ** 2 - 0.25).max() < 1e-5
    c = p.copy()
    c[..., 1] = 0
    c *= (1.5 / ring)[..., None]
    assert np.abs(tor.normals.reshape(-1, 3, 3) - (p - c) / 0.5).max() < 1e-5

    pot = obj_loader.load(os.path.join(MODELS, "utah-teapot-low.obj"))
    assert len(pot) == 3298 and pot.normals is not None and pot.uvs is not None
    assert np.abs(np.linalg.norm(pot.normals.reshape(-1, 3), axis=1) - 1).max() < 1e-3


def test_mesh_entities_in_the_scene_front_end():
    w = []
    d = scene_loader.load(os.path.join(HERE, "scenes", "mesh-trio.yaml"), 64, 48, warnings=w)
    assert w == []
    kinds, mesh = d.prims["kind"], d.prims["mesh"]
    # ground, sphere 960, cube 12, three tori 1536, teapot 3298, then the area light's proxy
    counts = [960, 12, 1536, 1536, 1536, 3298]
    assert len(d.prims) == 1 + sum(counts) + 1 and kinds[0] == abi.RT_PRIM_PLANE and kinds[-1] == abi.RT_PRIM_QUAD
    ids = mesh[1:-1]
    assert list(np.unique(ids)) == list(range(6)) and list(np.bincount(ids)) == counts
    # each mesh keeps its object-space faces under a TRANSFORM node
    nodes, _, roots = d.hierarchy

    def matrix_of(first):
        for r in roots:
            if nodes["kind"][r] == abi.RT_NODE_TRANSFORM and nodes["first"][nodes["child"][r]] == first:
                return nodes["matrix"][r].reshape(4, 4).astype(np.float64)
        raise AssertionError(first)

    def world(first, pts):
        M = matrix_of(first)
        return np.c_[pts.reshape(-1, 3).astype(np.float64), np.ones(pts.size // 3)] @ M

    sph = obj_loader.load(os.path.join(MODELS, "test-sphere.obj"))
    s = d.prims["v"][1:961, :9].reshape(-1, 3, 3)
    assert np.array_equal(s, tris_of(sph))
    assert np.allclose(world(1, s)[:, :3], (tris_of(sph) + np.float32([-3, 1, 0])).reshape(-1, 3), atol=1e-6)
    assert np.array_equal(d.normals[1:961], sph.normals) and np.array_equal(d.texcoords[1:961], sph.uvs)
    cube = world(961, d.prims["v"][961:973, :9])[:, :3]
    assert np.allclose(cube.min(0), [0.75, 0.0, -0.75]) and np.allclose(cube.max(0), [2.25, 1.5, 0.75])
    assert not d.normals[961:973].any()
    # the steel torus: rotate X 90 turns the ring's +y axis to +z (its normals
    # turn with it when the renderer maps them by the transposed inverse)
    first = 1 + 960 + 12 + 2 * 1536
    st = slice(first, first + 1536)
    tor = obj_loader.load(os.path.join(MODELS, "test-torus-mesh.obj"))
    R = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], np.float64)
    want = tor.tris.reshape(-1, 3).astype(np.float64) @ R.T * 0.6 + [0, 1.5, 3]
    assert np.array_equal(d.prims["v"][st, :9], tor.tris.reshape(-1, 9))
    assert np.allclose(world(first, d.prims["v"][st, :9])[:, :3], want, atol=1e-5)
    M = matrix_of(first)
    nn = d.normals[st].reshape(-1, 3).astype(np.float64) @ np.linalg.inv(M)[:3, :3].T
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    assert np.allclose(nn, tor.normals.reshape(-1, 3).astype(np.float64) @ R.T, atol=1e-6)


def test_textured_scene_loads():
    w = []
    d = scene_loader.load(os.path.join(HERE, "scenes", "textured.yaml"), 64, 48, warnings=w)
    assert w == [] and len(d.images) == 3 and d.images[0][1] == (4.0, 4.0) and d.images[2][1] == (2.0, 1.0)
    kinds = d.mats["kind"][d.prims["material"]]
    assert (d.mats["flags"][d.prims["material"]][:962] & abi.RT_MATF_IMAGE).all()
    assert kinds[1] == abi.RT_MAT_DISNEY and d.normals is not None and d.texcoords is not None


def test_mesh_entity_skips_and_refusals(tmp_path):
    cube = os.path.join(MODELS, "test-cube.obj")
    base = {"lights": [], "world": {"sky": {"color": [0, 0, 0]}},
            "materials": [{"id": "e", "type": "emissive", "color": [1, 1, 1]}]}
    w = []
    d = scene_loader.load_data(dict(base, entities=[{"type": "mesh", "name": "a"},
                                                    {"type": "mesh", "name": "b", "path": "missing.obj"},
                                                    {"type": "mesh", "name": "c", "path": cube}]), 8, 8, warnings=w)
    assert len(d.prims) == 12 and "requires a 'path'" in w[0] and "not found" in w[1]
    empty = tmp_path / "empty.obj"
    empty.write_text("# no faces\n")
    w = []
    scene_loader.load_data(dict(base, entities=[{"type": "mesh", "name": "z", "path": "empty.obj"}]), 8, 8,
                           base_dir=str(tmp_path), warnings=w)
    assert any("zero faces" in x for x in w) and any("failed to load" in x for x in w)
    with pytest.raises(scene_loader.SceneLoadError, match="subdivision"):
        scene_loader.load_data(dict(base, entities=[{"type": "mesh", "path": cube, "subdivision_iterations": 2}]),
                               8, 8)
    # any scale of a smooth mesh: its object-space faces and normals under a
    # TRANSFORM node (the renderer maps the interpolated normal by the
    # transposed inverse)
    sph = os.path.join(MODELS, "test-sphere.obj")
    d = scene_loader.load_data(dict(base, entities=[{"type": "mesh", "path": sph, "scale": [1, 2, 1]}]), 8, 8)
    nodes, _, roots = d.hierarchy
    assert nodes["kind"][roots[0]] == abi.RT_NODE_TRANSFORM
    assert np.allclose(np.diag(nodes["matrix"][roots[0]].reshape(4, 4)), [1, 2, 1, 1])
    assert np.array_equal(d.normals, obj_loader.load(sph).normals)
    # a flat mesh takes any scale; an emissive mesh cannot be a geometry light here
    d = scene_loader.load_data(dict(base, entities=[{"type": "mesh", "path": cube, "scale": [1, 2, 3]}]), 8, 8)
    assert np.allclose(d.prims["v"][:, :9].reshape(-1, 3).max(0), [0.5, 0.5, 0.5])
    assert np.allclose(np.diag(d.hierarchy[0]["matrix"][d.hierarchy[2][0]].reshape(4, 4)), [1, 2, 3, 1])
    with pytest.raises(ValueError, match="emissive meshes"):
        scene_loader.load_data(dict(base, entities=[{"type": "mesh", "path": cube, "material": "e"}]), 8, 8)
