"""SmoothTriangle (SmoothTriangle.cs:68-197) in the oracle, pinned by what
the geometry fixes rather than by retyping its formula:

  * the first-hit normal AOV of a one-sample render equals
    normalize(w0 N0 + w1 N1 + w2 N2) evaluated in float64 from the hit's
    barycentrics, with the sign of the face-normal front test (both sides);
  * vertex normals equal to the face normal reproduce the flat triangle;
  * the UV-gradient tangent frame: a UV map u = x, v = y gives the flat
    triangle's frame (its edges lie along x and y), u = y, v = x turns an
    anisotropic highlight (and changes the image), a degenerate map falls
    back to the edges;
  * a smooth icosphere's shading normals are far closer to the true sphere
    normal at the hit point than its face normals.
Plus the builder and YAML plumbing and the refusals (no GPU needed)."""
import numpy as np
import pytest

from raytracer import abi, scene_loader, scenes

V0, V1, V2 = (-2.0, -2.0, 0.0), (2.0, -2.0, 0.0), (-2.0, 2.0, 0.0)     # e1 = 4x, e2 = 4y, face normal +z
TILTED = [(0.3, 0.2, 1.0), (-0.4, 0.1, 1.0), (0.1, -0.5, 0.8)]


def plate(W, H, normals=None, uvs=None, back=False, mat=None):
    b = scenes.SceneBuilder()
    if mat is None:
        m = b.lambertian((0.7, 0.7, 0.7))
    else:
        m = b.disney((0.9, 0.8, 0.6), **mat)
    b.triangle(np.array(V0, np.float32), np.array(V1, np.float32), np.array(V2, np.float32), m,
               normals=normals, uvs=uvs)
    b.directional_light((-0.6, -0.3, -1.0) if not back else (-0.6, -0.3, 1.0), intensity=2.0)
    z = -3.0 if back else 3.0
    cam = scenes.make_camera((-0.2, -0.4, z), (-0.6, -0.6, 0.0), 60.0, W / H)
    return b.build(cam)


def expected_normals(normals, u, v):
    n = np.asarray(normals, np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    u = u.astype(np.float64)[:, None]
    v = v.astype(np.float64)[:, None]
    s = (1.0 - u - v) * n[0] + u * n[1] + v * n[2]
    return s / np.linalg.norm(s, axis=1, keepdims=True)


def film_normal(o):
    return o["normal_a"].reshape(3, -1).T


@pytest.mark.parametrize("back", [False, True])
def test_shading_normal_is_the_interpolated_vertex_normal(oracle, back):
    W, H = 40, 30
    desc = plate(W, H, TILTED, back=back)
    st = abi.settings_defaults(W, H, 1, 4)
    st.capture_aovs = 1
    osc = oracle.OracleScene(desc)
    prim, t, u, v, _, _ = osc.trace_primary(st, 0, 1)
    hit = prim == 0
    assert 0.3 < hit.mean() < 0.95
    want = expected_normals(TILTED, u[hit], v[hit]) * (-1.0 if back else 1.0)
    got = film_normal(osc.render(st, aovs=True))[hit]
    err = np.abs(got - want).max()
    print(f"back={back}: {hit.sum()} hits, max |normal - float64 formula| {err:.2e}", flush=True)
    assert err < 2e-6
    # a plausible slip (the weights of V0 and V1 exchanged) is far outside that bound
    assert np.abs(expected_normals(TILTED, 1.0 - u[hit] - v[hit], v[hit]) * (-1.0 if back else 1.0) - got).max() > 1e-2


def test_face_normals_reproduce_the_flat_triangle(oracle):
    W, H = 40, 30
    st = abi.settings_defaults(W, H, 8, 6)
    st.capture_aovs = 1
    flat = oracle.OracleScene(plate(W, H)).render(st, aovs=True)
    smooth = oracle.OracleScene(plate(W, H, [(0.0, 0.0, 2.0)] * 3)).render(st, aovs=True)
    assert np.array_equal(film_normal(flat), film_normal(smooth))
    rel = abs(smooth["beauty"].mean() - flat["beauty"].mean()) / flat["beauty"].mean()
    assert flat["beauty"].mean() > 0 and rel < 1e-6


ANISO = dict(metallic=1.0, roughness=0.35, anisotropic=0.9)


def test_uv_gradient_tangent_frame(oracle):
    W, H = 48, 36
    st = abi.settings_defaults(W, H, 16, 4)
    face = [(0.0, 0.0, 1.0)] * 3
    flat = oracle.OracleScene(plate(W, H, mat=ANISO)).render(st)["beauty"]
    xy = [(x, y) for x, y, _ in (V0, V1, V2)]
    same = oracle.OracleScene(plate(W, H, face, uvs=xy, mat=ANISO)).render(st)["beauty"]
    swapped = oracle.OracleScene(plate(W, H, face, uvs=[(y, x) for x, y in xy], mat=ANISO)).render(st)["beauty"]
    degenerate = oracle.OracleScene(plate(W, H, face, uvs=[(0.5, 0.5)] * 3, mat=ANISO)).render(st)["beauty"]
    bary = oracle.OracleScene(plate(W, H, face, mat=ANISO)).render(st)["beauty"]

    def rel(a):
        return float(np.abs(a - flat).mean() / flat.mean())

    print(f"rel diff: u=x,v=y {rel(same):.2e}; swapped {rel(swapped):.2e}; degenerate {rel(degenerate):.2e}; "
          f"barycentric {rel(bary):.2e}", flush=True)
    assert rel(same) < 1e-5 and rel(degenerate) < 1e-5 and rel(bary) < 1e-5
    assert rel(swapped) > 1e-2


def test_smooth_icosphere_normals_follow_the_sphere(oracle):
    W, H = 48, 48
    v, f = scenes.icosphere(2)

    def scene(smooth):
        b = scenes.SceneBuilder()
        m = b.lambertian((0.6, 0.6, 0.6))
        b.mesh(v, f, m, normals=v if smooth else None)
        b.directional_light((-1.0, -1.0, -1.0), intensity=2.0)
        return b.build(scenes.make_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 45.0, 1.0))

    st = abi.settings_defaults(W, H, 1, 2)
    st.capture_aovs = 1
    errs = []
    for smooth in (False, True):
        d = scene(smooth)
        osc = oracle.OracleScene(d)
        prim, _, u, w, _, _ = osc.trace_primary(st, 0, 1)
        hit = prim >= 0
        assert hit.mean() > 0.3
        n = film_normal(osc.render(st, aovs=True))[hit]
        # the hit point from its barycentrics on the hit face
        tri = d.prims["v"][prim[hit], :9].astype(np.float64).reshape(-1, 3, 3)
        uu, ww = u[hit].astype(np.float64)[:, None], w[hit].astype(np.float64)[:, None]
        p = (1.0 - uu - ww) * tri[:, 0] + uu * tri[:, 1] + ww * tri[:, 2]
        true = p / np.linalg.norm(p, axis=1, keepdims=True)
        errs.append(float(np.degrees(np.arccos(np.clip((n * true).sum(1), -1, 1))).mean()))
    print(f"mean angle to the sphere normal: flat {errs[0]:.3f} deg, smooth {errs[1]:.3f} deg", flush=True)
    assert errs[1] < errs[0] / 4


def test_builder_and_yaml_tables():
    d = plate(8, 8, TILTED, uvs=[(0, 0), (1, 0), (0, 1)])
    assert d.normals.shape == (1, 9) and np.allclose(d.normals[0], np.array(TILTED, np.float32).ravel())
    assert d.desc.normals and d.desc.texcoords
    doc = {"lights": [{"type": "sun", "direction": [0, -1, 0]}],
           "materials": [{"id": "m", "type": "lambertian"}],
           "entities": [{"type": "sphere", "material": "m"},
                        {"type": "smooth_triangle", "material": "m", "v0": [0, 0, 0], "v1": [1, 0, 0],
                         "v2": [0, 1, 0], "n0": [0, 0, 1], "n1": [0, 1, 1], "n2": [1, 0, 1],
                         "uv0": [0, 0], "uv1": [2, 0], "uv2": [0, 2], "rotate": [0, 0, 90], "scale": 2},
                        {"type": "triangle", "material": "m", "v0": [0, 0, 0], "v1": [1, 0, 0],
                         "v2": [0, 1, 0], "n0": [0, 0, 1], "n1": [0, 0, 1], "n2": [0, 0, 1]},
                        {"type": "triangle", "material": "m"}]}
    d = scene_loader.load_data(doc, 16, 16)
    n = d.normals
    assert n.shape == (4, 9) and not n[0].any() and not n[3].any()
    # the transformed smooth triangle keeps its object-space normals; its
    # TRANSFORM node's transposed inverse maps (0,1,1) -> (-1,0,1) and
    # (1,0,1) -> (0,1,1) (rotate Z 90; the uniform scale does not turn them)
    assert np.allclose(n[1].reshape(3, 3), [[0, 0, 1], [0, 1, 1], [1, 0, 1]], atol=1e-6)
    nodes, _, roots = d.hierarchy
    M = nodes["matrix"][roots[1]].reshape(4, 4).astype(np.float64)
    assert nodes["kind"][roots[1]] == abi.RT_NODE_TRANSFORM and nodes["first"][nodes["child"][roots[1]]] == 1
    w = n[1].reshape(3, 3).astype(np.float64) @ np.linalg.inv(M)[:3, :3].T
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    want = np.array([[0, 0, 1], [-1, 0, 1], [0, 1, 1]], np.float64)
    assert np.allclose(w, want / np.linalg.norm(want, axis=1, keepdims=True), atol=1e-6)
    assert np.allclose(n[2], [0, 0, 1] * 3)
    assert np.allclose(d.texcoords[1], [0, 0, 2, 0, 0, 2]) and np.allclose(d.texcoords[2], [0, 0, 1, 0, 0, 1])
    # any scale now: a non-uniform one stays on the TRANSFORM node
    d = scene_loader.load_data({"lights": [], "entities": [dict(doc["entities"][1], scale=[1, 2, 1])]}, 8, 8)
    assert np.allclose(np.diag(d.hierarchy[0]["matrix"][d.hierarchy[2][0]].reshape(4, 4))[:3] @ [0, 0, 1], 1.0)


def test_refusals(oracle):
    b = scenes.SceneBuilder()
    em = b.emissive((1.0, 1.0, 1.0), 5.0)
    b.triangle(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32), np.array([0, 1, 0], np.float32), em,
               normals=[(0, 0, 1)] * 3)
    with pytest.raises(ValueError, match="smooth"):
        b.build(scenes.make_camera((0, 0, 3), (0, 0, 0), 40.0, 1.0))
    with pytest.raises(RuntimeError, match="smooth"):
        oracle.OracleScene(smooth_light_scene())


def smooth_light_scene():
    """The tilted plate made an emissive geometry light (no SampleAt for
    interpolated normals: refused)."""
    d = plate(8, 8, TILTED)
    mats = d.mats.copy()
    mats["kind"][d.prims["material"][0]] = abi.RT_MAT_EMISSIVE
    lights = np.zeros(1, scenes.LIGHT_DTYPE)
    lights["kind"] = abi.RT_LIGHT_GEOMETRY
    lights["primitive"] = 0
    lights["shadow_samples"] = 1
    return scenes.SceneDesc(d.prims, mats, lights, d.camera, normals=d.normals)
