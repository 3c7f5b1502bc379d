"""YAML scene front-end (raytracer/scene_loader.py): the reference's
SceneLoader assembly (object order, light proxies, geometry lights,
defaults) on the reference's own test scene and on the plain-method Cornell
config written in its schema, plus transforms, camera selection, defaults
and the refusal of unsupported features (no GPU needed)."""
import math
import os

import numpy as np
import pytest

from raytracer import abi, scene_loader, scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def _same_tables(a, b):
    for k in ("prims", "mats", "lights"):
        x, y = getattr(a, k), getattr(b, k)
        assert x.dtype == y.dtype and len(x) == len(y), k
        assert x.tobytes() == y.tobytes(), k
    assert a.sky == b.sky and a.medium == b.medium
    for f, _ in abi.rt_camera._fields_:
        assert np.array_equal(np.asarray(getattr(a.camera, f)), np.asarray(getattr(b.camera, f))), f


def test_cornell_yaml_equals_the_builder_scene():
    got = scene_loader.load(os.path.join(HERE, "cornell-c1.yaml"), 256, 256)
    _same_tables(got, scenes.cornell(256, 256))


def test_caustic_lens_scene_assembly():
    d = scene_loader.load(os.path.join(HERE, "caustic-lens.yaml"), 320, 240, shadow_samples_override=None)
    kinds = list(d.prims["kind"])
    # ground plane, the glass sphere, then the area light's back-face-culled proxy quad
    assert kinds == [abi.RT_PRIM_PLANE, abi.RT_PRIM_SPHERE, abi.RT_PRIM_QUAD]
    assert d.prims["flags"][2] == abi.RT_PRIMF_BACKFACE_CULLED
    floor, glass, proxy = (d.mats[i] for i in d.prims["material"])
    assert floor["kind"] == abi.RT_MAT_LAMBERTIAN and np.allclose(floor["color"], 0.7)
    assert glass["kind"] == abi.RT_MAT_DIELECTRIC and glass["ior"] == np.float32(1.5)
    assert np.allclose(glass["color"], 0.5)          # no colour: CreateMaterial's SolidColor(0.5)
    assert proxy["kind"] == abi.RT_MAT_EMISSIVE and proxy["flags"] & abi.RT_MATF_LIGHT_PROXY
    assert len(d.lights) == 1
    L = d.lights[0]
    assert L["kind"] == abi.RT_LIGHT_AREA and L["shadow_samples"] == 4 and L["intensity"] == np.float32(12.0)
    assert L["primitive"] == 2
    assert np.allclose(d.prims["v"][1][:4], [0, 1.4, 0, 1.0])
    assert d.sky == (0.0, 0.0, 0.0)
    c = d.camera
    assert list(c.look_from) == [0.0, 4.5, 6.0] and c.vfov_deg == 40.0
    assert c.aspect == np.float32(320) / np.float32(240) and c.focus_dist == 1.0 and c.aperture == 0.0
    # -S override reaches the lights
    d4 = scene_loader.load(os.path.join(HERE, "caustic-lens.yaml"), 320, 240, shadow_samples_override=1)
    assert d4.lights[0]["shadow_samples"] == 1


def test_unsupported_phase_is_refused():
    with pytest.raises(scene_loader.SceneLoadError, match="phase 'hg'"):
        scene_loader.load(os.path.join(HERE, "firefly-stress.yaml"), 64, 64)


def test_transforms_are_applied_in_reference_order():
    doc = {"lights": [], "materials": [{"id": "m", "type": "lambertian", "color": [1, 1, 1]}],
           "entities": [{"type": "triangle", "material": "m", "v0": [1, 0, 0], "v1": [0, 1, 0], "v2": [0, 0, 1],
                         "scale": 2.0, "rotate": [90, 0, 90], "translate": [1, 2, 3]},
                        {"type": "sphere", "material": "m", "center": [1, 0, 0], "radius": 0.5, "scale": 3,
                         "translate": [0, 0, -1]},
                        {"type": "quad", "material": "m", "q": [0, 0, 0], "u": [1, 0, 0], "v": [0, 1, 0],
                         "rotate": [0, 90, 0]}]}
    d = scene_loader.load_data(doc, 64, 64)

    def ref(p):    # scale, then rotate about X (90), Y (0), Z (90), then translate
        x, y, z = (2.0 * c for c in p)
        y, z = y * math.cos(math.pi / 2) - z * math.sin(math.pi / 2), y * math.sin(math.pi / 2) + z * math.cos(math.pi / 2)
        x, y = x * math.cos(math.pi / 2) - y * math.sin(math.pi / 2), x * math.sin(math.pi / 2) + y * math.cos(math.pi / 2)
        return [x + 1, y + 2, z + 3]

    # each entity keeps its object-space primitive under a TRANSFORM node
    nodes, _, roots = d.hierarchy
    assert len(roots) == 3 and np.all(nodes["kind"][roots] == abi.RT_NODE_TRANSFORM)

    def world(k, p):      # Vector3.Transform with the node's row-vector matrix
        M = nodes["matrix"][roots[k]].reshape(4, 4).astype(np.float64)
        return np.array([*p, 1.0]) @ M

    def world_vec(k, v):
        return np.array(v, np.float64) @ nodes["matrix"][roots[k]].reshape(4, 4)[:3, :3]

    tri = d.prims["v"][0][:9].reshape(3, 3)
    assert np.allclose(tri, np.eye(3))
    for k, p in enumerate(([1, 0, 0], [0, 1, 0], [0, 0, 1])):
        assert np.allclose(world(0, p)[:3], ref(p), atol=1e-6)
    assert np.allclose(d.prims["v"][1][:4], [1, 0, 0, 0.5])
    assert np.allclose(world(1, [1, 0, 0])[:3], [3, 0, -1]) and np.allclose(world_vec(1, [0.5, 0, 0]), [1.5, 0, 0])
    assert np.allclose(world_vec(2, [1, 0, 0]), [0, 0, -1], atol=1e-6)   # rotate Y 90: u = +x -> -z
    for k in range(3):
        M = nodes["matrix"][roots[k]].reshape(4, 4).astype(np.float64)
        assert np.allclose(nodes["inverse"][roots[k]].reshape(4, 4) @ M, np.eye(4), atol=1e-6)
    # an emissive entity keeps the transform applied to its points (its geometry
    # light samples world-space primitives); so does every entity of a scene with
    # image textures, where a non-uniformly scaled sphere is refused
    doc["materials"].append({"id": "e", "type": "emissive", "color": [1, 1, 1], "intensity": 2})
    doc["entities"][0]["material"] = "e"
    d = scene_loader.load_data(doc, 64, 64)
    tri = d.prims["v"][0][:9].reshape(3, 3)
    for k, p in enumerate(([1, 0, 0], [0, 1, 0], [0, 0, 1])):
        assert np.allclose(tri[k], ref(p), atol=1e-6)
    assert len(d.hierarchy[2]) == 3        # the emissive triangle and the two transforms
    img = os.path.join(HERE, "..", "textures", "checkerboard.png")
    with pytest.raises(scene_loader.SceneLoadError, match="non-uniform"):
        scene_loader.load_data({"lights": [], "materials": [
            {"id": "t", "type": "lambertian", "texture": {"type": "image", "path": img}}],
            "entities": [{"type": "sphere", "scale": [1, 2, 1]}]}, 8, 8)


def test_defaults_camera_selection_and_focus():
    cams = [{"name": "a", "position": [0, 0, 5], "look_at": [0, 0, 0], "fov": 30},
            {"name": "Close", "position": [0, 0, 2], "look_at": [0, 0, 0], "focal_pos": [0.5, 0.0, 0.5],
             "aperture": 0.1}]
    doc = {"lights": [], "cameras": cams, "entities": [{"type": "sphere", "material": "missing"}]}
    d = scene_loader.load_data(doc, 100, 50)
    assert list(d.camera.look_from) == [0, 0, 5] and d.camera.vfov_deg == 30.0 and d.camera.aspect == 2.0
    m = d.mats[d.prims["material"][0]]
    assert m["kind"] == abi.RT_MAT_LAMBERTIAN and np.allclose(m["color"], 0.5)    # missing id: grey Lambertian
    d = scene_loader.load_data(doc, 100, 50, camera="close")
    assert list(d.camera.look_from) == [0, 0, 2] and d.camera.focus_dist == np.float32(1.5)
    assert d.camera.aperture == np.float32(0.1) and d.camera.vfov_deg == 60.0
    assert list(scene_loader.load_data(doc, 10, 10, camera="1").camera.look_from) == [0, 0, 2]
    d = scene_loader.load_data({"lights": []}, 10, 10)
    assert list(d.camera.look_from) == [0, 1, -5] and d.sky == scene_loader.DEFAULT_SKY and len(d.prims) == 0


def test_disney_fields_and_ground_material():
    doc = {"lights": [{"type": "sun", "direction": [0, -1, 0], "intensity": 3, "angular_radius": 0.5,
                       "shadow_samples": 9},
                      {"type": "sphere", "position": [0, 3, 0], "radius": -1, "intensity": 5,
                       "visible_to_camera": False},
                      {"type": "unknown_kind"}],
           "world": {"ground": {"y": -1, "color": [0.2, 0.3, 0.4]},
                     "medium": {"type": "homogeneous", "sigma_a": [0.1, -1, 0.2], "sigma_s": [0.5, 0.5, 0.5]}},
           "materials": [{"id": "d", "type": "disney", "color": [0.9, 0.1, 0.1], "metallic": 1, "roughness": 0.2,
                          "transmission_color": [0.5, 0.6, 0.7], "coat_roughness": 0.3, "thin_walled": True,
                          "clearcoat": 0.25}],
           "entities": [{"type": "quad", "material": "d"}]}
    d = scene_loader.load_data(doc, 10, 10)
    g = d.mats[d.prims["material"][0]]
    assert d.prims["kind"][0] == abi.RT_PRIM_PLANE and np.allclose(d.prims["v"][0][:6], [0, -1, 0, 0, 1, 0])
    assert g["kind"] == abi.RT_MAT_DISNEY and np.allclose(g["color"], [0.2, 0.3, 0.4]) and g["roughness"] == np.float32(0.8)
    m = d.mats[d.prims["material"][1]]
    assert m["kind"] == abi.RT_MAT_DISNEY and m["metallic"] == 1.0 and m["roughness"] == np.float32(0.2)
    assert m["flags"] & abi.RT_MATF_TRANSMISSION_COLOR and m["flags"] & abi.RT_MATF_COAT_ROUGHNESS
    assert m["flags"] & abi.RT_MATF_THIN_WALLED and m["coat_roughness"] == np.float32(0.3)
    assert m["sheen_tint"] == np.float32(0.5) and m["clearcoat_gloss"] == 1.0   # MaterialData defaults
    assert d.medium == ((float(np.float32(0.1)), 0.0, float(np.float32(0.2))), (0.5, 0.5, 0.5))
    kinds = list(d.lights["kind"])
    # the unknown light is skipped; the default sky (no world.sky) is sampled
    # by the EnvironmentLight SceneLoader appends last
    assert kinds == [abi.RT_LIGHT_DIRECTIONAL, abi.RT_LIGHT_SPHERE, abi.RT_LIGHT_ENVIRONMENT]
    assert d.lights[0]["shadow_samples"] == 0                               # directional: auto
    assert d.lights[1]["radius"] == np.float32(0.5)                          # radius <= 0: 0.5
    assert d.prims["flags"][d.lights[1]["primitive"]] == abi.RT_PRIMF_CAMERA_INVISIBLE


@pytest.mark.parametrize("doc,what", [
    ({"lights": [], "entities": [{"type": "box"}]}, "type 'box'"),
    ({"lights": [], "materials": [{"id": "m", "type": "metal"}]}, "metal"),
    ({"lights": [], "materials": [{"id": "m", "type": "lambertian", "texture": {"type": "checker"}}]}, "texture"),
    ({"lights": [{"type": "point"}]}, "light type 'point'"),
    ({}, "default point light"),
    ({"lights": [], "world": {"sky": {"type": "hdri"}}}, "sky type"),
    ({"lights": [], "world": {"ground": {"type": "quad"}}}, "ground type"),
    ({"lights": [], "entities": [{"type": "sphere", "motion": [{"t": 1}]}]}, "motion"),
    ({"lights": [], "entities": [{"type": "group", "children": [{"type": "csg"}]}]}, "type 'csg'"),
    ({"lights": [], "materials": [{"id": "e", "type": "emissive"}],
      "entities": [{"type": "group", "children": [{"type": "sphere", "material": "e"}]}]}, "emissive objects"),
])
def test_unsupported_features_are_refused(doc, what):
    with pytest.raises(scene_loader.SceneLoadError, match=what):
        scene_loader.load_data(doc, 8, 8)


def test_rotated_surface_frames_are_refused_where_they_matter():
    """A sphere or plane baked into world space keeps the UV / tangent frame
    of its world-space shape, while the reference's Transform wrapper turns
    it with the entity: rotating one whose material reads that frame (image
    texture, anisotropic Disney) is refused; a plain material loads."""
    mats = [{"id": "plain", "type": "lambertian", "color": [0.5, 0.5, 0.5]},
            {"id": "aniso", "type": "disney", "color": [0.5, 0.5, 0.5], "anisotropic": 0.6}]

    def doc(typ, mat, rotate):
        e = {"type": typ, "material": mat, "rotate": rotate}
        return {"lights": [], "materials": mats, "entities": [e]}

    for typ in ("sphere", "plane"):
        scene_loader.load_data(doc(typ, "plain", [0, 30, 0]), 8, 8)
        scene_loader.load_data(doc(typ, "aniso", [0, 360, 0]), 8, 8)
        with pytest.raises(scene_loader.SceneLoadError, match=f"rotated {typ}"):
            scene_loader.load_data(doc(typ, "aniso", [0, 30, 0]), 8, 8)


@pytest.mark.parametrize("key,value,ok", [("uv_scale", [2, 2], False), ("uv_scale", [1, 1], True),
                                          ("uv_offset", [0.5, 0], False), ("uv_rotation", 15, False),
                                          ("uv_rotation", 0, True)])
def test_ground_uv_transform_is_refused(key, value, ok):
    doc = {"lights": [], "world": {"ground": {"y": 0, key: value}}}
    if ok:
        scene_loader.load_data(doc, 8, 8)
    else:
        with pytest.raises(scene_loader.SceneLoadError, match=key):
            scene_loader.load_data(doc, 8, 8)


def _tree_doc():
    return {
        "lights": [{"type": "sphere", "position": [0, 5, 0], "radius": 0.2, "intensity": 20}],
        "materials": [{"id": "bark", "type": "lambertian", "color": [0.4, 0.3, 0.2]},
                      {"id": "leaf", "type": "disney", "color": [0.1, 0.6, 0.1], "roughness": 0.5},
                      {"id": "gold", "type": "lambertian", "color": [0.9, 0.7, 0.1]}],
        "templates": [{"name": "tree", "material": "bark", "translate": [0, 0.5, 0],
                       "children": [{"type": "sphere", "radius": 0.5, "material": "leaf", "translate": [0, 1, 0]},
                                    {"type": "quad", "q": [-0.1, 0, 0], "u": [0.2, 0, 0], "v": [0, 1, 0]}]}],
        "entities": [{"type": "instance", "template": "tree", "translate": [-2, 0, 0]},
                     {"type": "instance", "template": "Tree", "material": "gold", "rotate": [0, 45, 0],
                      "scale": [1, 2, 1]},
                     {"type": "instance", "template": "missing"},
                     {"type": "group", "translate": [2, 0, 0], "material": "leaf",
                      "children": [{"type": "sphere", "radius": 0.2, "center": [0, 0, k * 0.5]} for k in range(6)]},
                     {"type": "sphere", "radius": 0.3, "material": "bark"}],
    }


def test_templates_instances_and_groups():
    warnings = []
    d = scene_loader.load_data(_tree_doc(), 64, 48, warnings=warnings)
    nodes, children, roots = d.hierarchy
    kind = nodes["kind"]
    assert any("template 'missing' not found" in w for w in warnings)
    # roots: two instances, the group, the sphere, the light's proxy (object order)
    assert len(roots) == 5
    t0, t1, g, sph, proxy = roots
    assert kind[t0] == kind[t1] == kind[g] == abi.RT_NODE_TRANSFORM
    i0, i1 = nodes["child"][t0], nodes["child"][t1]
    assert kind[i0] == kind[i1] == abi.RT_NODE_INSTANCE
    # one template build shared by both instances (case-insensitive name); the
    # second instance overrides the material
    assert nodes["child"][i0] == nodes["child"][i1]
    assert nodes["material"][i0] == -1 and nodes["material"][i1] == 2
    shared = nodes["child"][i0]
    assert kind[shared] == abi.RT_NODE_TRANSFORM                 # the template's default pose
    assert np.allclose(nodes["matrix"][shared].reshape(4, 4)[3, :3], [0, 0.5, 0])
    grp = nodes["child"][shared]
    assert kind[grp] == abi.RT_NODE_GROUP and nodes["count"][grp] == 2
    kids = children[nodes["first"][grp]:nodes["first"][grp] + 2]
    assert kind[kids[0]] == abi.RT_NODE_TRANSFORM and kind[kids[1]] == abi.RT_NODE_PRIMITIVES
    leaf_prim = nodes["first"][nodes["child"][kids[0]]]
    quad_prim = nodes["first"][kids[1]]
    assert d.prims["material"][leaf_prim] == 1 and d.prims["material"][quad_prim] == 0   # own / fallback
    # the group: 6 children (a BVH in the renderer), the group's material as fallback
    gg = nodes["child"][g]
    assert kind[gg] == abi.RT_NODE_GROUP and nodes["count"][gg] == 6
    assert np.all(d.prims["material"][nodes["first"][children[nodes["first"][gg]:nodes["first"][gg] + 6]]] == 1)
    assert kind[sph] == abi.RT_NODE_PRIMITIVES and kind[proxy] == abi.RT_NODE_PRIMITIVES
    # the instanced template's primitives exist once
    assert len(d.prims) == 2 + 6 + 1 + 1


def test_imports_merge_before_local(tmp_path):
    lib = tmp_path / "lib.yaml"
    lib.write_text("materials:\n  - {id: red, type: lambertian, color: [1, 0, 0]}\n"
                   "  - {id: blue, type: lambertian, color: [0, 0, 1]}\n"
                   "entities:\n  - {type: sphere, material: red, radius: 0.25}\n"
                   "templates:\n  - name: pebble\n    children: [{type: sphere, radius: 0.1, material: blue}]\n")
    main = tmp_path / "main.yaml"
    main.write_text("imports: [{path: lib.yaml}, {path: lib.yaml}, {path: nowhere.yaml}]\n"
                    "lights: []\n"
                    "materials:\n  - {id: red, type: lambertian, color: [0, 1, 0]}\n"
                    "entities:\n  - {type: sphere, material: red, radius: 0.5, center: [1, 0, 0]}\n"
                    "  - {type: instance, template: pebble, translate: [0, 1, 0]}\n")
    warnings = []
    d = scene_loader.load(str(main), 32, 32, warnings=warnings)
    assert any("Circular import" in w for w in warnings) and any("not found" in w for w in warnings)
    # imported entity first, then the local ones; the local 'red' wins for both
    assert np.allclose(d.prims["v"][:2, 3], [0.25, 0.5])
    reds = d.mats["color"][d.prims["material"][:2]]
    assert np.allclose(reds, [[0, 1, 0], [0, 1, 0]])
    nodes, _, roots = d.hierarchy
    assert len(roots) == 3 and nodes["kind"][roots[2]] == abi.RT_NODE_TRANSFORM
