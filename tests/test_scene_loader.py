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

    tri = d.prims["v"][0][:9].reshape(3, 3)
    for k, p in enumerate(([1, 0, 0], [0, 1, 0], [0, 0, 1])):
        assert np.allclose(tri[k], ref(p), atol=1e-6)
    assert np.allclose(d.prims["v"][1][:4], [3, 0, -1, 1.5])
    q = d.prims["v"][2][:9].reshape(3, 3)       # rotate Y 90: u = +x -> -z
    assert np.allclose(q, [[0, 0, 0], [0, 0, -1], [0, 1, 0]], atol=1e-6)
    with pytest.raises(scene_loader.SceneLoadError, match="non-uniform"):
        scene_loader.load_data({"lights": [], "entities": [{"type": "sphere", "scale": [1, 2, 1]}]}, 8, 8)


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
    ({"lights": [], "imports": ["x.yaml"]}, "imports"),
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
