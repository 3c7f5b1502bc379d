"""Scene assembly (SceneLoader order, geometry-light extraction) and the
Renderer constructor's scene analysis, checked on the CPU against closed
forms."""
import math

import numpy as np

from raytracer import abi, scenes


def test_cornell_classifier_matches_closed_form(oracle):
    d = scenes.cornell()
    info = oracle.OracleScene(d).info()
    # Ē = π·A·L / (4π R²): A = 0.0625, L = 15, R = |extent|/2 of the box
    ext = np.array([2.0, 2.0, 2.0]) + 2e-4
    R = 0.5 * np.linalg.norm(ext)
    e = math.pi * 0.0625 * 15.0 / (4 * math.pi * R * R)
    assert abs(info.mean_irradiance - e) < 1e-4
    assert info.rr_min_bounces == 6 and abs(info.rr_min_survival - 0.40) < 1e-7


def test_geometry_lights_follow_object_order():
    b = scenes.SceneBuilder()
    e1 = b.emissive((1, 1, 1), 2.0)
    e2 = b.emissive((1, 0, 0), 3.0)
    m = b.lambertian((0.5, 0.5, 0.5))
    b.sphere((0, 0, 0), 1, m)
    q = b.quad((0, 2, 0), (1, 0, 0), (0, 0, 1), e1)
    b.sphere_light((0, 5, 0), 0.5, intensity=7.0)
    t = b.triangle(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32), np.array([0, 1, 0], np.float32), e2)
    d = b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40, 1))
    kinds = d.lights["kind"].tolist()
    # YAML lights first (with their proxy appended to the objects), then
    # geometry lights in object order
    assert kinds == [abi.RT_LIGHT_SPHERE, abi.RT_LIGHT_GEOMETRY, abi.RT_LIGHT_GEOMETRY]
    assert d.lights["primitive"][1] == q and d.lights["primitive"][2] == t
    proxy = d.lights["primitive"][0]
    assert d.prims["kind"][proxy] == abi.RT_PRIM_SPHERE
    assert d.mats["flags"][d.prims["material"][proxy]] & abi.RT_MATF_LIGHT_PROXY
    assert list(d.lights["shadow_samples"]) == [4, 4, 4]


def test_area_light_proxy_is_backface_culled():
    b = scenes.SceneBuilder()
    b.area_light((0, 2, 0), (1, 0, 0), (0, 0, 1), intensity=5.0)
    d = b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40, 1))
    assert d.prims["flags"][0] & abi.RT_PRIMF_BACKFACE_CULLED


def test_icosphere_counts():
    v, f = scenes.icosphere(3)
    assert len(f) == 20 * 4 ** 3 and len(v) == 10 * 4 ** 3 + 2
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)


def test_displaced_icosphere_workload_shape():
    d = scenes.displaced_icosphere_scene(64, 36, subdiv=4)
    n = 20 * 4 ** 4
    assert (d.prims["kind"] == abi.RT_PRIM_TRIANGLE).sum() == n + 2
    disney = d.mats[d.prims["material"][2:2 + n]]
    assert np.all(disney["kind"] == abi.RT_MAT_DISNEY)
    assert np.all((disney["roughness"] >= 0.05) & (disney["roughness"] <= 1.0))
    assert set(np.unique(disney["metallic"])) <= {0.0, 1.0}
    assert list(d.lights["kind"]) == [abi.RT_LIGHT_SPHERE] * 4 + [abi.RT_LIGHT_DIRECTIONAL]
