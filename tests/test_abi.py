"""The C ABI library: loads, exports every entry point include/rt3d.h
declares, keeps the struct layout, fills the reference defaults and rejects
bad input with status codes (no compute calls; no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np

from helpers import random_spheres
from raytracer import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "rt3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    fns = header_functions()
    for f in ("rt_create", "rt_scene_create", "rt_render", "rt_trace_primary", "rt_trace_rays", "rt_last_error"):
        assert f in fns


def test_library_exports_every_declared_symbol(native):
    L = native.load()
    for name in header_functions():
        assert hasattr(L, name), name
    assert set(native.EXPORTED_SYMBOLS) == set(header_functions())


def test_struct_sizes():
    assert C.sizeof(abi.rt_primitive) == 64
    assert C.sizeof(abi.rt_material) == 128
    assert C.sizeof(abi.rt_light) == 80
    assert C.sizeof(abi.rt_settings) == 96
    # the same sizes are static_asserts in the library (scene_build.hip)
    sizes = {"rt_camera": 56, "rt_medium": 32, "rt_scene_desc": 248, "rt_node": 160, "rt_image": 24, "rt_film": 88, "rt_render_stats": 152,
             "rt_scene_info": 96, "rt_batch_job": 32, "rt_exr_channel": 24, "rt_denoise_options": 32}
    for name, size in sizes.items():
        assert C.sizeof(getattr(abi, name)) == size, name


def test_version_and_defaults(native):
    L = native.load()
    assert L.rt_get_version() == 1
    s = abi.rt_settings()
    L.rt_settings_defaults(C.byref(s))
    assert (s.width, s.height, s.samples_per_pixel, s.max_depth) == (960, 540, 16, 8)
    assert s.max_sample_radiance == 10.0 and abs(s.indirect_clamp_factor - 0.25) < 1e-7
    assert s.rr_min_bounces == -1 and s.sampler == abi.RT_SAMPLER_SOBOL
    m = abi.rt_material()
    L.rt_material_defaults(C.byref(m), abi.RT_MAT_DISNEY)
    ref = abi.material_defaults(abi.RT_MAT_DISNEY)
    for f, _ in abi.rt_material._fields_:
        if f in ("color", "reserved"):
            continue
        a, b = getattr(m, f), getattr(ref, f)
        if isinstance(a, (int, float)):
            assert abs(a - b) < 1e-6, f


def test_build_rejects_bad_tables(native):
    d = random_spheres(8, 1)
    assert native.bvh_dump(d).size > 0
    bad = scenes.SceneDesc(d.prims.copy(), d.mats, d.lights, d.camera)
    bad.prims["material"][3] = 99
    L = native.load()
    assert L.rt_debug_bvh_dump(bad.ptr, None, 0) < 0
    bad2 = scenes.SceneDesc(d.prims.copy(), d.mats, d.lights, d.camera)
    bad2.prims["kind"][0] = 17
    assert L.rt_debug_bvh_dump(bad2.ptr, None, 0) < 0
    # a flat sky is accepted (with or without its environment light)
    sky = scenes.SceneDesc(d.prims, d.mats, d.lights, d.camera, sky=(0.5, 0.5, 0.5))
    assert L.rt_debug_bvh_dump(sky.ptr, None, 0) > 0
    badl = np.zeros(1, scenes.LIGHT_DTYPE)
    badl["kind"] = 9
    badl["shadow_samples"] = 1
    bad3 = scenes.SceneDesc(d.prims, d.mats, badl, d.camera)
    assert L.rt_debug_bvh_dump(bad3.ptr, None, 0) < 0
    assert L.rt_debug_bvh_dump(None, None, 0) < 0


def test_flat_world_has_no_bvh(native):
    # SceneLoader builds a BVH only above 4 objects (SceneLoader.cs:25,466)
    assert native.bvh_dump(random_spheres(4, 3)).size == 0
    assert native.bvh_dump(random_spheres(5, 3)).size > 0


def test_null_arguments_are_rejected(native):
    L = native.load()
    assert L.rt_create(0, None) == abi.RT_ERR_INVALID_ARGUMENT
    assert L.rt_render(None, None, None, None, None) == abi.RT_ERR_INVALID_ARGUMENT
    assert L.rt_trace_rays(None, None, None, 0, None, None, None, None) == abi.RT_ERR_INVALID_ARGUMENT
    assert L.rt_last_error(None) == b"null context"
