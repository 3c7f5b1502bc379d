"""ctypes mirror of include/rt3d.h (the C ABI of the HIP path tracer).

Every struct here has the same field order and size as its C counterpart;
sizes are asserted at import so a layout drift fails loudly.
"""
import ctypes as C

RT_OK = 0
RT_ERR_INVALID_ARGUMENT = -1
RT_ERR_NO_DEVICE = -2
RT_ERR_DEVICE = -3
RT_ERR_OUT_OF_MEMORY = -4
RT_ERR_UNSUPPORTED = -5
RT_ERR_INTERNAL = -6

RT_PRIM_TRIANGLE, RT_PRIM_SPHERE, RT_PRIM_QUAD, RT_PRIM_PLANE = 0, 1, 2, 3
RT_PRIMF_BACKFACE_CULLED, RT_PRIMF_CAMERA_INVISIBLE = 1, 2
RT_MAT_LAMBERTIAN, RT_MAT_EMISSIVE, RT_MAT_DIELECTRIC, RT_MAT_DISNEY = 0, 1, 2, 3
RT_MATF_LIGHT_PROXY, RT_MATF_THIN_WALLED, RT_MATF_TRANSMISSION_COLOR, RT_MATF_COAT_ROUGHNESS = 1, 2, 4, 8
RT_MATF_TEXTURED = 16
TEX_CHANNELS = 5   # texture set planes: base colour r, g, b, roughness, metallic
RT_LIGHT_SPHERE, RT_LIGHT_AREA, RT_LIGHT_DIRECTIONAL, RT_LIGHT_GEOMETRY, RT_LIGHT_ENVIRONMENT = 0, 1, 2, 3, 4
RT_SAMPLER_SOBOL, RT_SAMPLER_PRNG = 0, 1
RT_MIS_BALANCE, RT_MIS_POWER = 0, 1
RT_LIGHTS_ALL, RT_LIGHTS_POWER, RT_LIGHTS_UNIFORM = 0, 1, 2

F3 = C.c_float * 3


class rt_primitive(C.Structure):
    _fields_ = [("kind", C.c_int32), ("material", C.c_int32), ("flags", C.c_int32),
                ("mesh", C.c_int32), ("v", C.c_float * 12)]


class rt_material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flags", C.c_int32), ("color", F3), ("intensity", C.c_float),
                ("metallic", C.c_float), ("roughness", C.c_float), ("specular", C.c_float),
                ("specular_tint", C.c_float), ("sheen", C.c_float), ("sheen_tint", C.c_float),
                ("sheen_roughness", C.c_float), ("clearcoat", C.c_float), ("clearcoat_gloss", C.c_float),
                ("spec_trans", C.c_float), ("ior", C.c_float), ("anisotropic", C.c_float),
                ("anisotropic_rotation", C.c_float), ("diff_trans", C.c_float), ("coat_ior", C.c_float),
                ("coat_roughness", C.c_float), ("thin_film_thickness", C.c_float),
                ("thin_film_ior", C.c_float), ("transmission_depth", C.c_float),
                ("transmission_color", F3), ("image", C.c_int32), ("reserved", C.c_int32 * 3)]


class rt_light(C.Structure):
    _fields_ = [("kind", C.c_int32), ("shadow_samples", C.c_int32), ("primitive", C.c_int32),
                ("material", C.c_int32), ("color", F3), ("intensity", C.c_float), ("position", F3),
                ("radius", C.c_float), ("u", F3), ("soft_radius", C.c_float), ("v", F3),
                ("reserved", C.c_float)]


class rt_camera(C.Structure):
    _fields_ = [("look_from", F3), ("look_at", F3), ("vup", F3), ("vfov_deg", C.c_float),
                ("aspect", C.c_float), ("aperture", C.c_float), ("focus_dist", C.c_float),
                ("reserved", C.c_float)]


class rt_medium(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("sigma_a", F3), ("sigma_s", F3), ("reserved", C.c_float)]


class rt_scene_desc(C.Structure):
    _fields_ = [("primitives", C.POINTER(rt_primitive)), ("num_primitives", C.c_int64),
                ("materials", C.POINTER(rt_material)), ("num_materials", C.c_int64),
                ("lights", C.POINTER(rt_light)), ("num_lights", C.c_int64),
                ("camera", rt_camera), ("sky_color", F3), ("reserved0", C.c_int32),
                ("global_medium", rt_medium), ("texcoords", C.POINTER(C.c_float)),
                ("texels", C.POINTER(C.c_float)), ("tex_width", C.c_int32), ("tex_height", C.c_int32),
                ("normals", C.POINTER(C.c_float)), ("images", C.c_void_p), ("num_images", C.c_int64),
                ("nodes", C.c_void_p), ("num_nodes", C.c_int64), ("node_children", C.POINTER(C.c_int32)),
                ("num_node_children", C.c_int64), ("roots", C.POINTER(C.c_int32)), ("num_roots", C.c_int64)]


RT_NODE_PRIMITIVES, RT_NODE_GROUP, RT_NODE_TRANSFORM, RT_NODE_INSTANCE = 0, 1, 2, 3
RT_NODEF_CAMERA_INVISIBLE = 1


class rt_node(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flags", C.c_int32), ("first", C.c_int32), ("count", C.c_int32),
                ("child", C.c_int32), ("material", C.c_int32), ("reserved", C.c_int32 * 2),
                ("matrix", C.c_float * 16), ("inverse", C.c_float * 16)]


class rt_image(C.Structure):
    _fields_ = [("rgba8", C.POINTER(C.c_uint8)), ("width", C.c_int32), ("height", C.c_int32),
                ("scale_u", C.c_float), ("scale_v", C.c_float)]


RT_MATF_IMAGE = 32
TEXTURE_FILTERING = {"auto": 0, "on": 1, "off": 2}


class rt_settings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("samples_per_pixel", C.c_int32),
                ("max_depth", C.c_int32), ("sampler", C.c_int32), ("mis", C.c_int32),
                ("light_sampling", C.c_int32), ("capture_aovs", C.c_int32),
                ("max_sample_radiance", C.c_float), ("indirect_clamp_factor", C.c_float),
                ("exposure_ev", C.c_float), ("rr_min_bounces", C.c_int32),
                ("rr_min_survival", C.c_float), ("seed_xor", C.c_uint32), ("caustics", C.c_int32),
                ("caustic_photons", C.c_int32), ("caustic_gather_radius", C.c_float),
                ("paths_in_flight", C.c_int32), ("count_traversal", C.c_int32),
                ("texture_filtering", C.c_int32), ("sample_offset", C.c_int32),
                ("sample_total", C.c_int32), ("reserved", C.c_int32 * 2)]


FP = C.POINTER(C.c_float)


class rt_film(C.Structure):
    _fields_ = [("beauty", FP), ("beauty_a", FP), ("beauty_b", FP), ("albedo_a", FP), ("albedo_b", FP),
                ("normal_a", FP), ("normal_b", FP), ("depth_a", FP), ("depth_b", FP),
                ("sample_count", C.POINTER(C.c_int32)), ("display", FP)]


class rt_render_stats(C.Structure):
    _fields_ = [("wall_ms", C.c_double), ("device_ms", C.c_double), ("trace_ms", C.c_double),
                ("shadow_ms", C.c_double), ("shade_ms", C.c_double), ("camera_samples", C.c_int64),
                ("rays_closest", C.c_int64), ("rays_shadow", C.c_int64), ("trace_launches", C.c_int64),
                ("shadow_launches", C.c_int64), ("trace_node_tests", C.c_int64), ("trace_tri_tests", C.c_int64),
                ("trace_other_tests", C.c_int64), ("shadow_node_tests", C.c_int64), ("shadow_tri_tests", C.c_int64),
                ("shadow_other_tests", C.c_int64), ("any_node_fetches", C.c_int64), ("any_tri_tests", C.c_int64),
                ("any_other_tests", C.c_int64)]


class rt_batch_job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("settings", C.POINTER(rt_settings)), ("film", C.POINTER(rt_film)),
                ("stats", C.POINTER(rt_render_stats))]


class rt_scene_info(C.Structure):
    _fields_ = [("num_primitives", C.c_int64), ("num_objects", C.c_int64), ("num_nodes", C.c_int64),
                ("num_leaves", C.c_int64), ("max_depth", C.c_int32), ("has_bvh", C.c_int32),
                ("rr_min_bounces", C.c_int32), ("rr_min_survival", C.c_float),
                ("scene_radius", C.c_float), ("mean_irradiance", C.c_float),
                ("device_bytes", C.c_int64), ("reserved", C.c_int64 * 4)]


RT_EXR_HALF, RT_EXR_FLOAT = 1, 2


class rt_exr_channel(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int32), ("pad", C.c_int32), ("plane", C.POINTER(C.c_float))]


RT_DENOISE_NLM, RT_DENOISE_NFOR = 1, 2
RT_DENOISE_FAST, RT_DENOISE_HIGH = 0, 1


class rt_denoise_options(C.Structure):
    _fields_ = [("kind", C.c_int32), ("quality", C.c_int32), ("selection_margin", C.c_float),
                ("samples_a", C.c_int32), ("samples_b", C.c_int32), ("reserved", C.c_int32 * 3)]


for _t, _n in ((rt_primitive, 64), (rt_material, 128), (rt_light, 80), (rt_camera, 56),
               (rt_medium, 32), (rt_settings, 96), (rt_node, 160)):
    assert C.sizeof(_t) == _n, (_t.__name__, C.sizeof(_t))


def material_defaults(kind):
    """DisneyBsdf constructor defaults (DisneyBsdf.cs:167-212) for any kind."""
    m = rt_material()
    m.kind = kind
    m.color[:] = (0.8, 0.8, 0.8)
    m.intensity = 1.0
    m.metallic, m.roughness, m.specular, m.specular_tint = 0.0, 0.5, 0.5, 0.0
    m.sheen, m.sheen_tint, m.sheen_roughness, m.clearcoat = 0.0, 0.5, 0.3, 0.0
    m.clearcoat_gloss, m.spec_trans, m.ior, m.anisotropic = 1.0, 0.0, 1.5, 0.0
    m.anisotropic_rotation, m.diff_trans, m.coat_ior, m.coat_roughness = 0.0, 0.0, 1.5, 0.0
    m.thin_film_thickness, m.thin_film_ior, m.transmission_depth = 0.0, 1.5, 0.0
    m.transmission_color[:] = (1.0, 1.0, 1.0)
    return m


def settings_defaults(width=256, height=256, spp=16, max_depth=8):
    """Canonical parity settings: --sampler sobol --mis balance
    --light-sampling all -C 10 --indirect-clamp-factor 0.25 --exposure 0,
    scene-classified Russian roulette (Renderer.cs:300-378)."""
    s = rt_settings()
    s.width, s.height, s.samples_per_pixel, s.max_depth = width, height, spp, max_depth
    s.sampler, s.mis, s.light_sampling, s.capture_aovs = RT_SAMPLER_SOBOL, RT_MIS_BALANCE, RT_LIGHTS_ALL, 0
    s.max_sample_radiance, s.indirect_clamp_factor, s.exposure_ev = 10.0, 0.25, 0.0
    s.rr_min_bounces, s.rr_min_survival = -1, -1.0
    s.seed_xor = 0
    return s
