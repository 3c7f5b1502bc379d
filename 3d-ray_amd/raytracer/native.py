"""ctypes binding of lib3dray_amd.so, shaped like the reference's render
entry point:

  reference (C#)                                    here
  new Renderer(world, camera, lights, sky, spp,     Renderer(scene, samples_per_pixel, max_depth,
               maxDepth, globalMedium,                        max_sample_radiance=..., ...)
               maxSampleRadiance, verbose,
               misHeuristic, lightSamplingStrategy,
               indirectClampFactor, ..., exposureEv)
      -- src/RayTracer/Rendering/Renderer.cs:300-319
  renderer.Render(width, height, options)           renderer.render(width, height, capture_aovs=...)
      -- Renderer.cs:522-733, returns RenderResult      returns RenderResult(pixels, buffers)
         {Pixels, Buffers} (RenderBuffers.cs)

The world/camera/lights/sky/medium travel together as a scenes.SceneDesc
(the flattened SceneLoader output).  There is no CPU fallback: a missing
library or device raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RT3D_LIB") or os.path.join(os.path.dirname(_HERE), "lib3dray_amd.so")
_lib = None


class RtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rt error {code}: {msg}")
        self.code = code


def load():
    """Loads lib3dray_amd.so (fails loudly when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (or python 3d-ray_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int32)
    L.rt_get_version.restype = C.c_int
    L.rt_material_defaults.argtypes = [C.POINTER(abi.rt_material), C.c_int32]
    L.rt_material_defaults.restype = None
    L.rt_settings_defaults.argtypes = [C.POINTER(abi.rt_settings)]
    L.rt_settings_defaults.restype = None
    L.rt_create.argtypes = [C.c_int32, C.POINTER(vp)]
    L.rt_destroy.argtypes = [vp]
    L.rt_destroy.restype = None
    L.rt_last_error.argtypes = [vp]
    L.rt_last_error.restype = C.c_char_p
    L.rt_scene_create.argtypes = [vp, C.POINTER(abi.rt_scene_desc), C.POINTER(vp)]
    L.rt_scene_destroy.argtypes = [vp]
    L.rt_scene_destroy.restype = None
    L.rt_scene_get_info.argtypes = [vp, C.POINTER(abi.rt_scene_info)]
    L.rt_scene_set_textures.argtypes = [vp, vp, C.c_int32, C.c_int32, fp]
    L.rt_render.argtypes = [vp, vp, C.POINTER(abi.rt_settings), C.POINTER(abi.rt_film), C.POINTER(abi.rt_render_stats)]
    L.rt_trace_primary.argtypes = [vp, vp, C.POINTER(abi.rt_settings), C.c_int32, C.c_int32, ip, fp, fp, fp]
    L.rt_trace_rays.argtypes = [vp, vp, fp, C.c_int64, ip, fp, fp, fp]
    L.rt_occluded.argtypes = [vp, vp, fp, C.c_int64, ip]
    L.rt_occluded_noalloc.argtypes = [vp, vp, fp, C.c_int64, ip]
    i64p = C.POINTER(C.c_int64)
    L.rt_photon_map.argtypes = [vp, vp, C.POINTER(abi.rt_settings), C.c_int64, C.c_int64, i64p, i64p, i64p, fp, fp, fp,
                                ip, ip, ip, fp]
    L.rt_photon_gather.argtypes = [vp, vp, C.POINTER(abi.rt_settings), fp, C.c_int64, ip, fp, ip]
    L.rt_render_batch.argtypes = [vp, C.POINTER(abi.rt_batch_job), C.c_int32, C.c_int32, C.POINTER(abi.rt_render_stats)]
    L.rt_context_tune.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
    L.rt_adjoint.argtypes = [vp, vp, C.POINTER(abi.rt_settings), fp, fp, C.POINTER(abi.rt_render_stats)]
    L.rt_debug_bvh_dump.argtypes = [C.POINTER(abi.rt_scene_desc), ip, C.c_int64]
    L.rt_debug_bvh_dump.restype = C.c_int64
    u8p = C.POINTER(C.c_uint8)
    L.rt_exr_encode.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(abi.rt_exr_channel), C.c_int32, u8p, C.c_int64,
                                i64p]
    L.rt_pfm_encode.argtypes = [vp, fp, C.c_int32, C.c_int32, C.c_int32, u8p, C.c_int64, i64p]
    L.rt_png_encode.argtypes = [vp, fp, C.c_int32, C.c_int32, u8p, C.c_int64, i64p]
    L.rt_denoise.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(abi.rt_film), C.POINTER(abi.rt_denoise_options),
                             fp]
    _lib = L
    return L


GATHER_K = 40   # Renderer.CausticGatherK


EXPORTED_SYMBOLS = ("rt_get_version", "rt_material_defaults", "rt_settings_defaults", "rt_create", "rt_destroy",
                    "rt_last_error", "rt_scene_create", "rt_scene_destroy", "rt_scene_get_info",
                    "rt_scene_set_textures", "rt_render",
                    "rt_trace_primary", "rt_trace_rays", "rt_photon_map", "rt_photon_gather", "rt_render_batch",
                    "rt_context_tune",
                    "rt_adjoint", "rt_occluded", "rt_occluded_noalloc", "rt_exr_encode", "rt_pfm_encode",
                    "rt_png_encode",
                    "rt_denoise",
                    "rt_debug_bvh_dump")


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


def bvh_dump(desc):
    """Host-side build of `desc`; the flattened world BVH in pre-order
    (rt_debug_bvh_dump).  Needs no GPU."""
    L = load()
    n = L.rt_debug_bvh_dump(desc.ptr, None, 0)
    if n < 0:
        raise RtError(int(n), "scene build failed")
    out = np.empty(n, np.int32)
    L.rt_debug_bvh_dump(desc.ptr, _i(out), n)
    return out


class Context:
    """One HIP device + stream (rt_create)."""

    def __init__(self, device=0):
        L = load()
        h = C.c_void_p()
        rc = L.rt_create(device, C.byref(h))
        if rc != 0:
            raise RtError(rc, "rt_create failed (no HIP device?)")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise RtError(rc, load().rt_last_error(self.h).decode())

    def tune(self, shade_waves_per_simd=0, traversal_stack=0, bounce_graph=0):
        """rt_context_tune: kernel configuration (0 = keep); stack 1 private,
        2 LDS; bounce_graph 1 direct launches, 2 hipGraph replay."""
        self.check(load().rt_context_tune(self.h, shade_waves_per_simd, traversal_stack, bounce_graph))

    def close(self):
        if getattr(self, "h", None):
            load().rt_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Scene:
    """Device-resident scene (rt_scene_create): BVH, tables, LUT in HBM."""

    def __init__(self, ctx, desc):
        self.ctx = ctx
        self.desc = desc
        h = C.c_void_p()
        ctx.check(load().rt_scene_create(ctx.h, desc.ptr, C.byref(h)))
        self.h = h

    def info(self):
        i = abi.rt_scene_info()
        self.ctx.check(load().rt_scene_get_info(self.h, C.byref(i)))
        return i

    def set_textures(self, texture):
        """Replaces the texture set ((5, h, w) fp32: base r, g, b, roughness,
        metallic) of a scene with textured materials."""
        t = np.ascontiguousarray(texture, np.float32)
        assert t.ndim == 3 and t.shape[0] == abi.TEX_CHANNELS
        self.ctx.check(load().rt_scene_set_textures(self.ctx.h, self.h, t.shape[2], t.shape[1], _f(t)))
        self.desc.texture = t

    def trace_primary(self, settings, sample_begin=0, sample_count=1):
        n = settings.width * settings.height * sample_count
        prim = np.empty(n, np.int32)
        t, u, v = (np.empty(n, np.float32) for _ in range(3))
        self.ctx.check(load().rt_trace_primary(self.ctx.h, self.h, C.byref(settings), sample_begin, sample_count,
                                               _i(prim), _f(t), _f(u), _f(v)))
        return prim, t, u, v

    def trace_rays(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = len(rays)
        prim = np.empty(n, np.int32)
        t, u, v = (np.empty(n, np.float32) for _ in range(3))
        self.ctx.check(load().rt_trace_rays(self.ctx.h, self.h, _f(rays), n, _i(prim), _f(t), _f(u), _f(v)))
        return prim, t, u, v

    def occluded(self, segs):
        """Any-hit queries (rt_occluded): segs (n, 8) = origin, direction,
        tmin, tmax -> bool (n,)."""
        segs = np.ascontiguousarray(segs, np.float32).reshape(-1, 8)
        out = np.zeros(len(segs), np.int32)
        self.ctx.check(load().rt_occluded(self.ctx.h, self.h, _f(segs), len(segs), _i(out)))
        return out.astype(bool)

    def photon_map(self, settings):
        """Device photon map of the caustic pre-pass (rt_photon_map): dict of
        pos/inc/power (n,3), key/idx (n,), start (n_buckets+1,), grid (4,)."""
        L = load()
        n, em, nb = (C.c_int64(0) for _ in range(3))
        self.ctx.check(L.rt_photon_map(self.ctx.h, self.h, C.byref(settings), 0, 0, C.byref(n), C.byref(em), C.byref(nb),
                                       None, None, None, None, None, None, None))
        n, em, nb = n.value, em.value, nb.value
        out = {"n": n, "emitted": em, "n_buckets": nb}
        pos, inc, pw = (np.zeros((n, 3), np.float32) for _ in range(3))
        key, idx = np.zeros(n, np.int32), np.zeros(n, np.int32)
        start = np.zeros(nb + 1, np.int32)
        grid = np.zeros(4, np.float32)
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        self.ctx.check(L.rt_photon_map(self.ctx.h, self.h, C.byref(settings), n, nb + 1, C.byref(a), C.byref(b),
                                       C.byref(c), _f(pos), _f(inc), _f(pw), _i(key), _i(idx), _i(start), _f(grid)))
        out.update(pos=pos, inc=inc, power=pw, key=key, idx=idx, start=start, grid=grid)
        return out

    def photon_gather(self, settings, points):
        """k-nearest photon search of the render on the device
        (rt_photon_gather): count (n,), radius (n,), idx (n, 40) photon-order
        indices in heap order, -1 past the count."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = len(pts)
        cnt, rad = np.zeros(n, np.int32), np.zeros(n, np.float32)
        idx = np.full((n, GATHER_K), -1, np.int32)
        self.ctx.check(load().rt_photon_gather(self.ctx.h, self.h, C.byref(settings), _f(pts), n, _i(cnt), _f(rad),
                                               _i(idx)))
        return cnt, rad, idx

    def close(self):
        if getattr(self, "h", None):
            load().rt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class RenderBuffers:
    """Linear-HDR film (RenderBuffers.cs:52-169), planes shaped (C, H, W)."""

    def __init__(self, planes, spp, first=0):
        self.__dict__.update(planes)
        # even global sample indices in [first, first + spp) form half A
        self.samples_a = spp // 2 if first % 2 else (spp + 1) // 2
        self.samples_b = spp - self.samples_a


class RenderResult:
    def __init__(self, pixels, buffers, stats):
        self.pixels = pixels      # tone-mapped display (H, W, 3)
        self.buffers = buffers
        self.stats = stats


_MIS = {"balance": abi.RT_MIS_BALANCE, "power": abi.RT_MIS_POWER}
_LIGHTS = {"all": abi.RT_LIGHTS_ALL, "power": abi.RT_LIGHTS_POWER, "uniform": abi.RT_LIGHTS_UNIFORM}


_SAMPLERS = {"sobol": abi.RT_SAMPLER_SOBOL, "prng": abi.RT_SAMPLER_PRNG}


def render_spp(settings):
    """Camera samples per pixel a render takes: spp with the Sobol sampler,
    ceil(sqrt(spp))^2 strata with the PRNG sampler (Renderer.cs:556-558),
    except in a chunk (sample_offset / sample_total set), which takes exactly
    its spp samples of the whole render's grid."""
    chunk = settings.sample_offset != 0 or settings.sample_total != 0
    if settings.sampler == abi.RT_SAMPLER_PRNG and not chunk:
        q = int(np.ceil(np.sqrt(np.float32(settings.samples_per_pixel))))
        return q * q
    return settings.samples_per_pixel


class Renderer:
    """Mirror of Renderer(...) + Render(width, height, options)."""

    def __init__(self, scene, samples_per_pixel, max_depth, max_sample_radiance=10.0, mis_heuristic="balance",
                 light_sampling_strategy="all", indirect_clamp_factor=0.25, exposure_ev=0.0, rr_min_bounces=-1,
                 rr_min_survival=-1.0, seed_xor=0, paths_in_flight=0, enable_caustics=False, caustic_photons=0,
                 caustic_gather_radius=0.0, sample_offset=0, sample_total=0, sampler="sobol"):
        self.scene = scene
        s = abi.rt_settings()
        load().rt_settings_defaults(C.byref(s))
        s.samples_per_pixel = samples_per_pixel
        s.max_depth = max_depth
        s.max_sample_radiance = max_sample_radiance
        s.mis = _MIS[mis_heuristic]
        s.light_sampling = _LIGHTS[light_sampling_strategy]
        s.indirect_clamp_factor = indirect_clamp_factor
        s.exposure_ev = exposure_ev
        s.rr_min_bounces = rr_min_bounces
        s.rr_min_survival = rr_min_survival
        s.seed_xor = seed_xor
        s.paths_in_flight = paths_in_flight
        s.caustics = 1 if enable_caustics else 0
        s.caustic_photons = caustic_photons
        s.caustic_gather_radius = caustic_gather_radius
        s.sampler = _SAMPLERS[sampler]
        s.sample_offset = sample_offset     # resume / chunking: samples [offset, offset + spp)
        s.sample_total = sample_total
        self.settings = s

    def _film(self, width, height, capture_aovs):
        npx = width * height
        planes = {k: np.zeros(3 * npx, np.float32) for k in ("beauty", "beauty_a", "beauty_b", "display")}
        if capture_aovs:
            for k in ("albedo_a", "albedo_b", "normal_a", "normal_b"):
                planes[k] = np.zeros(3 * npx, np.float32)
            planes["depth_a"] = np.zeros(npx, np.float32)
            planes["depth_b"] = np.zeros(npx, np.float32)
        count = np.zeros(npx, np.int32)
        film = abi.rt_film()
        for k, a in planes.items():
            setattr(film, k, _f(a))
        film.sample_count = _i(count)
        return film, planes, count

    def _result(self, width, height, planes, count, stats):
        npx = width * height
        shaped = {k: (a.reshape(3, height, width) if a.size == 3 * npx else a.reshape(height, width))
                  for k, a in planes.items()}
        shaped["sample_count"] = count.reshape(height, width)
        pixels = np.transpose(shaped["display"], (1, 2, 0))
        return RenderResult(pixels, RenderBuffers(shaped, render_spp(self.settings), self.settings.sample_offset), stats)

    def render(self, width, height, capture_aovs=False):
        s = self.settings
        s.width, s.height, s.capture_aovs = width, height, 1 if capture_aovs else 0
        film, planes, count = self._film(width, height, capture_aovs)
        stats = abi.rt_render_stats()
        ctx = self.scene.ctx
        ctx.check(load().rt_render(ctx.h, self.scene.h, C.byref(s), C.byref(film), C.byref(stats)))
        return self._result(width, height, planes, count, stats)

    def adjoint(self, width, height, d_beauty):
        """Texture gradient (rt_adjoint): dLoss/dTexels (5, th, tw) of a loss
        with dLoss/dBeauty = d_beauty (3, height, width); returns (gradient,
        rt_render_stats)."""
        s = self.settings
        s.width, s.height, s.capture_aovs = width, height, 0
        db = np.ascontiguousarray(d_beauty, np.float32).reshape(3, height, width)
        tex = self.scene.desc.texture
        out = np.zeros(tex.shape, np.float32)
        stats = abi.rt_render_stats()
        ctx = self.scene.ctx
        ctx.check(load().rt_adjoint(ctx.h, self.scene.h, C.byref(s), _f(db), _f(out), C.byref(stats)))
        return out, stats


def render_batch(items, pool_paths=0, films=True):
    """Renders several (Renderer, width, height[, capture_aovs]) items through
    one path pool (rt_render_batch).  Returns (list of RenderResult or None
    per item, total rt_render_stats)."""
    L = load()
    jobs = (abi.rt_batch_job * len(items))()
    keep, outs = [], []
    ctx = None
    for k, it in enumerate(items):
        r, w, h = it[0], it[1], it[2]
        aov = bool(it[3]) if len(it) > 3 else False
        st = abi.rt_settings()
        C.memmove(C.byref(st), C.byref(r.settings), C.sizeof(st))
        st.width, st.height, st.capture_aovs = w, h, 1 if aov else 0
        jobs[k].scene = r.scene.h
        jobs[k].settings = C.pointer(st)
        keep.append(st)
        if films:
            film, planes, count = r._film(w, h, aov)
            keep.append(film)
            jobs[k].film = C.pointer(film)
            outs.append((r, w, h, planes, count))
        ctx = r.scene.ctx if ctx is None else ctx
    total = abi.rt_render_stats()
    ctx.check(L.rt_render_batch(ctx.h, jobs, len(items), pool_paths, C.byref(total)))
    results = [r._result(w, h, planes, count, None) for r, w, h, planes, count in outs] if films else None
    return results, total
