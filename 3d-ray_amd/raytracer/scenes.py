"""Scene assembly for the C ABI: the host-side stand-in for SceneLoader.Load.

SceneBuilder appends primitives, materials and lights in the order the
reference's SceneLoader produces them (SceneLoader.cs:282-553):

  * entities in declaration order (each its own object; a mesh is one object
    whose faces are contiguous, Mesh.cs:75-116);
  * sphere/area lights append a visible emissive proxy primitive to the
    object list as they are created (CreateSphereLight / CreateAreaLight,
    SceneLoader.cs:3550-3636) and register themselves in the light list;
  * after all entities and lights, every non-proxy emissive primitive becomes
    a GeometryLight, in object order (ExtractGeometryLights,
    SceneLoader.cs:805-898), with 4 shadow samples unless overridden.

The generators at the bottom build the benchmark workloads procedurally with
seeded numpy streams (there is no network for asset downloads).
"""
import ctypes as C

import numpy as np

from . import abi

PRIM_DTYPE = np.dtype([("kind", "<i4"), ("material", "<i4"), ("flags", "<i4"), ("mesh", "<i4"),
                       ("v", "<f4", (12,))])
MAT_DTYPE = np.dtype([("kind", "<i4"), ("flags", "<i4"), ("color", "<f4", (3,)), ("intensity", "<f4"),
                      ("metallic", "<f4"), ("roughness", "<f4"), ("specular", "<f4"), ("specular_tint", "<f4"),
                      ("sheen", "<f4"), ("sheen_tint", "<f4"), ("sheen_roughness", "<f4"), ("clearcoat", "<f4"),
                      ("clearcoat_gloss", "<f4"), ("spec_trans", "<f4"), ("ior", "<f4"), ("anisotropic", "<f4"),
                      ("anisotropic_rotation", "<f4"), ("diff_trans", "<f4"), ("coat_ior", "<f4"),
                      ("coat_roughness", "<f4"), ("thin_film_thickness", "<f4"), ("thin_film_ior", "<f4"),
                      ("transmission_depth", "<f4"), ("transmission_color", "<f4", (3,)), ("image", "<i4"), ("reserved", "<i4", (3,))])
LIGHT_DTYPE = np.dtype([("kind", "<i4"), ("shadow_samples", "<i4"), ("primitive", "<i4"), ("material", "<i4"),
                        ("color", "<f4", (3,)), ("intensity", "<f4"), ("position", "<f4", (3,)), ("radius", "<f4"),
                        ("u", "<f4", (3,)), ("soft_radius", "<f4"), ("v", "<f4", (3,)), ("reserved", "<f4")])
NODE_DTYPE = np.dtype([("kind", "<i4"), ("flags", "<i4"), ("first", "<i4"), ("count", "<i4"), ("child", "<i4"),
                       ("material", "<i4"), ("reserved", "<i4", (2,)), ("matrix", "<f4", (16,)),
                       ("inverse", "<f4", (16,))])
assert PRIM_DTYPE.itemsize == 64 and MAT_DTYPE.itemsize == 128 and LIGHT_DTYPE.itemsize == 80
assert NODE_DTYPE.itemsize == 160

_DISNEY_DEFAULTS = dict(metallic=0.0, roughness=0.5, specular=0.5, specular_tint=0.0, sheen=0.0, sheen_tint=0.5,
                        sheen_roughness=0.3, clearcoat=0.0, clearcoat_gloss=1.0, spec_trans=0.0, ior=1.5,
                        anisotropic=0.0, anisotropic_rotation=0.0, diff_trans=0.0, coat_ior=1.5,
                        coat_roughness=0.0, thin_film_thickness=0.0, thin_film_ior=1.5, transmission_depth=0.0)


def material_records(n):
    """n material records holding the DisneyBsdf constructor defaults."""
    m = np.zeros(n, MAT_DTYPE)
    for k, v in _DISNEY_DEFAULTS.items():
        m[k] = v
    m["color"] = 0.8
    m["intensity"] = 1.0
    m["transmission_color"] = 1.0
    return m


class SceneDesc:
    """Owns the numpy tables behind one rt_scene_desc (keeps them alive)."""

    def __init__(self, prims, mats, lights, camera, sky=(0.0, 0.0, 0.0), medium=None, texcoords=None,
                 texture=None, normals=None, images=None, hierarchy=None):
        self.prims = np.ascontiguousarray(prims, PRIM_DTYPE)
        self.mats = np.ascontiguousarray(mats, MAT_DTYPE)
        self.lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
        self.camera = camera
        self.sky = tuple(float(x) for x in sky)
        self.medium = medium
        d = abi.rt_scene_desc()
        d.primitives = C.cast(self.prims.ctypes.data, C.POINTER(abi.rt_primitive))
        d.num_primitives = len(self.prims)
        d.materials = C.cast(self.mats.ctypes.data, C.POINTER(abi.rt_material))
        d.num_materials = len(self.mats)
        d.lights = C.cast(self.lights.ctypes.data, C.POINTER(abi.rt_light)) if len(self.lights) else None
        d.num_lights = len(self.lights)
        d.camera = camera
        d.sky_color[:] = self.sky
        if medium is not None:
            d.global_medium.enabled = 1
            d.global_medium.sigma_a[:] = medium[0]
            d.global_medium.sigma_s[:] = medium[1]
        # texture set (5, h, w) fp32 and optional per-primitive vertex UVs (n, 6)
        self.texture = None if texture is None else np.ascontiguousarray(texture, np.float32)
        self.texcoords = None if texcoords is None else np.ascontiguousarray(texcoords, np.float32).reshape(-1, 6)
        if self.texture is not None:
            assert self.texture.ndim == 3 and self.texture.shape[0] == abi.TEX_CHANNELS
            d.texels = C.cast(self.texture.ctypes.data, C.POINTER(C.c_float))
            d.tex_height, d.tex_width = self.texture.shape[1], self.texture.shape[2]
        if self.texcoords is not None:
            assert len(self.texcoords) == len(self.prims)
            d.texcoords = C.cast(self.texcoords.ctypes.data, C.POINTER(C.c_float))
        # optional per-primitive vertex normals (n, 9); an all-zero row is a flat triangle
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        if self.normals is not None:
            assert len(self.normals) == len(self.prims)
            d.normals = C.cast(self.normals.ctypes.data, C.POINTER(C.c_float))
        # image textures: [(rgba8 (h, w, 4) uint8, (scale_u, scale_v))]
        self.images = [(np.ascontiguousarray(px, np.uint8), tuple(float(np.float32(x)) for x in sc))
                       for px, sc in (images or [])]
        if self.images:
            self._image_recs = (abi.rt_image * len(self.images))()
            for k, (px, sc) in enumerate(self.images):
                assert px.ndim == 3 and px.shape[2] == 4
                r = self._image_recs[k]
                r.rgba8 = C.cast(px.ctypes.data, C.POINTER(C.c_uint8))
                r.height, r.width = px.shape[0], px.shape[1]
                r.scale_u, r.scale_v = sc
            d.images = C.cast(self._image_recs, C.c_void_p)
            d.num_images = len(self.images)
        # object hierarchy: (nodes NODE_DTYPE, group child lists int32, roots int32)
        self.hierarchy = None
        if hierarchy is not None:
            nodes, children, roots = hierarchy
            self.hierarchy = (np.ascontiguousarray(nodes, NODE_DTYPE), np.ascontiguousarray(children, np.int32),
                              np.ascontiguousarray(roots, np.int32))
            n, ch, r = self.hierarchy
            d.nodes = n.ctypes.data if len(n) else None
            d.num_nodes = len(n)
            d.node_children = C.cast(ch.ctypes.data, C.POINTER(C.c_int32)) if len(ch) else None
            d.num_node_children = len(ch)
            d.roots = C.cast(r.ctypes.data, C.POINTER(C.c_int32)) if len(r) else None
            d.num_roots = len(r)
        self.desc = d

    @property
    def ptr(self):
        return C.byref(self.desc)


def make_camera(look_from, look_at, vfov, aspect, vup=(0.0, 1.0, 0.0), aperture=0.0, focus_dist=1.0):
    c = abi.rt_camera()
    c.look_from[:] = look_from
    c.look_at[:] = look_at
    c.vup[:] = vup
    c.vfov_deg, c.aspect, c.aperture, c.focus_dist = vfov, aspect, aperture, focus_dist
    return c


class SceneBuilder:
    def __init__(self):
        self._prims = []          # list of numpy chunks (PRIM_DTYPE)
        self._count = 0
        self._mats = []
        self._mat_count = 0
        self._lights = []
        self._next_mesh = 0
        self._normals = []        # (first primitive, (n, 9) vertex normals) of smooth triangles
        self._uvs = []            # (first primitive, (n, 6) vertex UVs)
        self._images = []         # image textures: (rgba8, (scale_u, scale_v))
        self.geometry_shadow_samples = 4
        self.shadow_samples_override = None
        # object hierarchy (rt_node): top-level items in object order, each a
        # flat object [first, count] or a root node
        self._objects = []        # [first, count, claimed]
        self._top = []            # ("obj", object index) | ("node", node index)
        self._nodes = []
        self._children = []

    # ── materials ───────────────────────────────────────────────────────
    def add_materials(self, recs):
        recs = np.ascontiguousarray(recs, MAT_DTYPE)
        first = self._mat_count
        self._mats.append(recs)
        self._mat_count += len(recs)
        return first

    def material(self, kind, color=(0.8, 0.8, 0.8), **kw):
        m = material_records(1)
        m["kind"] = kind
        m["color"][0] = color
        flags = 0
        for k, v in kw.items():
            if k == "light_proxy":
                flags |= abi.RT_MATF_LIGHT_PROXY if v else 0
            elif k == "thin_walled":
                flags |= abi.RT_MATF_THIN_WALLED if v else 0
            elif k == "transmission_color":
                flags |= abi.RT_MATF_TRANSMISSION_COLOR
                m["transmission_color"][0] = v
            elif k == "coat_roughness":
                flags |= abi.RT_MATF_COAT_ROUGHNESS
                m["coat_roughness"] = v
            elif k == "textured":
                flags |= abi.RT_MATF_TEXTURED if v else 0
            elif k == "image":
                flags |= abi.RT_MATF_IMAGE
                m["image"] = v
            else:
                m[k] = v
        m["flags"] = flags
        return self.add_materials(m)

    @property
    def material_count(self):
        return self._mat_count

    def material_record(self, index):
        """The MAT_DTYPE record of material `index`."""
        k = int(index)
        for chunk in self._mats:
            if k < len(chunk):
                return chunk[k]
            k -= len(chunk)
        raise IndexError(index)

    def image(self, rgba8, scale=(1.0, 1.0)):
        """An ImageTexture from 8-bit RGBA pixels (h, w, 4), top row first,
        with the YAML uv_scale; returns its index for image=."""
        self._images.append((np.ascontiguousarray(rgba8, np.uint8), tuple(scale)))
        return len(self._images) - 1

    def lambertian(self, albedo, image=None):
        if image is not None:
            return self.material(abi.RT_MAT_LAMBERTIAN, albedo, image=image)
        return self.material(abi.RT_MAT_LAMBERTIAN, albedo)

    def emissive(self, color, intensity=1.0, light_proxy=False):
        return self.material(abi.RT_MAT_EMISSIVE, color, intensity=intensity, light_proxy=light_proxy)

    def dielectric(self, ior, albedo=(1.0, 1.0, 1.0)):
        return self.material(abi.RT_MAT_DIELECTRIC, albedo, ior=ior)

    def disney(self, base_color, **kw):
        return self.material(abi.RT_MAT_DISNEY, base_color, **kw)

    # ── primitives ──────────────────────────────────────────────────────
    def _append(self, kind, mat, values, flags=0, mesh=-1):
        values = np.asarray(values, np.float32).ravel()
        p = np.zeros(1, PRIM_DTYPE)
        p["kind"] = kind
        p["material"] = mat
        p["flags"] = flags
        p["mesh"] = mesh
        p["v"][0, :len(values)] = values
        first = self._count
        self._prims.append(p)
        self._count += 1
        self._object(first, 1)
        return first

    def _object(self, first, count):
        self._top.append(("obj", len(self._objects)))
        self._objects.append([first, count, False])

    # ── object hierarchy (rt_node: Group / Transform / Instance) ────────
    def _node(self, **kw):
        n = np.zeros(1, NODE_DTYPE)
        n["child"] = -1
        n["material"] = -1
        for k, v in kw.items():
            n[k] = v
        self._nodes.append(n)
        return len(self._nodes) - 1

    def node_prims(self, first):
        """The object starting at primitive `first` (a lone primitive or a
        mesh) as a PRIMITIVES node; it leaves the top-level object list."""
        for o in self._objects:
            if o[0] == first:
                if o[2]:
                    raise ValueError("object already wrapped")
                o[2] = True
                return self._node(kind=abi.RT_NODE_PRIMITIVES, first=first, count=o[1])
        raise ValueError(f"no object starts at primitive {first}")

    def node_group(self, children, camera_invisible=False):
        first = len(self._children)
        self._children.extend(int(c) for c in children)
        return self._node(kind=abi.RT_NODE_GROUP, first=first, count=len(children),
                          flags=abi.RT_NODEF_CAMERA_INVISIBLE if camera_invisible else 0)

    def node_transform(self, child, matrix, inverse=None, camera_invisible=False):
        """Transform(child, matrix): matrix is the 4x4 object-to-world matrix
        in System.Numerics layout (row vectors, translation in row 4); the
        inverse defaults to the float64 inverse rounded to float32."""
        m = np.asarray(matrix, np.float32).reshape(4, 4)
        inv = (np.linalg.inv(m.astype(np.float64)).astype(np.float32) if inverse is None
               else np.asarray(inverse, np.float32).reshape(4, 4))
        return self._node(kind=abi.RT_NODE_TRANSFORM, child=child, matrix=m.ravel(), inverse=inv.ravel(),
                          flags=abi.RT_NODEF_CAMERA_INVISIBLE if camera_invisible else 0)

    def node_instance(self, child, material=-1):
        return self._node(kind=abi.RT_NODE_INSTANCE, child=child, material=material)

    def add_root(self, node):
        """Appends `node` to the world's object list at this point of the
        object order."""
        self._top.append(("node", int(node)))

    def _hierarchy(self):
        if not self._nodes:
            return None
        roots = []
        for kind, i in self._top:
            if kind == "node":
                roots.append(i)
            elif not self._objects[i][2]:
                first, count, _ = self._objects[i]
                roots.append(self._node(kind=abi.RT_NODE_PRIMITIVES, first=first, count=count))
        return (np.concatenate(self._nodes), np.asarray(self._children, np.int32), np.asarray(roots, np.int32))

    def triangle(self, v0, v1, v2, mat, flags=0, normals=None, uvs=None):
        """A Triangle, or a SmoothTriangle when vertex normals (n0, n1, n2)
        are given; uvs = (uv0, uv1, uv2) are its vertex texture coordinates."""
        first = self._append(abi.RT_PRIM_TRIANGLE, mat, np.concatenate([v0, v1, v2]), flags)
        if normals is not None:
            self._normals.append((first, np.asarray(normals, np.float32).reshape(1, 9)))
        if uvs is not None:
            self._uvs.append((first, np.asarray(uvs, np.float32).reshape(1, 6)))
        return first

    def triangles(self, tris, mats, flags=0):
        """tris: (n, 3, 3) standalone triangle entities; mats: int or (n,)."""
        tris = np.asarray(tris, np.float32).reshape(-1, 9)
        first = self._count
        p = np.zeros(len(tris), PRIM_DTYPE)
        p["kind"] = abi.RT_PRIM_TRIANGLE
        p["material"] = mats
        p["flags"] = flags
        p["mesh"] = -1
        p["v"][:, :9] = tris
        self._prims.append(p)
        self._count += len(tris)
        for k in range(len(tris)):
            self._object(first + k, 1)
        return first

    def mesh(self, vertices, faces, mats, normals=None, normal_faces=None, uvs=None, uv_faces=None):
        """A Mesh object (nested BVH): faces contiguous, in face order.  With
        vertex normals (indexed by normal_faces, default faces) the faces are
        SmoothTriangles; uvs (indexed by uv_faces) are vertex texture
        coordinates."""
        vertices = np.asarray(vertices, np.float32)
        faces = np.asarray(faces, np.int64)
        tris = vertices[faces].reshape(-1, 9)
        fn = fu = None
        if normals is not None:
            nf = faces if normal_faces is None else np.asarray(normal_faces, np.int64)
            fn = np.asarray(normals, np.float32)[nf].reshape(-1, 9)
        if uvs is not None:
            tf = faces if uv_faces is None else np.asarray(uv_faces, np.int64)
            fu = np.asarray(uvs, np.float32)[tf].reshape(-1, 6)
        return self.mesh_triangles(tris, mats, fn, fu)

    def mesh_triangles(self, tris, mats, normals=None, uvs=None, flags=0):
        """A Mesh object from its triangle list (n, 9) in Mesh order, with
        per-triangle vertex normals (n, 9; zero rows stay flat) and vertex
        UVs (n, 6); flags apply to the whole mesh."""
        tris = np.asarray(tris, np.float32).reshape(-1, 9)
        if normals is not None:
            self._normals.append((self._count, np.asarray(normals, np.float32).reshape(-1, 9)))
        if uvs is not None:
            self._uvs.append((self._count, np.asarray(uvs, np.float32).reshape(-1, 6)))
        mesh_id = self._next_mesh
        self._next_mesh += 1
        first = self._count
        p = np.zeros(len(tris), PRIM_DTYPE)
        p["kind"] = abi.RT_PRIM_TRIANGLE
        p["material"] = mats
        p["flags"] = flags
        p["mesh"] = mesh_id
        p["v"][:, :9] = tris
        self._prims.append(p)
        self._count += len(tris)
        self._object(first, len(tris))
        return first

    def sphere(self, center, radius, mat, flags=0):
        return self._append(abi.RT_PRIM_SPHERE, mat, list(center) + [radius], flags)

    def quad(self, q, u, v, mat, flags=0):
        return self._append(abi.RT_PRIM_QUAD, mat, list(q) + list(u) + list(v), flags)

    def plane(self, point, normal, mat):
        return self._append(abi.RT_PRIM_PLANE, mat, list(point) + list(normal))

    # ── lights (SceneLoader.CreateLight, SceneLoader.cs:3463-3636) ──────
    def _light(self, **kw):
        L = np.zeros(1, LIGHT_DTYPE)
        L["primitive"] = -1
        L["material"] = -1
        for k, v in kw.items():
            L[k] = v
        self._lights.append(L)
        return sum(len(x) for x in self._lights) - 1

    def sphere_light(self, position, radius, color=(1.0, 1.0, 1.0), intensity=20.0, shadow_samples=4,
                     visible_to_camera=True):
        ss = self.shadow_samples_override or shadow_samples
        if radius <= 0:
            radius = 0.5
        pm = self.emissive(color, intensity, light_proxy=True)
        flags = 0 if visible_to_camera else abi.RT_PRIMF_CAMERA_INVISIBLE
        prim = self.sphere(position, radius, pm, flags)
        return self._light(kind=abi.RT_LIGHT_SPHERE, shadow_samples=ss, primitive=prim, material=pm, color=color,
                           intensity=intensity, position=position, radius=radius)

    def area_light(self, corner, u, v, color=(1.0, 1.0, 1.0), intensity=20.0, shadow_samples=4, soft_radius=0.0,
                   visible_to_camera=True):
        ss = self.shadow_samples_override or shadow_samples
        pm = self.emissive(color, intensity, light_proxy=True)
        flags = abi.RT_PRIMF_BACKFACE_CULLED | (0 if visible_to_camera else abi.RT_PRIMF_CAMERA_INVISIBLE)
        prim = self.quad(corner, u, v, pm, flags)
        return self._light(kind=abi.RT_LIGHT_AREA, shadow_samples=ss, primitive=prim, material=pm, color=color,
                           intensity=intensity, position=corner, u=u, v=v, soft_radius=soft_radius)

    def directional_light(self, direction, color=(1.0, 1.0, 1.0), intensity=1.0, angular_radius_deg=0.0):
        ss = self.shadow_samples_override or 0
        return self._light(kind=abi.RT_LIGHT_DIRECTIONAL, shadow_samples=ss, color=color, intensity=intensity,
                           position=direction, radius=angular_radius_deg)

    # ── finish ──────────────────────────────────────────────────────────
    def build(self, camera, sky=(0.0, 0.0, 0.0), medium=None):
        prims = np.concatenate(self._prims) if self._prims else np.zeros(0, PRIM_DTYPE)
        mats = np.concatenate(self._mats) if self._mats else material_records(0)
        lights = list(self._lights)
        # ExtractGeometryLights: non-proxy emissive primitives, object order
        kinds = mats["kind"][prims["material"]]
        proxy = (mats["flags"][prims["material"]] & abi.RT_MATF_LIGHT_PROXY) != 0
        emissive = np.nonzero((kinds == abi.RT_MAT_EMISSIVE) & ~proxy)[0]
        normals = None
        if self._normals:
            normals = np.zeros((len(prims), 9), np.float32)
            for first, n in self._normals:
                normals[first:first + len(n)] = n
        texcoords = None
        if self._uvs:
            texcoords = np.tile(np.array([0, 0, 1, 0, 0, 1], np.float32), (len(prims), 1))   # barycentric UVs
            for first, t in self._uvs:
                texcoords[first:first + len(t)] = t
        if len(emissive):
            if np.any(prims["mesh"][emissive] >= 0) or np.any(prims["kind"][emissive] == abi.RT_PRIM_PLANE):
                raise ValueError("emissive meshes/planes are not supported as geometry lights")
            if normals is not None and np.any(normals[emissive] != 0):
                raise ValueError("emissive smooth triangles are not supported as geometry lights")
            g = np.zeros(len(emissive), LIGHT_DTYPE)
            g["kind"] = abi.RT_LIGHT_GEOMETRY
            g["shadow_samples"] = self.shadow_samples_override or self.geometry_shadow_samples
            g["primitive"] = emissive
            g["material"] = prims["material"][emissive]
            lights.append(g)
        # AddEnvironmentLight: a flat sky that can be sampled (luminance above
        # 1e-6, FlatSky.HasImportanceSampling) becomes the last light
        # (SceneLoader.cs:524-528)
        sk = np.asarray(sky, np.float32)
        lum = (np.float32(0.2126) * sk[0] + np.float32(0.7152) * sk[1]) + np.float32(0.0722) * sk[2]
        if lum > np.float32(1e-6):
            env = np.zeros(1, LIGHT_DTYPE)
            env["kind"] = abi.RT_LIGHT_ENVIRONMENT
            env["shadow_samples"] = self.shadow_samples_override or 1
            env["primitive"] = -1
            env["material"] = -1
            env["color"] = sk
            env["intensity"] = 1.0
            lights.append(env)
        lights = np.concatenate(lights) if lights else np.zeros(0, LIGHT_DTYPE)
        return SceneDesc(prims, mats, lights, camera, sky, medium, texcoords=texcoords, normals=normals,
                         images=self._images, hierarchy=self._hierarchy())


# ═══════════════════════════════════════════════════════════════════════
# Workload generators
# ═══════════════════════════════════════════════════════════════════════

def cornell(width=256, height=256, light="geometry"):
    """Small Cornell-style box: 5 walls as 10 triangles of the
    box x∈[-1,1], y∈[0,2], z∈[-1,1] (open +z), a 0.25×0.25 emissive ceiling
    quad of radiance 15, a Lambertian sphere (0.73) and a Disney glass sphere
    (ior 1.5, roughness 0, spec_trans 1), both r = 0.3."""
    b = SceneBuilder()
    white = b.lambertian((0.73, 0.73, 0.73))
    red = b.lambertian((0.65, 0.05, 0.05))
    green = b.lambertian((0.12, 0.45, 0.15))
    P = lambda x, y, z: np.array([x, y, z], np.float32)
    walls = [
        (P(-1, 0, -1), P(1, 0, -1), P(1, 0, 1), white), (P(-1, 0, -1), P(1, 0, 1), P(-1, 0, 1), white),   # floor
        (P(-1, 2, -1), P(-1, 2, 1), P(1, 2, 1), white), (P(-1, 2, -1), P(1, 2, 1), P(1, 2, -1), white),   # ceiling
        (P(-1, 0, -1), P(-1, 2, -1), P(1, 2, -1), white), (P(-1, 0, -1), P(1, 2, -1), P(1, 0, -1), white),  # back
        (P(-1, 0, -1), P(-1, 0, 1), P(-1, 2, 1), red), (P(-1, 0, -1), P(-1, 2, 1), P(-1, 2, -1), red),      # left
        (P(1, 0, -1), P(1, 2, -1), P(1, 2, 1), green), (P(1, 0, -1), P(1, 2, 1), P(1, 0, 1), green),        # right
    ]
    for a, c, d, m in walls:
        b.triangle(a, c, d, m)
    b.sphere((-0.45, 0.3, -0.3), 0.3, white)
    glass = b.disney((1.0, 1.0, 1.0), spec_trans=1.0, ior=1.5, roughness=0.0)
    b.sphere((0.45, 0.3, 0.2), 0.3, glass)
    if light == "geometry":
        em = b.emissive((1.0, 1.0, 1.0), 15.0)
        b.quad((-0.125, 1.999, -0.125), (0.25, 0.0, 0.0), (0.0, 0.0, 0.25), em)
    else:
        b.area_light((-0.125, 1.999, -0.125), (0.25, 0.0, 0.0), (0.0, 0.0, 0.25), intensity=15.0, shadow_samples=4)
    cam = make_camera((0.0, 1.0, 3.9), (0.0, 1.0, 0.0), 39.0, width / height)
    return b.build(cam)


def icosphere(subdiv):
    """Unit icosphere: 20·4^subdiv faces, vertices on the unit sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(subdiv):
        nv = len(v)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b_ = uk // nv, uk % nv
        mid = v[a] + v[b_]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        m = inv.reshape(3, -1).T + nv   # midpoints of edges (01, 12, 20)
        f = np.concatenate([np.stack([f[:, 0], m[:, 0], m[:, 2]], 1), np.stack([f[:, 1], m[:, 1], m[:, 0]], 1),
                            np.stack([f[:, 2], m[:, 2], m[:, 1]], 1), m])
    return v, f


def fbm(p, octaves, seed):
    """Seeded 3D value-noise fBm in [-1, 1] (lacunarity 2, gain 0.5)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(4096).astype(np.int64)
    vals = rng.uniform(-1.0, 1.0, 4096)

    def lattice(ix, iy, iz):
        h = perm[(ix & 4095)]
        h = perm[(h + iy) & 4095]
        h = perm[(h + iz) & 4095]
        return vals[h]

    total = np.zeros(len(p))
    amp, freq, norm = 1.0, 2.0, 0.0
    for _ in range(octaves):
        q = p * freq + 17.0
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3.0 - 2.0 * fr)
        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    c = lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
                    wx = w[:, 0] if dx else 1.0 - w[:, 0]
                    wy = w[:, 1] if dy else 1.0 - w[:, 1]
                    wz = w[:, 2] if dz else 1.0 - w[:, 2]
                    acc = acc + c * wx * wy * wz
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def displaced_icosphere_scene(width=1920, height=1080, subdiv=8, seed=0xC0FFEE, sphere_lights=4, sun=True):
    """Single large scene: icosphere subdivided to 20·4^subdiv
    triangles (1,310,720 at subdiv 8), radius 1 at (0,1,0), displaced along
    the normal by 6-octave fBm of amplitude 0.05, on a 2-triangle ground;
    per-triangle Disney parameters drawn from the seed (baseColor U[0.05,0.95]^3,
    roughness U[0.05,1], metallic Bernoulli(0.3), specular 0.5); 4 sphere
    lights (r 0.1, radiance U[5,20]) and a hard sun.  Every triangle is its
    own entity, as the reference represents per-triangle materials."""
    rng = np.random.default_rng(seed)
    v, f = icosphere(subdiv)
    d = fbm(v, 6, seed)
    v = v * (1.0 + 0.05 * d)[:, None]
    v = v + np.array([0.0, 1.0, 0.0])
    tris = v[f].astype(np.float32)
    n = len(tris)
    b = SceneBuilder()
    b.shadow_samples_override = 1
    ground = b.disney((0.5, 0.5, 0.5), roughness=0.8)
    g = np.array([[[-10, 0, -10], [-10, 0, 10], [10, 0, 10]], [[-10, 0, -10], [10, 0, 10], [10, 0, -10]]], np.float32)
    b.triangles(g, ground)
    mats = material_records(n)
    mats["kind"] = abi.RT_MAT_DISNEY
    mats["color"] = rng.uniform(0.05, 0.95, (n, 3))
    mats["roughness"] = rng.uniform(0.05, 1.0, n)
    mats["metallic"] = (rng.random(n) < 0.3).astype(np.float32)
    mats["specular"] = 0.5
    m0 = b.add_materials(mats)
    b.triangles(tris, m0 + np.arange(n))
    for k in range(sphere_lights):
        ang = 2.0 * np.pi * (k + 0.25) / max(sphere_lights, 1)
        pos = (1.8 * np.cos(ang), 2.3 + 0.2 * (k % 2), 1.8 * np.sin(ang))
        b.sphere_light(pos, 0.1, (1.0, 1.0, 1.0), float(rng.uniform(5.0, 20.0)), shadow_samples=1)
    if sun:
        b.directional_light((-0.4, -1.0, -0.3), (1.0, 1.0, 1.0), 3.0)
    cam = make_camera((0.0, 1.2, 3.5), (0.0, 0.9, 0.0), 40.0, width / height)
    return b.build(cam)


def sphere_uv(dirs):
    """Sphere.GetSphereUV (Sphere.cs:260-269) of unit directions (n, 3)."""
    theta = np.arccos(np.clip(-dirs[:, 1], -1.0, 1.0))
    phi = np.arctan2(-dirs[:, 2], dirs[:, 0]) + np.pi
    return np.stack([phi / (2.0 * np.pi), theta / np.pi], 1)


def material_texture(size=512, seed=0x7E47, cell=16, base=(0.05, 0.95), rough=(0.1, 1.0), p_metal=0.3):
    """Seeded texture set (5, size, size): piecewise-constant cells of
    `cell` texels with baseColor U[base]^3, roughness U[rough] and metallic
    Bernoulli(p_metal) in {0, 1}."""
    rng = np.random.default_rng(seed)
    n = size // cell
    c = rng.uniform(base[0], base[1], (3, n, n))
    r = rng.uniform(rough[0], rough[1], (1, n, n))
    m = (rng.random((1, n, n)) < p_metal).astype(np.float64)
    t = np.concatenate([c, r, m]).repeat(cell, 1).repeat(cell, 2)
    return np.ascontiguousarray(t, np.float32)


def textured_icosphere_scene(width=256, height=256, subdiv=6, seed=0xC0FFEE, texture=None, sphere_lights=4,
                             sun=True):
    """Differentiation workload: the displaced icosphere of the single large
    scene at 20·4^subdiv triangles (81,920 at subdiv 6) with ONE textured
    Disney material (specular 0.5) whose baseColor, roughness and metallic
    come from a texture set mapped spherically (vertex UVs = GetSphereUV of
    the undisplaced vertex direction, u unwrapped across the seam per
    triangle); same ground, lights and camera.  texture: (5, h, w), default
    the constant 0.5 start texture at 512^2."""
    rng = np.random.default_rng(seed)
    v, f = icosphere(subdiv)
    uvv = sphere_uv(v)
    d = fbm(v, 6, seed)
    v = v * (1.0 + 0.05 * d)[:, None]
    v = v + np.array([0.0, 1.0, 0.0])
    tris = v[f].astype(np.float32)
    n = len(tris)
    tuv = uvv[f]                                   # (n, 3, 2)
    u = tuv[:, :, 0]
    wrap = (u.max(1) - u.min(1)) > 0.5
    u[wrap] = np.where(u[wrap] < 0.5, u[wrap] + 1.0, u[wrap])
    b = SceneBuilder()
    b.shadow_samples_override = 1
    ground = b.disney((0.5, 0.5, 0.5), roughness=0.8)
    g = np.array([[[-10, 0, -10], [-10, 0, 10], [10, 0, 10]], [[-10, 0, -10], [10, 0, 10], [10, 0, -10]]], np.float32)
    b.triangles(g, ground)
    first = b._count
    mat = b.disney((0.5, 0.5, 0.5), specular=0.5, textured=True)
    b.triangles(tris, mat)
    for k in range(sphere_lights):
        ang = 2.0 * np.pi * (k + 0.25) / max(sphere_lights, 1)
        pos = (1.8 * np.cos(ang), 2.3 + 0.2 * (k % 2), 1.8 * np.sin(ang))
        b.sphere_light(pos, 0.1, (1.0, 1.0, 1.0), float(rng.uniform(5.0, 20.0)), shadow_samples=1)
    if sun:
        b.directional_light((-0.4, -1.0, -0.3), (1.0, 1.0, 1.0), 3.0)
    cam = make_camera((0.0, 1.2, 3.5), (0.0, 0.9, 0.0), 40.0, width / height)
    desc = b.build(cam)
    tc = np.zeros((len(desc.prims), 6), np.float32)
    tc[first:first + n] = tuv.reshape(n, 6)
    if texture is None:
        texture = np.full((abi.TEX_CHANNELS, 512, 512), 0.5, np.float32)
    return SceneDesc(desc.prims, desc.mats, desc.lights, desc.camera, desc.sky, desc.medium, texcoords=tc,
                     texture=texture)


def gerstner_water(n_cells, half_extent=2.0, height=0.45, waves=8, seed=0x5EA):
    """Water heightfield: an n_cells x n_cells grid over [-half, half]^2
    (2 triangles per cell) displaced by the sum of `waves` seeded Gerstner
    waves (horizontal + vertical displacement).  Returns (vertices, faces)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(-half_extent, half_extent, n_cells + 1)
    X, Z = np.meshgrid(g, g, indexing="xy")
    x, z = X.ravel().astype(np.float64), Z.ravel().astype(np.float64)
    px, py, pz = x.copy(), np.full_like(x, height), z.copy()
    lam = rng.uniform(0.4, 1.6, waves)
    amp = rng.uniform(0.006, 0.02, waves)
    ang = rng.uniform(0.0, 2.0 * np.pi, waves)
    phase = rng.uniform(0.0, 2.0 * np.pi, waves)
    for k in range(waves):
        kk = 2.0 * np.pi / lam[k]
        dx, dz = np.cos(ang[k]), np.sin(ang[k])
        q = 0.6 / (kk * amp[k] * waves)          # steepness kept below folding
        th = kk * (dx * x + dz * z) + phase[k]
        px += q * amp[k] * dx * np.cos(th)
        pz += q * amp[k] * dz * np.cos(th)
        py += amp[k] * np.sin(th)
    v = np.stack([px, py, pz], 1).astype(np.float32)
    n1 = n_cells + 1
    i, j = np.meshgrid(np.arange(n_cells), np.arange(n_cells), indexing="xy")
    a = (j * n1 + i).ravel()
    b, c, d = a + 1, a + n1, a + n1 + 1
    f = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)]).astype(np.int64)
    return v, f


def caustics_scene(width=1920, height=1080, water_cells=200, seed=0x5EA):
    """Caustics + volumes workload: three Disney glass spheres (ior 1.5,
    Beer-Lambert absorption 0.5 per unit: transmission colour e^-0.5 at
    depth 1), an 80k-triangle Gerstner-wave water sheet (Disney glass, ior
    1.33) over a diffuse floor, a sun plus one sphere light, and a global
    homogeneous fog with sigma_t = 0.05 and albedo 0.9."""
    b = SceneBuilder()
    floor = b.lambertian((0.75, 0.72, 0.68))
    t = float(np.exp(-0.5))
    glass = b.disney((1.0, 1.0, 1.0), spec_trans=1.0, ior=1.5, roughness=0.0, transmission_color=(t, t, t),
                     transmission_depth=1.0)
    water = b.disney((1.0, 1.0, 1.0), spec_trans=1.0, ior=1.33, roughness=0.0)
    b.quad((-6.0, 0.0, -6.0), (0.0, 0.0, 12.0), (12.0, 0.0, 0.0), floor)
    v, f = gerstner_water(water_cells, seed=seed)
    b.mesh(v, f, water)
    for cx, cz, r in ((-0.9, -0.3, 0.35), (0.1, 0.4, 0.3), (1.0, -0.5, 0.4)):
        b.sphere((cx, 0.45 + r + 0.35, cz), r, glass)
    b.directional_light((-0.35, -1.0, -0.25), intensity=3.0)
    b.sphere_light((1.6, 3.0, 1.2), 0.15, intensity=40.0, shadow_samples=1)
    cam = make_camera((0.0, 2.4, 5.2), (0.0, 0.6, 0.0), 40.0, width / height)
    sig_t, albedo = 0.05, 0.9
    return b.build(cam, medium=((sig_t * (1 - albedo),) * 3, (sig_t * albedo,) * 3))


_BOX_TRIS = np.array([
    [[0, 0, 0], [1, 1, 0], [1, 0, 0]], [[0, 0, 0], [0, 1, 0], [1, 1, 0]],      # z = 0
    [[0, 0, 1], [1, 0, 1], [1, 1, 1]], [[0, 0, 1], [1, 1, 1], [0, 1, 1]],      # z = 1
    [[0, 0, 0], [1, 0, 0], [1, 0, 1]], [[0, 0, 0], [1, 0, 1], [0, 0, 1]],      # y = 0
    [[0, 1, 0], [0, 1, 1], [1, 1, 1]], [[0, 1, 0], [1, 1, 1], [1, 1, 0]],      # y = 1
    [[0, 0, 0], [0, 0, 1], [0, 1, 1]], [[0, 0, 0], [0, 1, 1], [0, 1, 0]],      # x = 0
    [[1, 0, 0], [1, 1, 0], [1, 1, 1]], [[1, 0, 0], [1, 1, 1], [1, 0, 1]],      # x = 1
], np.float64) - 0.5


def _random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def batch_scene(target_tris, width, height, seed, lights):
    """One scene of the batched workload: a displaced icosphere of the
    largest subdivision within `target_tris` plus randomly rotated, scaled
    and placed boxes (each box transform applied on the host) making up the
    remaining triangle count, on a ground quad; random per-triangle Disney
    materials; `lights` sphere lights (r 0.1, radiance U[5,20])."""
    rng = np.random.default_rng(seed)
    s = 3
    while s < 8 and 20 * 4 ** (s + 1) <= target_tris:
        s += 1
    v, f = icosphere(s)
    v = v * (1.0 + 0.05 * fbm(v, 6, seed))[:, None] + np.array([0.0, 1.0, 0.0])
    tris = [v[f]]
    nbox = max(0, (int(target_tris) - len(f)) // 12)
    if nbox:
        rot = _random_rotations(rng, nbox)
        scale = rng.uniform(0.04, 0.18, nbox)
        ang = rng.uniform(0.0, 2.0 * np.pi, nbox)
        rad = rng.uniform(1.2, 4.0, nbox)
        pos = np.stack([rad * np.cos(ang), rng.uniform(0.05, 1.5, nbox), rad * np.sin(ang)], 1)
        boxes = np.einsum("nij,tkj->ntki", rot, _BOX_TRIS) * scale[:, None, None, None] + pos[:, None, None, :]
        tris.append(boxes.reshape(-1, 3, 3))
    tris = np.concatenate(tris).astype(np.float32)
    n = len(tris)
    b = SceneBuilder()
    b.shadow_samples_override = 1
    ground = b.disney((0.5, 0.5, 0.5), roughness=0.8)
    b.quad((-10.0, 0.0, -10.0), (0.0, 0.0, 20.0), (20.0, 0.0, 0.0), ground)
    mats = material_records(n)
    mats["kind"] = abi.RT_MAT_DISNEY
    mats["color"] = rng.uniform(0.05, 0.95, (n, 3))
    mats["roughness"] = rng.uniform(0.05, 1.0, n)
    mats["metallic"] = (rng.random(n) < 0.3).astype(np.float32)
    mats["specular"] = 0.5
    m0 = b.add_materials(mats)
    b.triangles(tris, m0 + np.arange(n))
    for k in range(lights):
        ang = rng.uniform(0.0, 2.0 * np.pi)
        pos = (rng.uniform(1.0, 2.5) * np.cos(ang), rng.uniform(2.0, 3.0), rng.uniform(1.0, 2.5) * np.sin(ang))
        b.sphere_light(pos, 0.1, (1.0, 1.0, 1.0), float(rng.uniform(5.0, 20.0)), shadow_samples=1)
    cam = make_camera((0.0, 1.4, 4.5), (0.0, 0.9, 0.0), 45.0, width / height)
    return b.build(cam)


def batch_workload(count=64, seed=0xBA7C4, min_tris=1e4, max_tris=2e6,
                   resolutions=((256, 256), (512, 512), (1024, 1024), (1920, 1080)), spps=(4, 16, 64),
                   depths=(4, 8, 12)):
    """The batched workload: `count` scenes with log-uniform triangle counts
    in [min_tris, max_tris], resolution / spp / depth drawn from the given
    sets, 1-8 sphere lights each.  Returns a list of dicts with the scene
    description and its render settings."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        tris = float(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        w, h = resolutions[rng.integers(len(resolutions))]
        spp = int(spps[rng.integers(len(spps))])
        depth = int(depths[rng.integers(len(depths))])
        lights = int(rng.integers(1, 9))
        out.append({"desc": batch_scene(tris, w, h, int(rng.integers(1 << 30)), lights), "width": int(w),
                    "height": int(h), "spp": spp, "depth": depth, "target_tris": int(tris)})
    return out
