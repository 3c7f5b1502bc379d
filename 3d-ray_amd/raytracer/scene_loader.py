"""YAML scene front-end, shaped like the reference's SceneLoader.Load
(src/RayTracer/Scene/SceneLoader.cs:122-553) for the part of the schema
(SceneData.cs) the renderer implements:

  world.sky        flat colour (BuildSkySettings :1595-1625; no sky = (0.5, 0.7, 1.0))
  world.ground     infinite plane (BuildGround :1324-1374, ResolveGroundMaterial :1528-1564)
  world.medium     homogeneous, isotropic phase (BuildMedium :4135, ParseSigmas :4206-4217)
  materials        lambertian, dielectric, emissive, disney / disney_bsdf / pbr
                   (CreateMaterial :1027-1074, MaterialData defaults SceneData.cs:470-717);
                   an unknown type is a grey-default Lambertian, as in the reference
  entities         sphere, quad, triangle / smooth_triangle (n0-n2 and uv0-uv2 make it
                   a SmoothTriangle), plane / infinite_plane (CreateEntity :2282-2325,
                   CreateTriangleEntity :3430-3461), mesh / obj (CreateMeshEntity
                   :2828-2935 through raytracer/obj_loader.py, no subdivision), group
                   (CreateGroupEntity :3226-3244, BuildChildList :3345-3418) and
                   instance of a named template (CreateInstanceEntity :3259-3337),
                   with scale / rotate / translate (ComputeTransformMatrix :3788-3822)
                   and visible_to_camera
  templates        named groups with a default pose, built once and shared by their
                   instances (:294-321)
  imports          other scene files' materials, entities, lights, templates and
                   mediums, placed before the local ones (ProcessImports :563-675)
  lights           sphere, area, directional / sun (CreateLight :3463-3636, with the
                   emissive proxies it appends); unknown types are skipped
  camera/cameras   ResolveCamera :731-781, focal_pos via ComputeFocusDistance :3744-3786
  geometry lights  every non-proxy emissive primitive (ExtractGeometryLights :805),
                   default 4 shadow samples (GeometryLight.DefaultShadowSamples)

The assembled tables keep SceneLoader's object order — ground, entities,
then the light proxies — which is the primitive order the renderer's IDs
follow.  A transformed entity, a group and an instance become object
hierarchy nodes (rt_node: Transform, Group, Instance and
CameraInvisibleHittable wrappers, the ray mapped into object space as
Transform.Hit does); two cases keep the older form, the transform applied
to the primitive's defining points and vectors (the same surfaces up to
float rounding): emissive entities, whose geometry lights sample
world-space primitives, and every entity of a scene with image textures,
which the library filters only without wrappers (groups and instances are
then refused).  Everything else in the schema (CSG, mesh subdivision,
procedural textures and maps, mix materials, metal, bound media,
non-isotropic phases, point / spot / portal lights, motion, sky models
other than flat, ground shapes other than the plane, emissive objects
inside groups or templates) raises SceneLoadError naming the feature, so a
scene is either loaded as the reference would assemble it or refused.
"""
import math
import os

import numpy as np

from . import abi, obj_loader, scenes

DEFAULT_SKY = (0.5, 0.7, 1.0)           # SceneLoader.DefaultSkyColor
GEOMETRY_SHADOW_SAMPLES = 4             # GeometryLight.DefaultShadowSamples


class SceneLoadError(ValueError):
    pass


def _v3(x, default=None):
    if x is None:
        return default
    if len(x) < 3:
        raise SceneLoadError(f"expected 3 components, got {x!r}")
    return tuple(float(np.float32(c)) for c in x[:3])


def _f(d, key, default):
    v = d.get(key, default)
    return default if v is None else float(np.float32(v))


_UNSUPPORTED_MATERIAL_KEYS = ("normal_map", "bump_map", "coat_normal_map", "subsurface_radius",
                              "displacement", "emission_texture")
MISSING_TEXTURE = (1.0, 0.0, 1.0)        # CreateImageTexture's magenta fallback


def read_image(path):
    """8-bit RGBA pixels (h, w, 4) of an image file, top row first (the
    Rgba32 rows ImageSharp hands ImageTexture)."""
    from PIL import Image
    with Image.open(path) as im:
        if im.mode in ("I;16", "I;16B", "I", "F"):
            raise SceneLoadError(f"image texture '{path}': {im.mode} images are not supported (8-bit only)")
        return np.asarray(im.convert("RGBA"), np.uint8).copy()


def _image_texture(b, t, base_dir, warnings, mid):
    """CreateImageTexture (SceneLoader.cs:1879-1910): the image at `path`
    (relative to the scene file) with uv_scale; a missing path or file is
    the magenta SolidColor, with a warning.  Returns an image index or None."""
    for k in t:
        if k not in ("type", "path", "uv_scale") and t[k] is not None:
            raise SceneLoadError(f"material '{mid}': image texture key '{k}' is not supported")
    path = t.get("path")
    if path is None or not str(path).strip():
        warnings.append("Image texture requires a 'path' field. Using fallback magenta.")
        return None
    path = str(path)
    full = path if os.path.isabs(path) else os.path.join(base_dir or ".", path)
    if not os.path.isfile(full):
        warnings.append(f"Image texture file not found: {full}. Using fallback magenta.")
        return None
    try:
        px = read_image(full)
    except SceneLoadError:
        raise
    except Exception as err:          # the reference warns and falls back
        warnings.append(f"Failed to load image texture '{full}': {err}. Using fallback magenta.")
        return None
    sc = t.get("uv_scale")
    su = float(np.float32(sc[0])) if sc else 1.0
    sv = float(np.float32(sc[1])) if sc and len(sc) > 1 else su
    return b.image(px, (su, sv))


def _material(b, m, base_dir=None, warnings=None):
    """CreateMaterial for the supported kinds; returns the table index."""
    warnings = [] if warnings is None else warnings
    typ = (m.get("type") or "").lower()
    if typ in ("mix", "blend"):
        raise SceneLoadError(f"material '{m.get('id')}': mix materials are not supported")
    if typ == "metal":
        raise SceneLoadError(f"material '{m.get('id')}': the metal material is not supported")
    for k in m:
        if k in _UNSUPPORTED_MATERIAL_KEYS or (k.endswith("_texture") and m[k] is not None):
            raise SceneLoadError(f"material '{m.get('id')}': '{k}' (textures and maps) is not supported")
    color = _v3(m.get("color"), (0.5, 0.5, 0.5))
    img = None
    if m.get("texture") is not None:
        t = m["texture"]
        tt = (t.get("type") or "").lower() if isinstance(t, dict) else ""
        if tt != "image":
            raise SceneLoadError(f"material '{m.get('id')}': texture type '{tt}' is not supported (image only)")
        if typ not in ("lambertian", "disney", "disney_bsdf", "pbr", ""):
            raise SceneLoadError(f"material '{m.get('id')}': image textures on '{typ}' are not supported")
        img = _image_texture(b, t, base_dir, warnings, m.get("id"))
        if img is None:
            color = MISSING_TEXTURE
    if typ == "lambertian":
        return b.lambertian(color, image=img)
    if typ == "dielectric":
        return b.dielectric(_f(m, "refraction_index", 1.5), color)
    if typ == "emissive":
        return b.emissive(color, _f(m, "intensity", 1.0))
    if typ in ("disney", "disney_bsdf", "pbr"):
        kw = dict(metallic=_f(m, "metallic", 0.0), roughness=_f(m, "roughness", 0.5),
                  specular=_f(m, "specular", 0.5), specular_tint=_f(m, "specular_tint", 0.0),
                  sheen=_f(m, "sheen", 0.0), sheen_tint=_f(m, "sheen_tint", 0.5),
                  sheen_roughness=_f(m, "sheen_roughness", 0.3), clearcoat=_f(m, "clearcoat", 0.0),
                  clearcoat_gloss=_f(m, "clearcoat_gloss", 1.0), spec_trans=_f(m, "spec_trans", 0.0),
                  ior=_f(m, "ior", 1.5), anisotropic=_f(m, "anisotropic", 0.0),
                  anisotropic_rotation=_f(m, "anisotropic_rotation", 0.0),
                  transmission_depth=_f(m, "transmission_depth", 0.0), diff_trans=_f(m, "diff_trans", 0.0),
                  coat_ior=_f(m, "coat_ior", 1.5), thin_film_thickness=_f(m, "thin_film_thickness", 0.0),
                  thin_film_ior=_f(m, "thin_film_ior", 1.5))
        if m.get("thin_walled"):
            kw["thin_walled"] = True
        if m.get("transmission_color") is not None:
            kw["transmission_color"] = _v3(m["transmission_color"])
        cr = _f(m, "coat_roughness", -1.0)
        if cr >= 0.0:
            kw["coat_roughness"] = max(cr, 0.0)
        if img is not None:
            if kw["spec_trans"] != 0.0:
                raise SceneLoadError(f"material '{m.get('id')}': an image-textured Disney needs spec_trans = 0")
            kw["image"] = img
        return b.disney(color, **kw)
    return b.lambertian(color, image=img)


def _transform(e):
    """ComputeTransformMatrix: scale, then rotate about X, Y, Z (degrees),
    then translate; as a 4x4 acting on column vectors, plus the uniform
    scale factor (None when non-uniform)."""
    M = np.eye(4)
    uniform = 1.0
    s = e.get("scale")
    if s is not None:
        sv = [float(s)] * 3 if not isinstance(s, (list, tuple)) else [float(x) for x in s[:3]]
        M = np.diag(sv + [1.0]) @ M
        uniform = sv[0] if sv[0] == sv[1] == sv[2] else None
    r = e.get("rotate")
    if r is not None and len(r) >= 3:
        ax, ay, az = (math.radians(float(x)) for x in r[:3])
        Rx = np.array([[1, 0, 0, 0], [0, math.cos(ax), -math.sin(ax), 0], [0, math.sin(ax), math.cos(ax), 0],
                       [0, 0, 0, 1]])
        Ry = np.array([[math.cos(ay), 0, math.sin(ay), 0], [0, 1, 0, 0], [-math.sin(ay), 0, math.cos(ay), 0],
                       [0, 0, 0, 1]])
        Rz = np.array([[math.cos(az), -math.sin(az), 0, 0], [math.sin(az), math.cos(az), 0, 0], [0, 0, 1, 0],
                       [0, 0, 0, 1]])
        M = Rz @ Ry @ Rx @ M
    t = e.get("translate")
    if t is not None and len(t) >= 3:
        T = np.eye(4)
        T[:3, 3] = [float(x) for x in t[:3]]
        M = T @ M
    return M, uniform


def _rotated(e):
    """True when the entity's transform rotates (a non-zero rotate angle)."""
    r = e.get("rotate")
    return r is not None and any(float(x) % 360.0 != 0.0 for x in list(r)[:3])


def _reads_surface_frame(b, mat):
    """True when the material's shading reads the hit's UV or tangent frame:
    an image texture, or a Disney material with anisotropy."""
    rec = b.material_record(mat)
    if int(rec["flags"]) & abi.RT_MATF_IMAGE:
        return True
    return int(rec["kind"]) == abi.RT_MAT_DISNEY and float(rec["anisotropic"]) != 0.0


def _pt(M, p):
    return tuple((M @ np.array([*p, 1.0]))[:3])


def _vec(M, v):
    return tuple(M[:3, :3] @ np.array(v))


_UNSUPPORTED_ENTITY_KEYS = ("motion", "interior_medium", "exterior_medium", "displacement")


class _Ctx:
    """Per-load state of the object hierarchy: whether wrappers may be used
    (no image texture anywhere: the library filters image textures only in
    scenes without wrappers), the named templates and their built nodes
    (SceneLoader.cs:294-321: built once, on first use)."""

    def __init__(self, hier_ok, templates, get_material, base_dir, warnings):
        self.hier_ok = hier_ok
        self.templates = templates
        self.cache = {}
        self.get_material = get_material
        self.base_dir = base_dir
        self.warnings = warnings


def _row_matrix(M):
    """The column-vector 4x4 of _transform as System.Numerics lays it out
    (row vectors, translation in row 4)."""
    return np.asarray(M, np.float64).T.astype(np.float32)


def _emissive(b, mat):
    return int(b.material_record(mat)["kind"]) == abi.RT_MAT_EMISSIVE


def _mesh_entity(b, e, mat, flags, M, uniform, ident, base_dir, warnings):
    """CreateMeshEntity (SceneLoader.cs:2828-2935) without subdivision or
    displacement: the OBJ at `path` (relative to the scene file) as one
    Mesh; a missing path or file, or a file without triangles, is a warning
    and the entity is skipped, as in the reference."""
    name = e.get("name") or "(unnamed)"
    for k in ("subdivision_iterations", "subdivision_pixel_error"):
        if float(e.get(k) or 0) > 0:
            raise SceneLoadError(f"mesh '{name}': '{k}' (mesh subdivision) is not supported")
    if str(e.get("subdivision_scheme") or "").strip():
        raise SceneLoadError(f"mesh '{name}': 'subdivision_scheme' (mesh subdivision) is not supported")
    path = e.get("path")
    if path is None or not str(path).strip():
        warnings.append(f"Mesh entity '{name}' requires a 'path' to an OBJ file. Skipping.")
        return None
    path = str(path)
    full = path if os.path.isabs(path) else os.path.join(base_dir or ".", path)
    if not os.path.isfile(full):
        warnings.append(f"Mesh entity '{name}': OBJ file not found at '{full}'. Skipping.")
        return None
    w = []
    try:
        mesh = obj_loader.load(full, w)
    except obj_loader.ObjLoadError as err:
        raise SceneLoadError(f"mesh '{name}': {err}") from None
    warnings.extend(f"Mesh '{name}': {x}" for x in w)
    if mesh is None:
        warnings.append(f"Mesh entity '{name}': failed to load '{full}'. Skipping.")
        return None
    tris, normals = mesh.tris, mesh.normals
    if not ident:
        P = tris.reshape(-1, 3).astype(np.float64)
        tris = (P @ M[:3, :3].T + M[:3, 3]).astype(np.float32).reshape(-1, 9)
        if normals is not None:
            # the Transform wrapper maps the interpolated normal by the inverse
            # transpose, which commutes with the interpolation only for a
            # uniform scale
            if uniform is None:
                raise SceneLoadError(f"mesh '{name}': non-uniform scale of a mesh with vertex normals "
                                     "is not supported")
            Nm = np.linalg.inv(M[:3, :3]).T
            normals = (normals.reshape(-1, 3).astype(np.float64) @ Nm.T).astype(np.float32).reshape(-1, 9)
    return b.mesh_triangles(tris, mat, normals, mesh.uvs, flags)


def _entity(b, e, mat, base_dir=None, warnings=None, ctx=None):
    """One top-level entity (the entity loop of SceneLoader.Load, :333-447):
    its object, then Transform and CameraInvisibleHittable around it.  With
    wrappers allowed a transformed entity keeps its object-space primitive
    under a TRANSFORM node (Transform.Hit maps the ray); otherwise, and for
    emissive entities (their geometry lights sample world-space
    primitives), the transform is applied to the primitive's points."""
    typ = (e.get("type") or "").lower()
    for k in _UNSUPPORTED_ENTITY_KEYS:
        if e.get(k) is not None:
            raise SceneLoadError(f"entity '{e.get('name')}': '{k}' is not supported")
    warnings = [] if warnings is None else warnings
    visible = bool(e.get("visible_to_camera", True))
    M, uniform = _transform(e)
    ident = np.array_equal(M, np.eye(4))
    if typ in ("group", "instance"):
        if ctx is None or not ctx.hier_ok:
            raise SceneLoadError(f"entity '{e.get('name')}': groups and instances in a scene with image "
                                 "textures are not supported")
        node = _group_node(b, e, mat, ctx) if typ == "group" else _instance_node(b, e, ctx)
        if node is None:
            return None
        if not ident:
            node = b.node_transform(node, _row_matrix(M), camera_invisible=not visible)
        elif not visible:
            node = b.node_group([node], camera_invisible=True)
        b.add_root(node)
        return node
    if not ident and ctx is not None and ctx.hier_ok and not _emissive(b, mat):
        if typ in ("plane", "infinite_plane") and not visible:
            raise SceneLoadError(f"entity '{e.get('name')}': camera-invisible planes are not supported")
        first = _object_entity(b, e, typ, mat, 0, np.eye(4), 1.0, True, base_dir, warnings)
        if first is None:
            return None
        node = b.node_transform(b.node_prims(first), _row_matrix(M), camera_invisible=not visible)
        b.add_root(node)
        return node
    flags = 0 if visible else abi.RT_PRIMF_CAMERA_INVISIBLE
    return _object_entity(b, e, typ, mat, flags, M, uniform, ident, base_dir, warnings)


def _object_entity(b, e, typ, mat, flags, M, uniform, ident, base_dir, warnings):
    """CreateEntity / CreateMeshEntity: one primitive object (a mesh is one
    object), with M applied to its points when not the identity; returns
    its first primitive (None when skipped)."""
    if typ in ("mesh", "obj"):
        return _mesh_entity(b, e, mat, flags, M, uniform, ident, base_dir, warnings)
    if typ == "sphere":
        c = _v3(e.get("center"), (0.0, 0.0, 0.0))
        r = _f(e, "radius", 1.0)
        if _rotated(e) and _reads_surface_frame(b, mat):
            # the reference's Transform wrapper hits the sphere in object space,
            # so its UV and tangent (Sphere.cs:51-58) turn with the entity; the
            # sphere here is baked into world space and would keep them fixed
            raise SceneLoadError(f"entity '{e.get('name')}': a rotated sphere with an image-textured or "
                                 "anisotropic material is not supported")
        if not ident:
            if uniform is None:
                raise SceneLoadError(f"entity '{e.get('name')}': non-uniform scale of a sphere is not supported")
            c, r = _pt(M, c), r * abs(uniform)
        return b.sphere(c, r, mat, flags)
    if typ == "quad":
        q = _v3(e.get("q"), (0.0, 0.0, 0.0))
        u = _v3(e.get("u"), (1.0, 0.0, 0.0))
        v = _v3(e.get("v"), (0.0, 1.0, 0.0))
        if not ident:
            q, u, v = _pt(M, q), _vec(M, u), _vec(M, v)
        return b.quad(q, u, v, mat, flags)
    if typ in ("triangle", "smooth_triangle"):
        vs = [_v3(e.get("v0"), (0.0, 0.0, 0.0)), _v3(e.get("v1"), (1.0, 0.0, 0.0)), _v3(e.get("v2"), (0.0, 1.0, 0.0))]
        normals = uvs = None
        if all(e.get(k) is not None for k in ("n0", "n1", "n2")):
            # SmoothTriangle; its UVs only when all three are given
            normals = [_v3(e.get(k)) for k in ("n0", "n1", "n2")]
            if all(e.get(k) is not None for k in ("uv0", "uv1", "uv2")):
                uvs = [tuple(float(np.float32(c)) for c in e.get(k)[:2]) for k in ("uv0", "uv1", "uv2")]
            if not ident:
                # the Transform wrapper maps the interpolated normal by the
                # inverse transpose; with a uniform scale that commutes with
                # the interpolation, so the vertex normals are mapped instead
                if uniform is None:
                    raise SceneLoadError(f"entity '{e.get('name')}': non-uniform scale of a smooth triangle "
                                         "is not supported")
                Nm = np.linalg.inv(M[:3, :3]).T
                normals = [tuple(Nm @ np.array(n)) for n in normals]
        if not ident:
            vs = [_pt(M, p) for p in vs]
        return b.triangle(np.array(vs[0], np.float32), np.array(vs[1], np.float32), np.array(vs[2], np.float32),
                          mat, flags, normals=normals, uvs=uvs)
    if typ in ("plane", "infinite_plane"):
        if flags:
            raise SceneLoadError(f"entity '{e.get('name')}': camera-invisible planes are not supported")
        if _rotated(e) and _reads_surface_frame(b, mat):
            # InfinitePlane derives its UV axes from the normal it holds: the
            # reference's rotated object-space axes are not those of the baked
            # world-space normal
            raise SceneLoadError(f"entity '{e.get('name')}': a rotated plane with an image-textured or "
                                 "anisotropic material is not supported")
        p = _v3(e.get("point"), (0.0, 0.0, 0.0))
        n = _v3(e.get("normal"), (0.0, 1.0, 0.0))
        if not ident:
            p = _pt(M, p)
            n = tuple(np.linalg.inv(M[:3, :3]).T @ np.array(n))
        return b.plane(p, n, mat)
    raise SceneLoadError(f"entity '{e.get('name')}': type '{typ}' is not supported")


def _child_node(b, c, fallback, ctx):
    """One child of a group or template (BuildChildList, SceneLoader.cs:3345-3418):
    its own material or the parent's, its object, then its Transform and
    CameraInvisibleHittable; None when the child yields no object."""
    for k in _UNSUPPORTED_ENTITY_KEYS:
        if c.get(k) is not None:
            raise SceneLoadError(f"entity '{c.get('name')}': '{k}' is not supported")
    typ = (c.get("type") or "").lower()
    if c.get("motion"):
        ctx.warnings.append(f"Child entity '{c.get('name') or typ}': 'motion' is only supported on top-level "
                            "entities — ignored.")
    mat = ctx.get_material(c["material"]) if c.get("material") is not None else fallback
    if typ == "group":
        node = _group_node(b, c, mat, ctx)
    elif typ == "instance":
        node = None                        # CreateEntity returns null for an instance child
    elif typ in ("csg",) or typ.startswith("heightfield") or typ == "terrain":
        raise SceneLoadError(f"entity '{c.get('name')}': type '{typ}' is not supported")
    else:
        if _emissive(b, mat):
            raise SceneLoadError(f"entity '{c.get('name')}': emissive objects inside groups or templates "
                                 "are not supported")
        first = _object_entity(b, c, typ, mat, 0, np.eye(4), 1.0, True, ctx.base_dir, ctx.warnings)
        node = None if first is None else b.node_prims(first)
    if node is None:
        return None
    M, _ = _transform(c)
    visible = bool(c.get("visible_to_camera", True))
    if not np.array_equal(M, np.eye(4)):
        node = b.node_transform(node, _row_matrix(M), camera_invisible=not visible)
    elif not visible:
        node = b.node_group([node], camera_invisible=True)
    return node


def _group_node(b, e, mat, ctx):
    """CreateGroupEntity (SceneLoader.cs:3226-3244): a Group of the children."""
    children = e.get("children") or []
    if not children:
        ctx.warnings.append(f"Group '{e.get('name') or '(unnamed)'}' has no children. Skipping.")
        return None
    nodes = [n for n in (_child_node(b, c, mat, ctx) for c in children) if n is not None]
    if not nodes:
        ctx.warnings.append(f"Group '{e.get('name') or '(unnamed)'}': all children failed to load. Skipping.")
        return None
    return b.node_group(nodes)


def _instance_node(b, e, ctx):
    """CreateInstanceEntity + BuildTemplateGeometry (SceneLoader.cs:3259-3337):
    the named template's Group (with its default-pose Transform) built once
    and shared; an Instance around it with the entity's material as override
    when the entity names one."""
    name = e.get("name") or "(unnamed)"
    tname = e.get("template")
    if tname is None or not str(tname).strip():
        ctx.warnings.append(f"Instance '{name}' requires a 'template' name. Skipping.")
        return None
    key = str(tname).lower()
    tdef = ctx.templates.get(key)
    if tdef is None:
        ctx.warnings.append(f"Instance '{name}': template '{tname}' not found. Skipping.")
        return None
    if key not in ctx.cache:
        fallback = ctx.get_material(tdef.get("material"))
        nodes = [n for n in (_child_node(b, c, fallback, ctx) for c in tdef.get("children") or []) if n is not None]
        if not nodes:
            ctx.warnings.append(f"Instance '{name}' of '{tname}': all template children failed to load. Skipping.")
            return None
        shared = b.node_group(nodes)
        M, _ = _transform(tdef)
        if not np.array_equal(M, np.eye(4)):
            shared = b.node_transform(shared, _row_matrix(M))
        ctx.cache[key] = shared
    over = ctx.get_material(e["material"]) if e.get("material") is not None else -1
    return b.node_instance(ctx.cache[key], material=over)


def _merge_imports(data, base_dir, visited, warnings):
    """ProcessImports (SceneLoader.cs:563-675): imported materials, entities,
    lights, templates and mediums go before the local ones (a local id or
    name wins); nested imports first; a repeated file is skipped."""
    imports = data.get("imports") or []
    if not imports:
        return
    import yaml
    got = {k: [] for k in ("materials", "entities", "lights", "templates", "mediums")}
    for imp in imports:
        path = imp.get("path") if isinstance(imp, dict) else imp
        if path is None or not str(path).strip():
            warnings.append("Import entry has empty 'path'. Skipping.")
            continue
        full = os.path.abspath(path if os.path.isabs(str(path)) else os.path.join(base_dir, str(path)))
        if full in visited:
            warnings.append(f"Circular import detected: '{path}'. Skipping.")
            continue
        visited.add(full)
        if not os.path.isfile(full):
            warnings.append(f"Import file not found: '{path}' (resolved to '{full}'). Skipping.")
            continue
        with open(full, "r", encoding="utf-8") as f:
            sub = yaml.safe_load(f)
        if not isinstance(sub, dict):
            warnings.append(f"Failed to parse import file: '{path}'. Skipping.")
            continue
        _merge_imports(sub, os.path.dirname(full), visited, warnings)
        for k in got:
            got[k].extend(sub.get(k) or [])
    for k, v in got.items():
        if v:
            data[k] = v + list(data.get(k) or [])


def _camera(data, selector, width, height):
    """ResolveCamera, then the Camera constructor's inputs."""
    cams = data.get("cameras") or []
    if not cams:
        cd = data.get("camera") or {}
    elif selector is None:
        cd = cams[0]
    else:
        by_name = [c for c in cams if str(c.get("name", "")).lower() == str(selector).lower()]
        if by_name:
            cd = by_name[0]
        elif str(selector).lstrip("-").isdigit() and 0 <= int(selector) < len(cams):
            cd = cams[int(selector)]
        else:
            cd = cams[0]
    for k in ("motion", "shutter"):
        if cd.get(k) is not None:
            raise SceneLoadError(f"camera '{k}' (motion blur) is not supported")
    pos = _v3(cd.get("position"), (0.0, 1.0, -5.0))
    look = _v3(cd.get("look_at"), (0.0, 0.0, 0.0))
    vup = _v3(cd.get("vup"), (0.0, 1.0, 0.0))
    focus = _f(cd, "focal_dist", 1.0)
    fp = cd.get("focal_pos")
    if fp is not None and len(fp) >= 3:
        # ComputeFocusDistance in float32: the focus point projected on the optical axis
        f32 = np.float32
        fwd = np.array(look, f32) - np.array(pos, f32)
        ln = f32(np.sqrt(f32(fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2])))
        if ln >= f32(1e-6):
            fwd = fwd / ln
            d = np.array([float(x) for x in fp[:3]], f32) - np.array(pos, f32)
            proj = f32(f32(d[0] * fwd[0] + d[1] * fwd[1]) + f32(d[2] * fwd[2]))
            if proj > f32(1e-4):
                focus = float(proj)
    aspect = float(np.float32(width) / np.float32(height))
    return scenes.make_camera(pos, look, _f(cd, "fov", 60.0), aspect, vup=vup, aperture=_f(cd, "aperture", 0.0),
                              focus_dist=focus)


def load_data(data, width, height, shadow_samples_override=None, camera=None, base_dir=None, warnings=None):
    """SceneLoader.Load on an already parsed YAML document (a dict); mesh
    paths resolve against base_dir, and the loader's warnings (skipped
    meshes, malformed OBJ lines) are appended to `warnings`."""
    if not isinstance(data, dict):
        raise SceneLoadError("not a scene document")
    warnings = [] if warnings is None else warnings
    if data.get("imports"):
        data = dict(data)
        _merge_imports(data, base_dir or ".", set(), warnings)
    # named media (SceneLoader.cs:235-256) only act through an entity's
    # interior / exterior binding, which is refused per entity
    b = scenes.SceneBuilder()
    b.shadow_samples_override = shadow_samples_override
    b.geometry_shadow_samples = GEOMETRY_SHADOW_SAMPLES
    world = data.get("world") or {}

    materials = {}
    for m in data.get("materials") or []:
        if m.get("id") is None:
            continue
        materials[str(m["id"])] = _material(b, m, base_dir, warnings)

    def get_material(mid):
        if mid is not None and str(mid) in materials:
            return materials[str(mid)]
        return b.lambertian((0.5, 0.5, 0.5))

    sky = DEFAULT_SKY
    sd = world.get("sky")
    if sd is not None:
        st = (sd.get("type") or "").lower()
        if st not in ("flat", ""):
            raise SceneLoadError(f"sky type '{st}' is not supported (flat only)")
        if sd.get("background") is not None:
            raise SceneLoadError("a separate sky background is not supported")
        sky = _v3(sd.get("color"), DEFAULT_SKY)

    medium = None
    md = world.get("medium")
    if md is not None:
        mt = (md.get("type") or "homogeneous").lower()
        if mt != "homogeneous":
            raise SceneLoadError(f"medium type '{mt}' is not supported (homogeneous only)")
        ph = (md.get("phase") or "isotropic").lower()
        if ph != "isotropic":
            raise SceneLoadError(f"medium phase '{ph}' is not supported (isotropic only)")
        sa = tuple(max(x, 0.0) for x in _v3(md.get("sigma_a"), (0.0, 0.0, 0.0)))
        ss = tuple(max(x, 0.0) for x in _v3(md.get("sigma_s"), (0.0, 0.0, 0.0)))
        medium = (sa, ss)

    # objects: the ground first
    gd = world.get("ground")
    if gd is not None:
        gt = (gd.get("type") or "infinite_plane").strip().lower()
        if gt not in ("infinite_plane", "plane", ""):
            raise SceneLoadError(f"ground type '{gt}' is not supported (infinite plane only)")
        if gd.get("visibility") is not None:
            raise SceneLoadError("ground visibility masks are not supported")
        # UvTransformedHittable (SceneLoader.cs:1360-1367) is not built: a
        # ground UV transform other than the identity is refused
        for k, ident_v in (("uv_scale", (1.0, 1.0)), ("uv_offset", (0.0, 0.0)), ("uv_rotation", 0.0)):
            val = gd.get(k)
            if val is None:
                continue
            vals = [float(x) for x in val] if isinstance(val, (list, tuple)) else [float(val)] * (
                2 if isinstance(ident_v, tuple) else 1)
            want = list(ident_v) if isinstance(ident_v, tuple) else [ident_v]
            if vals[:len(want)] != want:
                raise SceneLoadError(f"ground '{k}' (a UV transform of the ground) is not supported")
        if gd.get("material"):
            gm = get_material(gd["material"])
        elif gd.get("color") is not None or gd.get("roughness") is not None or gd.get("metallic") is not None:
            gm = b.disney(_v3(gd.get("color"), (0.5, 0.5, 0.5)), metallic=_f(gd, "metallic", 0.0),
                          roughness=_f(gd, "roughness", 0.8))
        else:
            sdd = world.get("sky") or {}
            alb = _v3(sdd.get("ground_albedo")) or _v3(sdd.get("ground_color"))
            gm = b.lambertian(alb if alb is not None else (0.5, 0.5, 0.5))
        anchor = _v3(gd.get("point"), (0.0, float(np.float32(gd.get("y", 0.0) or 0.0)), 0.0))
        n = np.array(_v3(gd.get("normal"), (0.0, 1.0, 0.0)), np.float32)
        if float(n @ n) < 1e-12:
            n = np.array([0.0, 1.0, 0.0], np.float32)
        n = n / np.float32(np.sqrt(np.float32(n @ n)))
        b.plane(anchor, tuple(float(x) for x in n), gm)

    templates = {}
    for t in data.get("templates") or []:
        if not t.get("name") or not str(t["name"]).strip():
            warnings.append("Template without a 'name' field. Skipping.")
            continue
        if not t.get("children"):
            warnings.append(f"Template '{t['name']}' has no 'children'. Skipping.")
            continue
        templates[str(t["name"]).lower()] = t          # last one wins
    has_images = any(int(r["flags"]) & abi.RT_MATF_IMAGE for r in (b.material_record(m) for m in range(b.material_count)))
    ctx = _Ctx(not has_images, templates, get_material, base_dir, warnings)
    for e in data.get("entities") or []:
        _entity(b, e, get_material(e.get("material")), base_dir, warnings, ctx)

    # lights (proxies appended after the entities)
    lights_key = "lights" in data
    for L in data.get("lights") or []:
        lt = (L.get("type") or "").lower()
        color = _v3(L.get("color"), (1.0, 1.0, 1.0))
        inten = _f(L, "intensity", 1.0)
        ss = int(L.get("shadow_samples", 4) or 0)
        vis = bool(L.get("visible_to_camera", True))
        if lt in ("directional", "sun"):
            b.directional_light(_v3(L.get("direction"), (-1.0, -1.0, -1.0)), color, inten,
                                angular_radius_deg=_f(L, "angular_radius", 0.0))
        elif lt in ("area", "area_light", "rect", "rect_light"):
            corner, u, v = _v3(L.get("corner")), _v3(L.get("u")), _v3(L.get("v"))
            if corner is None or u is None or v is None:
                continue   # the reference warns and skips the light
            b.area_light(corner, u, v, color, inten, shadow_samples=ss, soft_radius=_f(L, "soft_radius", 0.0),
                         visible_to_camera=vis)
        elif lt in ("sphere", "sphere_light", "ball", "ball_light"):
            b.sphere_light(_v3(L.get("position"), (0.0, 10.0, 0.0)), _f(L, "radius", 0.5), color, inten,
                           shadow_samples=ss, visible_to_camera=vis)
        elif lt in ("point", "spot", "spotlight", "portal", "portal_light"):
            raise SceneLoadError(f"light type '{lt}' is not supported")
        # other types: CreateLight returns null and the light is skipped
    if not lights_key:
        raise SceneLoadError("a scene without a 'lights' section gets the reference's default point light, "
                             "which is not supported")
    cam = _camera(data, camera, width, height)
    return b.build(cam, sky=sky, medium=medium)


def load(path, width, height, shadow_samples_override=None, camera=None, warnings=None):
    """SceneLoader.Load(yamlPath, imageWidth, imageHeight, shadowSamplesOverride,
    cameraSelector): a scenes.SceneDesc for the renderer."""
    import yaml
    with open(path, "r", encoding="utf-8") as f:
        data = yaml.safe_load(f)
    return load_data(data, width, height, shadow_samples_override, camera,
                     base_dir=os.path.dirname(os.path.abspath(path)), warnings=warnings)
