/*
 * rt3d.h — C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * This is the drop-in boundary for the render entry point of the C# path
 * tracer fiorenzobrioni/3d-ray.  The reference exposes no FFI; its render
 * entry point is the pair
 *
 *   new Renderer(world, camera, lights, sky, samplesPerPixel, maxDepth,
 *                globalMedium, maxSampleRadiance, verbose, misHeuristic,
 *                lightSamplingStrategy, indirectClampFactor, ...)
 *       -- src/RayTracer/Rendering/Renderer.cs:300-319
 *   RenderResult Render(int width, int height, RenderCaptureOptions options)
 *       -- src/RayTracer/Rendering/Renderer.cs:522-733
 *
 * and the world it receives is assembled by SceneLoader.Load
 * (src/RayTracer/Scene/SceneLoader.cs:122-553).  A host (the C# P/Invoke shim,
 * the Python ctypes harness, or C++) flattens that object graph into the POD
 * tables below, in the SceneLoader's object order, and calls this library.
 * The library builds the binned-SAH BVH on the host (restating
 * src/RayTracer/Acceleration/BvhNode.cs:91-363), uploads everything to HBM
 * once per scene, and renders with hand-written HIP kernels.
 *
 * Conventions
 *   - every entry point returns an rt_status (0 = OK, < 0 = error); on error
 *     rt_last_error(ctx) holds a message until the next call on that context;
 *   - the caller owns all host memory; tables are copied during the call and
 *     never retained; the library owns device memory inside rt_scene/rt_ctx;
 *   - calls on one rt_ctx must be serialised by the caller; distinct contexts
 *     are independent; there is no process-global render state (the
 *     reference's Sampler.SetKind and ShadowRay.BlockSpecularTransmission
 *     statics are per-call settings here);
 *   - all structs are plain C with fixed sizes (checked by static asserts in
 *     the implementation) so C# [StructLayout(Sequential)] and ctypes map 1:1.
 */
#ifndef RT3D_H
#define RT3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT3D_ABI_VERSION 1

typedef enum rt_status {
    RT_OK = 0,
    RT_ERR_INVALID_ARGUMENT = -1,
    RT_ERR_NO_DEVICE = -2,
    RT_ERR_DEVICE = -3,
    RT_ERR_OUT_OF_MEMORY = -4,
    RT_ERR_UNSUPPORTED = -5,
    RT_ERR_INTERNAL = -6
} rt_status;

/* ── Primitives ───────────────────────────────────────────────────────────
 * One record per intersectable primitive, in SceneLoader object order
 * (ground, YAML entities, then light proxies appended by CreateLight,
 * SceneLoader.cs:282-457).  The index of a record is its primitive ID — the
 * ID the bit-exact primary-hit parity gate compares.
 *
 *   TRIANGLE  v[0..8]  = v0, v1, v2           (Triangle.cs:35-89)
 *   SPHERE    v[0..3]  = centre, radius        (Sphere.cs:13-75)
 *   QUAD      v[0..8]  = Q, U, V               (Quad.cs:302-362)
 *   PLANE     v[0..5]  = point, normal         (InfinitePlane.cs:28)
 *
 * mesh = -1 for a standalone entity.  mesh = k >= 0 marks the faces of one
 * Mesh object (Mesh.cs:75-116, a nested BVH); its faces must be contiguous
 * and in face order, and the mesh takes the object-list position of its
 * first face.  Only TRIANGLE faces may belong to a mesh.
 */
typedef enum rt_prim_kind {
    RT_PRIM_TRIANGLE = 0,
    RT_PRIM_SPHERE = 1,
    RT_PRIM_QUAD = 2,
    RT_PRIM_PLANE = 3
} rt_prim_kind;

typedef enum rt_prim_flags {
    RT_PRIMF_BACKFACE_CULLED = 1,   /* BackFaceCulledHittable (area-light proxy) */
    RT_PRIMF_CAMERA_INVISIBLE = 2   /* CameraInvisibleHittable */
} rt_prim_flags;

typedef struct rt_primitive {
    int32_t kind;       /* rt_prim_kind */
    int32_t material;   /* index into rt_scene_desc.materials */
    int32_t flags;      /* rt_prim_flags */
    int32_t mesh;       /* -1 or mesh object index */
    float v[12];
} rt_primitive;         /* 64 bytes */

/* ── Materials ─────────────────────────────────────────────────────────────
 *   LAMBERTIAN  color = albedo                         (Lambertian.cs:37-96)
 *   EMISSIVE    color, intensity                       (Emissive.cs:56-104)
 *   DIELECTRIC  color = albedo, ior                    (Dielectric.cs:46-105)
 *   DISNEY      color = baseColor + every Disney knob  (DisneyBsdf.cs:213-293)
 * rt_material_defaults() fills the DisneyBsdf constructor defaults.
 */
typedef enum rt_material_kind {
    RT_MAT_LAMBERTIAN = 0,
    RT_MAT_EMISSIVE = 1,
    RT_MAT_DIELECTRIC = 2,
    RT_MAT_DISNEY = 3
} rt_material_kind;

typedef enum rt_material_flags {
    RT_MATF_LIGHT_PROXY = 1,          /* Emissive.IsLightProxy */
    RT_MATF_THIN_WALLED = 2,          /* DisneyBsdf.ThinWalled */
    RT_MATF_TRANSMISSION_COLOR = 4,   /* DisneyBsdf.TransmissionColor != null */
    RT_MATF_COAT_ROUGHNESS = 8,       /* DisneyBsdf.CoatRoughness != null */
    RT_MATF_TEXTURED = 16,            /* DISNEY only: baseColor, roughness and metallic
                                         come from the scene's texture set at the
                                         hit's texture coordinate (no reference
                                         counterpart; spec_trans and diff_trans
                                         must be 0) */
    RT_MATF_IMAGE = 32                /* LAMBERTIAN albedo / DISNEY baseColor is the
                                         ImageTexture rt_material.image of the scene's
                                         image table (SceneLoader.cs:1027-1031,
                                         ImageTexture.cs:42-289); DISNEY needs
                                         spec_trans = 0 */
} rt_material_flags;

typedef struct rt_material {
    int32_t kind;
    int32_t flags;
    float color[3];
    float intensity;
    float metallic, roughness, specular, specular_tint;
    float sheen, sheen_tint, sheen_roughness, clearcoat;
    float clearcoat_gloss, spec_trans, ior, anisotropic;
    float anisotropic_rotation, diff_trans, coat_ior, coat_roughness;
    float thin_film_thickness, thin_film_ior, transmission_depth;
    float transmission_color[3];
    int32_t image;      /* RT_MATF_IMAGE: index into rt_scene_desc.images */
    int32_t reserved[3];
} rt_material;          /* 128 bytes */

/* ── Lights (final ILight list order of SceneLoader.Load) ─────────────────
 *   SPHERE       position = centre, radius, color, intensity, shadow_samples,
 *                primitive/material = visible proxy  (SphereLight.cs:112-271)
 *   AREA         position = corner, u, v, color, intensity, shadow_samples,
 *                soft_radius, primitive/material = proxy (AreaLight.cs:97-226)
 *   DIRECTIONAL  position = direction, color, intensity, radius = angular
 *                radius in degrees, shadow_samples (0 = reference default)
 *                                                   (DirectionalLight.cs:102-181)
 *   GEOMETRY     primitive = emissive TRIANGLE/QUAD/SPHERE, material = its
 *                Emissive, shadow_samples, soft_radius (GeometryLight.cs:133-279)
 *   ENVIRONMENT  the EnvironmentLight over the scene's flat sky
 *                (rt_scene_desc.sky_color; the record's colour is not read),
 *                shadow_samples (EnvironmentLight.cs:20-124).  SceneLoader
 *                appends it after the geometry lights when the sky can be
 *                sampled (SceneLoader.cs:524-528); the first one in the list
 *                also weighs BSDF-sampled sky hits by MIS (Renderer.cs:426,
 *                1820-1861).  Without one the sky is seen at full weight and
 *                never sampled by NEE, as the reference's Renderer does.
 */
typedef enum rt_light_kind {
    RT_LIGHT_SPHERE = 0,
    RT_LIGHT_AREA = 1,
    RT_LIGHT_DIRECTIONAL = 2,
    RT_LIGHT_GEOMETRY = 3,
    RT_LIGHT_ENVIRONMENT = 4
} rt_light_kind;

typedef struct rt_light {
    int32_t kind;
    int32_t shadow_samples;
    int32_t primitive;
    int32_t material;
    float color[3];
    float intensity;
    float position[3];
    float radius;
    float u[3];
    float soft_radius;
    float v[3];
    float reserved;
} rt_light;             /* 80 bytes */

/* Look-at thin-lens camera (Camera.cs:26-106).  aspect = width/height as the
 * SceneLoader computes it (SceneLoader.cs:491). */
typedef struct rt_camera {
    float look_from[3];
    float look_at[3];
    float vup[3];
    float vfov_deg;
    float aspect;
    float aperture;
    float focus_dist;
    float reserved;
} rt_camera;            /* 56 bytes */

/* Global homogeneous medium with isotropic phase (HomogeneousMedium.cs:36-76). */
typedef struct rt_medium {
    int32_t enabled;
    float sigma_a[3];
    float sigma_s[3];
    float reserved;
} rt_medium;            /* 32 bytes */

typedef struct rt_scene_desc {
    const rt_primitive* primitives;
    int64_t num_primitives;
    const rt_material* materials;
    int64_t num_materials;
    const rt_light* lights;
    int64_t num_lights;
    rt_camera camera;
    float sky_color[3];   /* flat sky radiance (FlatSky.cs): seen by escaping rays
                             (SampleSky, Renderer.cs:1820-1861), sampled by an
                             RT_LIGHT_ENVIRONMENT light when the list holds one */
    int32_t reserved0;
    rt_medium global_medium;
    /* Texture set of RT_MATF_TEXTURED materials (no reference counterpart:
     * the reference's Disney parameters are constant FloatTexture/SolidColor
     * values, DisneyBsdf.cs:55-73).  texels: planar fp32
     * [5][tex_height][tex_width] = base colour r, g, b, roughness, metallic;
     * bilinear, texel centres at (i + 0.5) / width, u wraps, v clamps.
     * texcoords (optional): 6 floats per primitive, the (u, v) of a
     * TRIANGLE's three vertices, interpolated as SmoothTriangle does
     * (SmoothTriangle.cs:186: (1-b1-b2) uv0 + b1 uv1 + b2 uv2); without them a
     * hit's texture coordinate is HitRecord.U/V (barycentrics, quad
     * (alpha, beta), Sphere.GetSphereUV, the plane's tiled UV). */
    const float* texcoords;
    const float* texels;
    int32_t tex_width;
    int32_t tex_height;
    /* Vertex normals (optional): 9 floats per primitive, N0, N1, N2 of a
     * TRIANGLE.  A triangle whose triple is not all zero is a SmoothTriangle
     * (SmoothTriangle.cs:68-197): normals normalised once, shading normal
     * normalize(w0 N0 + w1 N1 + w2 N2), front face from the face normal,
     * tangent frame from the UV gradients of its texcoords (barycentric UVs
     * without them).  Such a triangle may not be a GEOMETRY light. */
    const float* normals;
    /* Image textures of RT_MATF_IMAGE materials (ImageTexture.cs:42-289). */
    const struct rt_image* images;
    int64_t num_images;
    /* Object hierarchy (optional; rt_node below).  num_nodes == 0: the world
     * list is the flat object list (each primitive, or each run of
     * consecutive primitives sharing a mesh id, is one object, in order).
     * num_nodes > 0: the world list is nodes[roots[0 .. num_roots)], in that
     * order (SceneLoader.cs:460-487 assembles it into the world BVH plus the
     * infinite planes); node_children holds the child lists of GROUP nodes. */
    const struct rt_node* nodes;
    int64_t num_nodes;
    const int32_t* node_children;
    int64_t num_node_children;
    const int32_t* roots;
    int64_t num_roots;
} rt_scene_desc;

/* ── Object hierarchy: the reference's IHittable wrappers ─────────────────
 * A node is one IHittable of the loaded scene.  Nodes may be shared: an
 * INSTANCE's template (SceneLoader.cs:3259-3337 builds a template once and
 * wraps it in one Instance per use) is one node referenced by every
 * instance, so its BVHs and primitives are resident once.
 *   PRIMITIVES  primitive `first` when count == 1 and its mesh id < 0, else
 *               the `count` faces [first, first + count) of one mesh (a Mesh:
 *               nested BVH over > 4 faces, a list otherwise, Mesh.cs:75-116);
 *   GROUP       Group.cs:20-60: children node_children[first .. first+count);
 *               a BVH over the finite children when more than 4, the
 *               infinite planes after it;
 *   TRANSFORM   Transform.cs:74-129: `child` hit in object space — the ray
 *               mapped by `inverse`, the hit point by `matrix`, normals by
 *               the transpose of `inverse` (normalised), tangents and DpDu /
 *               DpDv by `matrix` (MapHitToWorld, :162-207); its box is the
 *               child's box with its 8 corners mapped by `matrix`;
 *   INSTANCE    Instance.cs:54-82: `child` unchanged, with every hit's
 *               material replaced by `material` when it is >= 0.
 * Matrices use System.Numerics.Matrix4x4's field order M11, M12, ... M44
 * and its row-vector convention: a point maps to
 * (x M11 + y M21 + z M31 + M41, ...), evaluated as Vector3.Transform does
 * (fma(x, row1, fma(y, row2, fma(z, row3, row4)))); a direction drops row4.
 * `inverse` is the host's Matrix4x4.Invert of `matrix` (passed, not
 * recomputed, so the ray transform uses the host's inverse bit for bit).
 * RT_NODEF_CAMERA_INVISIBLE wraps the node in CameraInvisibleHittable.
 * Limits: wrapper nesting (TRANSFORM / INSTANCE / camera-invisible nodes on
 * one root-to-primitive path) at most 4; a GROUP's list part at most 15
 * members; no GEOMETRY light, caustics, rt_adjoint or image texture in a
 * scene with wrappers (RT_ERR_UNSUPPORTED). */
typedef enum rt_node_kind {
    RT_NODE_PRIMITIVES = 0,
    RT_NODE_GROUP = 1,
    RT_NODE_TRANSFORM = 2,
    RT_NODE_INSTANCE = 3
} rt_node_kind;
#define RT_NODEF_CAMERA_INVISIBLE 1

typedef struct rt_node {
    int32_t kind;           /* rt_node_kind */
    int32_t flags;          /* RT_NODEF_* */
    int32_t first;          /* PRIMITIVES: first primitive; GROUP: first index into node_children */
    int32_t count;          /* PRIMITIVES: primitive count; GROUP: child count */
    int32_t child;          /* TRANSFORM / INSTANCE: the wrapped node */
    int32_t material;       /* INSTANCE: override material, -1 = the template's own materials */
    int32_t reserved[2];
    float matrix[16];       /* TRANSFORM: object -> world, M11 .. M44 */
    float inverse[16];      /* TRANSFORM: Matrix4x4.Invert(matrix) as the host computed it */
} rt_node;              /* 160 bytes */

/* One ImageTexture: the decoded 8-bit image (ImageSharp's Rgba32 rows, top
 * row first; alpha ignored) and the YAML uv_scale.  The library decodes
 * sRGB with ImageTexture's gamma-2.2 table, builds the 2x2 box mip chain
 * and filters as ImageTexture.Value does (bilinear at level 0 without a
 * footprint; ratio-trilinear with at most 8 taps from the camera ray's
 * differentials, FootprintMath.cs:38-119). */
typedef struct rt_image {
    const uint8_t* rgba8;   /* width * height * 4 bytes */
    int32_t width;
    int32_t height;
    float scale_u;
    float scale_v;
} rt_image;             /* 24 bytes */

/* ── Render settings (Renderer ctor arguments + Render(width, height)) ───── */
typedef enum rt_sampler_kind { RT_SAMPLER_SOBOL = 0, RT_SAMPLER_PRNG = 1 } rt_sampler_kind;
typedef enum rt_mis_kind { RT_MIS_BALANCE = 0, RT_MIS_POWER = 1 } rt_mis_kind;
typedef enum rt_light_sampling { RT_LIGHTS_ALL = 0, RT_LIGHTS_POWER = 1, RT_LIGHTS_UNIFORM = 2 } rt_light_sampling;

typedef struct rt_settings {
    int32_t width;
    int32_t height;
    int32_t samples_per_pixel;
    int32_t max_depth;
    int32_t sampler;                /* rt_sampler_kind; PRNG: ceil(sqrt(spp))^2 stratified
                                       samples with hashed draws (Renderer.cs:556-558,
                                       677-682); a chunk of a PRNG render takes exactly
                                       spp samples, strata of the whole render's
                                       sample_total = q * q grid (sample s in
                                       stratum (s mod q, s div q)) */
    int32_t mis;                    /* rt_mis_kind */
    int32_t light_sampling;         /* rt_light_sampling */
    int32_t capture_aovs;           /* RenderCaptureOptions.CaptureAovs */
    float max_sample_radiance;      /* --clamp, reference default 10 */
    float indirect_clamp_factor;    /* --indirect-clamp-factor, default 0.25 */
    float exposure_ev;              /* --exposure, default 0 */
    int32_t rr_min_bounces;         /* < 0: scene classifier (Renderer.cs:361-378) */
    float rr_min_survival;          /* < 0: scene classifier */
    uint32_t seed_xor;              /* 0 = the reference's Sobol streams; any other
                                       value XORs every pixel seed (independent
                                       streams for statistical tests) */
    int32_t caustics;               /* --caustics: photon-map caustics (Renderer.cs:380-406) */
    int32_t caustic_photons;        /* --caustic-photons: emitted photon budget */
    float caustic_gather_radius;    /* gather-radius cap; 0 = 6 % of the scene radius */
    int32_t paths_in_flight;        /* path slots in flight; 0 = library default (2^25) */
    int32_t count_traversal;        /* 1: count BVH node/primitive tests (slower
                                       instrumented kernels; fills rt_render_stats) */
    int32_t texture_filtering;      /* --texture-filtering: 0 auto (= on), 1 on, 2 off:
                                       camera-ray differentials for image textures
                                       (Renderer.cs:344-348) */
    int32_t sample_offset;          /* first sample index: the call renders samples
                                       [sample_offset, sample_offset + spp) of every
                                       pixel (Sobol index, A/B half by the index's
                                       parity); 0 = the reference's render.  Chunks
                                       [0, n1) and [n1, n1 + n2) merge into the
                                       [0, n1 + n2) render by sample-weighted means
                                       (resume / chunking; no reference counterpart) */
    int32_t sample_total;           /* the whole render's spp when this call is one
                                       chunk of it (0 = samples_per_pixel): sets the
                                       ray-differential pixel footprint, which shrinks
                                       by ceil(sqrt(spp)) of the render (Renderer.cs:556);
                                       with the PRNG sampler a chunk needs it, a square */
    int32_t reserved[2];
} rt_settings;          /* 96 bytes */

/* ── Film (RenderBuffers.cs:52-169; FrameBuffer planar layout
 *    Data[c*W*H + y*W + x], FrameBuffer.cs:20-41).  Caller-allocated host
 *    arrays; any NULL plane is not written.  Colour planes hold 3*W*H floats,
 *    depth planes W*H floats, sample_count W*H ints.  display is the
 *    tone-mapped RenderResult.Pixels (exposure, ACES, gamma 2.2). */
typedef struct rt_film {
    float* beauty;
    float* beauty_a;
    float* beauty_b;
    float* albedo_a;
    float* albedo_b;
    float* normal_a;
    float* normal_b;
    float* depth_a;
    float* depth_b;
    int32_t* sample_count;
    float* display;
} rt_film;

/* Per-render counters and timings filled by rt_render (optional). */
typedef struct rt_render_stats {
    double wall_ms;              /* whole rt_render call */
    double device_ms;            /* HIP-event time of all render kernels */
    double trace_ms;             /* HIP-event time of the closest-hit kernels */
    double shadow_ms;            /* HIP-event time of the shadow kernels */
    double shade_ms;             /* HIP-event time of the shading kernels */
    int64_t camera_samples;      /* width*height*spp */
    int64_t rays_closest;        /* primary + extension segments traced */
    int64_t rays_shadow;         /* shadow segments (each transparent-shadow
                                    continuation counts once) */
    int64_t trace_launches;      /* closest-hit kernel launches (bounce iterations
                                    enqueued, trailing empty ones included) */
    int64_t shadow_launches;     /* shadow kernel launches (same count) */
    /* with rt_settings.count_traversal: the reference traversal's work —
       BvhNode.Hit calls (node box tests) and primitive Hit calls — of the
       closest-hit and shadow kernels */
    int64_t trace_node_tests, trace_tri_tests, trace_other_tests;
    int64_t shadow_node_tests, shadow_tri_tests, shadow_other_tests;
    /* with count_traversal: the work the shadow kernel's own any-hit
       traversal does on the 4-wide BVH (128-B node fetches, triangle/quad/
       plane tests, sphere tests); zero when it uses the binary tree */
    int64_t any_node_fetches, any_tri_tests, any_other_tests;
} rt_render_stats;

typedef struct rt_scene_info {
    int64_t num_primitives;
    int64_t num_objects;         /* top-level objects (meshes count once) */
    int64_t num_nodes;           /* internal BVH nodes, all levels */
    int64_t num_leaves;
    int32_t max_depth;           /* deepest root-to-leaf path */
    int32_t has_bvh;             /* 0 when the world is a flat list (<= 4 objects) */
    int32_t rr_min_bounces;      /* resolved by the scene classifier */
    float rr_min_survival;
    float scene_radius;
    float mean_irradiance;
    int64_t device_bytes;        /* scene bytes resident in HBM */
    int64_t reserved[4];
} rt_scene_info;

typedef struct rt_ctx rt_ctx;
typedef struct rt_scene rt_scene;

/* Library ABI version (RT3D_ABI_VERSION). */
int rt_get_version(void);

/* Fills the DisneyBsdf constructor defaults (DisneyBsdf.cs:167-212) and the
 * reference's per-kind defaults into *m for the given kind. */
void rt_material_defaults(rt_material* m, int32_t kind);

/* Fills the reference's CLI/ctor defaults (Renderer.cs:300-319,
 * Program.cs Appendix-A canonical flags). */
void rt_settings_defaults(rt_settings* s);

/* Creates a context on HIP device `device` (one HIP stream, default priority).
 * Replaces: the implicit process-wide state of Renderer.Render's Parallel.For. */
int rt_create(int32_t device, rt_ctx** out);
void rt_destroy(rt_ctx* ctx);
const char* rt_last_error(const rt_ctx* ctx);

/* Builds the world (SceneLoader.cs:460-487: BVH over finite objects when there
 * are more than 4, infinite planes kept beside it), the binned-SAH BVH
 * (BvhNode.cs:91-363), the light tables and the Kulla–Conty energy LUT
 * (EnergyCompensationLut.cs:34-171), and uploads them.
 * Replaces: SceneLoader.Load's world assembly + the Renderer constructor's
 * scene analysis (Renderer.cs:321-441). */
int rt_scene_create(rt_ctx* ctx, const rt_scene_desc* desc, rt_scene** out);
void rt_scene_destroy(rt_scene* scene);
int rt_scene_get_info(const rt_scene* scene, rt_scene_info* info);

/* Replaces the texture set of a scene with textured materials (same layout
 * as rt_scene_desc.texels; the size may change).  No reference counterpart:
 * the optimisation loop of the adjoint updates textures between renders
 * without rebuilding the BVH. */
int rt_scene_set_textures(rt_ctx* ctx, rt_scene* scene, int32_t width, int32_t height, const float* texels);

/* Renders settings->samples_per_pixel samples per pixel into the film.
 * Replaces: Renderer.Render(width, height, options) (Renderer.cs:522-733). */
int rt_render(rt_ctx* ctx, const rt_scene* scene, const rt_settings* settings,
              const rt_film* film, rt_render_stats* stats);

/* Path-replay adjoint of the texture set (no reference counterpart).
 * Given dLoss/dBeauty (d_beauty: planar [3][height][width], host memory) of
 * a loss on the beauty film rendered by rt_render with the same scene and
 * settings, writes dLoss/dTexels (d_texels: planar [5][tex_height][tex_width],
 * host memory; overwritten).  Each window of samples is traced twice with
 * the same sampler streams: a primal pass storing every sample's radiance,
 * then a replay pass adding, at textured Disney hits, the derivative of the
 * NEE BSDF terms and of the sampled-bounce weight (detached sampling: pdfs,
 * MIS weights, lobe choice and Russian roulette held fixed; delta and
 * fallback bounces add no term).  Requires: no medium, no transmissive
 * material, no caustics, inactive clamps (max_sample_radiance *
 * indirect_clamp_factor >= 1e30); else RT_ERR_UNSUPPORTED.  Texel gradients
 * are summed with float atomics (order not fixed). */
int rt_adjoint(rt_ctx* ctx, const rt_scene* scene, const rt_settings* settings, const float* d_beauty,
               float* d_texels, rt_render_stats* stats);

/* Kernel configuration of a context (no reference counterpart; set by an
 * autotuner): shade_waves_per_simd = minimum waves per SIMD the NEE shading
 * kernel is compiled for (2 or 3; 1 runs as 2 and 4 as 3), traversal_stack = 1 private scratch, 2 LDS,
 * bounce_graph = 1 direct launches, 2 replay of hipGraph-captured bounce
 * iterations (no per-kernel timings then); 0 keeps a setting.  Results do
 * not depend on any of them. */
int rt_context_tune(rt_ctx* ctx, int32_t shade_waves_per_simd, int32_t traversal_stack, int32_t bounce_graph);

/* ── Batched rendering (no reference counterpart: the reference renders one
 *    scene per process run, Program.cs:474-490) ─────────────────────────── */
typedef struct rt_batch_job {
    const rt_scene* scene;
    const rt_settings* settings;   /* caustics are not supported in a batch */
    const rt_film* film;           /* NULL: the film stays on the device */
    rt_render_stats* stats;        /* optional: per-job camera_samples */
} rt_batch_job;

/* Renders n_jobs (scene, settings) pairs through one path pool: the camera
 * samples of the small jobs (fewer than two pools of samples) stream through
 * the same device-driven bounce loop together, every path carrying its job,
 * so they fill the GPU together; a larger job runs alone in the pool with
 * the single-scene kernels, in job order, without a host round trip.  Each job's film is identical to rt_render of that job (same
 * per-sample values, same summation order).  pool_paths = path slots of the
 * pool (0 = the batch default, 2^24).  total (optional) receives summed counts and
 * the batch's timings. */
int rt_render_batch(rt_ctx* ctx, const rt_batch_job* jobs, int32_t n_jobs, int32_t pool_paths,
                    rt_render_stats* total);

/* Parity hook: for samples [sample_begin, sample_begin+sample_count) of every
 * pixel, the camera ray of Renderer.cs:630-653 (+ the camera-visibility skip
 * of TraceRay, Renderer.cs:1609-1644) is traced and its closest hit written as
 * prim_id (-1 = miss), t, u, v.  Output index = s_local*W*H + y*W + x.
 * NULL outputs are skipped; settings.sample_offset is not applied (the
 * sample indices are sample_begin's).  No reference counterpart (HitRecord
 * carries no primitive index); this is the bit-exact primary-hit gate. */
int rt_trace_primary(rt_ctx* ctx, const rt_scene* scene, const rt_settings* settings,
                     int32_t sample_begin, int32_t sample_count,
                     int32_t* prim_id, float* t, float* u, float* v);

/* Occlusion of n segments (rays: n x 8 floats = origin, direction, tmin,
 * tmax): occluded[i] = 1 when any primitive is hit with tmin <= t <= tmax.
 * The any-hit query the shadow kernel runs for opaque scenes (a 4-wide BVH
 * collapsed from the binary one); the answer equals the existence of a
 * closest hit on [tmin, tmax].
 * Replaces: ShadowRay.Transmittance's occlusion test for opaque materials
 * (src/RayTracer/Rendering/ShadowRay.cs:59-158, BvhNode.Hit with the segment). */
int rt_occluded(rt_ctx* ctx, const rt_scene* scene, const float* rays, int64_t num_rays, int32_t* occluded);

/* rt_occluded without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and answers as rt_occluded; the segments and results are
 * staged in the library's static per-GPU array (about 36 MiB, held until the
 * process ends), so at most 1,048,576 segments per call (more is
 * RT_ERR_INVALID_ARGUMENT).  Calls that overlap on one GPU share that array
 * and race: serialise them. */
int rt_occluded_noalloc(rt_ctx* ctx, const rt_scene* scene, const float* rays, int64_t num_rays, int32_t* occluded);

/* Closest hit of arbitrary rays (origin xyz, direction xyz per ray, 6 floats)
 * against the scene, tMin = 1e-4, tMax = float.MaxValue, exactly as
 * world.Hit(ray, MathUtils.Epsilon, MathUtils.Infinity) (Renderer.cs:1629).
 * Replaces: IHittable.Hit on the assembled world (HittableList.cs:19-38). */
int rt_trace_rays(rt_ctx* ctx, const rt_scene* scene, const float* rays, int64_t num_rays,
                  int32_t* prim_id, float* t, float* u, float* v);

/* Parity hook: runs the caustic photon pre-pass rt_render runs for
 * `settings` (caustic_photons; caustic_gather_radius, or 6 % of the scene
 * radius when 0, Renderer.cs:394; grid cell = half of it) and copies the
 * device photon map out.  Counts are always written; the arrays (NULL =
 * skip) are filled when cap_photons >= n_photons (pos/inc/power 3 floats
 * and key/idx 1 int per photon) and cap_buckets >= n_buckets + 1 (start) —
 * call once with zero caps to size them.  key = bucket of each photon in
 * photon order, idx/start = the CSR grid, grid = origin xyz and cell size.
 * Photons are in emission order (global photon index).  No reference
 * counterpart (PhotonMap keeps these private, PhotonMap.cs:27-35); it lets
 * a test check the device grid bit-exactly against a host sort. */
int rt_photon_map(rt_ctx* ctx, const rt_scene* scene, const rt_settings* settings, int64_t cap_photons,
                  int64_t cap_buckets, int64_t* n_photons, int64_t* n_emitted, int64_t* n_buckets, float* pos,
                  float* inc, float* power, int32_t* key, int32_t* idx, int32_t* start, float* grid);

/* Parity hook: builds the photon map rt_photon_map builds for `settings`
 * and runs the render's k-nearest search on the device at each of n_points
 * points (3 floats each), with k = 40 (Renderer.CausticGatherK) and the
 * gather radius as the radius cap.  count[i] = photons found, radius[i] =
 * distance of the farthest (0 when none), photon_idx[40 * i + j] = photon j
 * of point i in the order the search's max-heap holds them (the order the
 * render sums them in), as a photon-order index (rt_photon_map's order);
 * -1 past count[i].
 * Replaces: PhotonMap.GatherKNearest's result list and radius
 * (PhotonMap.cs:209-240, heap rules :296-317). */
int rt_photon_gather(rt_ctx* ctx, const rt_scene* scene, const rt_settings* settings, const float* points,
                     int64_t n_points, int32_t* count, float* radius, int32_t* photon_idx);

/* ── Film files.  Each encoder writes a whole file image into `out` (the
 * caller writes it to disk); with out = NULL it only stores in *size an upper
 * bound of the bytes needed, and a capacity below that bound is
 * RT_ERR_INVALID_ARGUMENT.  On success *size = bytes written.  Planes are
 * plane-major host floats, row 0 = top (FrameBuffer.cs:20-41); they are
 * converted on the device and compressed on host threads. */

/* EXR channel (ExrImage.Channel, ExrImage.cs:45-46; ExrPixelType, :8-12). */
#define RT_EXR_HALF 1
#define RT_EXR_FLOAT 2
typedef struct rt_exr_channel {
    const char* name;      /* 1..31 printable ASCII characters, unique */
    int32_t type;          /* RT_EXR_HALF or RT_EXR_FLOAT */
    int32_t pad;
    const float* plane;    /* width*height floats */
} rt_exr_channel;

/* Replaces ExrImage.Write (ExrImage.cs:52-137): single-part scanline OpenEXR
 * v2, ZIP compression in 16-line blocks (zlib framing, a block that does not
 * shrink is stored raw), channels stored in ordinal name order, half samples
 * clamped to +-65504 with NaN kept, Rec.709 chromaticities.  Invalid channel
 * sets (none, empty / too long / non-printable / duplicate names, bad pixel
 * type) and dimensions are RT_ERR_INVALID_ARGUMENT, as the reference throws
 * ArgumentException. */
int rt_exr_encode(rt_ctx* ctx, int32_t width, int32_t height, const rt_exr_channel* channels, int32_t n_channels,
                  uint8_t* out, int64_t capacity, int64_t* size);

/* Replaces PfmImage.Write (PfmImage.cs:29-65): "PF" (3 channels) or "Pf"
 * (1), "W H", scale -1.0 (little-endian), rows bottom-to-top, channels
 * interleaved.  channels other than 1 or 3 are RT_ERR_INVALID_ARGUMENT. */
int rt_pfm_encode(rt_ctx* ctx, const float* planes, int32_t width, int32_t height, int32_t channels, uint8_t* out,
                  int64_t capacity, int64_t* size);

/* Replaces Program.SaveImage for .png (Program.cs:723-756): the display
 * planes (3 x W*H, rt_film.display) as 8-bit RGBA, channel byte =
 * clamp((int)(clamp(v, 0, 1) * 256), 0, 255) (NaN -> 0), alpha 255.  The
 * reference's PNG encoder is a library (ImageSharp); the pixels, not the
 * compressed bytes, are the contract. */
int rt_png_encode(rt_ctx* ctx, const float* display, int32_t width, int32_t height, uint8_t* out, int64_t capacity,
                  int64_t* size);

/* ── Denoiser (replaces NforDenoiser.Denoise, NforDenoiser.cs:29-57, with
 * DenoiserOptions, DenoiserOptions.cs:1-51).  Consumes the dual-buffer film
 * of a render with AOVs (beauty_a/b, albedo_a/b, normal_a/b, depth_a/b host
 * planes of rt_film; the other planes are not read) and writes the denoised
 * linear beauty (3 x W*H floats).  A film without AOV planes is
 * RT_ERR_INVALID_ARGUMENT, as the reference throws InvalidOperationException. */
#define RT_DENOISE_NLM 1     /* DenoiserKind.Nlm: joint NL-means */
#define RT_DENOISE_NFOR 2    /* DenoiserKind.Nfor: first-order regression (the default) */
#define RT_DENOISE_FAST 0    /* DenoiseQuality.Fast */
#define RT_DENOISE_HIGH 1    /* DenoiseQuality.High (the default) */
typedef struct rt_denoise_options {
    int32_t kind;              /* RT_DENOISE_NFOR or RT_DENOISE_NLM */
    int32_t quality;           /* RT_DENOISE_HIGH or RT_DENOISE_FAST */
    float selection_margin;    /* NFOR candidate-selection margin (NforRegression.SelectionMargin);
                                  < 0: the Sobol sampler's 0.5 */
    int32_t samples_a;         /* RenderBuffers.SamplesA = ceil(spp / 2) */
    int32_t samples_b;         /* RenderBuffers.SamplesB = floor(spp / 2) */
    int32_t reserved[3];
} rt_denoise_options;          /* 32 bytes */

int rt_denoise(rt_ctx* ctx, int32_t width, int32_t height, const rt_film* film, const rt_denoise_options* options,
               float* beauty_out);

/* Diagnostics (host only, no device needed): builds the scene on the host and
 * serialises the flattened world BVH in pre-order — internal node: (0, axis,
 * 6 box floats as int bits); leaf: (1, count, 6 box floats, entry ids) with
 * an entry id = primitive id or -(mesh ordinal + 2).  Returns the number of
 * int32 values (written only when cap is large enough), 0 when the world has
 * no BVH, < 0 on error.  No reference counterpart: it lets a test compare the
 * tree with the reference's BvhNode tree (BvhNode.cs:91-363) node by node. */
int64_t rt_debug_bvh_dump(const rt_scene_desc* desc, int32_t* out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* RT3D_H */
