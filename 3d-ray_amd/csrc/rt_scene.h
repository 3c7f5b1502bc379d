// rt_scene.h — host-side scene representation and the builder interface.
#pragma once
#include <string>
#include <vector>

#include "../../include/rt3d.h"
#include "rt_internal.h"
#include "material.h"

namespace rt {

struct Box {
    f3 mn, mx;
};

// Host copy of everything rt_scene_create uploads.
struct HostScene {
    std::vector<GNode> nodes;
    std::vector<GNode4> nodes4;        // any-hit BVH (collapsed binary tree)
    int any_stack = 0;                 // traversal stack bound of the any-hit BVH
    std::vector<GEntry> entries;
    std::vector<float4> quad_aux;      // 4 per quad: (U, area) (V, 0) (tangent, 0) (bitangent, 0)
    std::vector<GMat> mats;
    std::vector<GLight> lights;
    std::vector<int32_t> light_prim;   // geometry light → primitive id (diagnostics)
    uint32_t top_start = 0, top_count = 0;   // top-level HittableList entries
    bool has_bvh = false;
    int max_depth = 0;                 // deepest root-to-leaf path over all BVHs
    int world_depth = 0;               // deepest leaf of the world BVH
    int mesh_depth = -1;               // deepest leaf over nested mesh BVHs (-1: none)
    std::vector<float> light_power;    // ILight.ApproximatePower per light
    std::vector<uint32_t> mesh_refs;   // root reference of each mesh object, object order
    uint32_t world_root = 0;           // root reference of the world BVH (has_bvh)
    Box world_box;
    int64_t num_internal = 0, num_leaves = 0, num_objects = 0, num_prims = 0;
    float lut_e[32 * 32];
    float lut_eavg[32];
    // camera basis (Camera.cs:26-51)
    f3 cam_origin, cam_llc, cam_horiz, cam_vert, cam_u, cam_v;
    float lens_radius = 0.0f;
    f3 sky;
    int env_light = -1;                // first EnvironmentLight of the list (Renderer.cs:426)
    bool thin_film = false;            // some Disney material has a thin film
    bool medium = false;
    bool opaque_shadows = true;        // no material transmits shadow rays
    f3 sig_a, sig_s, sig_t;
    // scene analysis (Renderer.cs:361-378)
    Box bounds;
    float total_flux = 0, scene_radius = 0, mean_irradiance = 0;
    int rr_min_bounces = 3;
    float rr_min_survival = 0.1f;
    int max_shadow_per_vertex = 0;     // Σ_l shadow_samples
    int max_light_samples = 0;         // max_l shadow_samples
    // texture set of RT_MATF_TEXTURED materials
    std::vector<rt_material> raw_mats; // the material records as given
    bool textured = false;
    int tex_w = 0, tex_h = 0;
    std::vector<float> texels;         // planar [kTexChannels][tex_h][tex_w]
    std::vector<float4> tex_uv;        // 2 per primitive: (uv0, uv1), (uv2, -, -); empty = HitRecord U/V
    // SmoothTriangle data, 4 per primitive: (N0, T.x) (N1, T.y) (N2, T.z)
    // (B, 1); a zero record is a flat triangle; empty = no smooth triangles
    std::vector<float4> smooth;
    // image textures (ImageTexture): headers, mip levels, linear RGB texels
    std::vector<GImage> images;
    std::vector<GLevel> img_levels;
    std::vector<float4> img_texels;
    // object hierarchy (rt_node): wrapper transforms, 8 per TRANSFORM node
    // (matrix rows 1-4, inverse rows 1-4); `hier` = the world was assembled
    // from nodes (wrapper entries may occur)
    std::vector<float4> xf;
    bool hier = false;
    int hier_stack = 0;                // traversal stack bound of the hierarchy
};

// Builds the host scene; returns RT_OK or an error with `err` set.
int build_host_scene(const rt_scene_desc* d, HostScene& hs, std::string& err);

// Kulla–Conty energy table (EnergyCompensationLut.cs:34-130)
void build_energy_lut(float* E, float* Eavg);
float lut_sample_eavg(const float* Eavg, float alpha);

}  // namespace rt
