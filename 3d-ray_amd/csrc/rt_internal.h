// rt_internal.h — in-HBM scene layout shared by the host builder and the
// gfx950 kernels.
//
// BVH nodes (64 B, one half of a 128-B L2 line): the two children's boxes
// live in the parent, so one node fetch gives both slab tests.
//   a = (c0.min.x, c0.min.y, c0.min.z, c0.max.x)
//   b = (c0.max.y, c0.max.z, c1.min.x, c1.min.y)
//   c = (c1.min.z, c1.max.x, c1.max.y, c1.max.z)
//   d = (child0, child1, split axis, 0)
// A child reference >= 0 is an internal node index; bit 31 set marks a leaf
// range of entries: count in bits 27..30, first entry in bits 0..26.
//
// Entries (64 B) are the primitives of a leaf in the reference's leaf order
// (BvhNode.cs:95-111 keeps the partitioned list order).  .a.w carries
// kind | flags << 8 (bit-cast int).
//   TRI    a = (v0, tag)     b = (e1, prim id) c = (e2, material) d = (n, 0)
//   SPHERE a = (centre, tag) b = (r, -, -, id) c = (-, -, -, mat)   d = 0
//   QUAD   a = (Q, tag)      b = (n, id)       c = (w, mat)         d = (D, aux, -, -)
//   PLANE  a = (point, tag)  b = (normal, id)  c = (-, -, -, mat)   d = 0
//   BVH    a = (box.min, tag) b = (box.max, child ref)              — a nested
//          BVH (the world BVH root or a Mesh's own BVH, Mesh.cs:93-96);
//          flag kEntryNoBox = a Mesh whose faces are a plain list (<= 4).
//   WRAP   a = (box.min, tag) b = (box.max, child entry) c = (transform,
//          override material, -, -) — Transform / Instance /
//          CameraInvisibleHittable around the entry at index `child`
//          (transform -1: none, material -1: none); no box test of its own
//          (Transform.Hit and Instance.Hit call the child directly); the box
//          is the builder's.
// Transforms: 8 float4 each, rows 1-4 of the object-to-world matrix then
// rows 1-4 of its inverse (System.Numerics layout, row vectors; w unused).
#pragma once
#include <stdint.h>
#include "rt_common.h"

namespace rt {

enum EntryKind : int32_t { kTri = 0, kSphere = 1, kQuad = 2, kPlane = 3, kBvh = 4, kWrap = 5 };

// Most wrapper entries on one root-to-primitive path, and the stack entry
// that marks a wrapper's frame (a leaf reference of count 0).
constexpr int kMaxWrap = 4;
enum EntryFlags : int32_t { kEntryCulled = 1, kEntryCamInvisible = 2, kEntryNoBox = 4 };

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kLeafCountShift = 27;
constexpr uint32_t kLeafStartMask = (1u << 27) - 1u;
RT_HD uint32_t leaf_ref(uint32_t start, uint32_t count) { return kLeafBit | (count << kLeafCountShift) | start; }
RT_HD bool ref_is_leaf(uint32_t r) { return (r & kLeafBit) != 0u; }
RT_HD uint32_t leaf_start(uint32_t r) { return r & kLeafStartMask; }
RT_HD uint32_t leaf_count(uint32_t r) { return (r >> kLeafCountShift) & 15u; }
constexpr uint32_t kFrameRef = kLeafBit;

// System.Numerics Vector3.Transform / TransformNormal with a Matrix4x4 given
// as its rows R[0..3] (row vectors: p' = p M), in the runtime's multiply-add
// order fma(x, row1, fma(y, row2, fma(z, row3, row4))); a direction drops
// row4.  xf_normal multiplies by the transposed inverse given the inverse's
// rows (Transform.cs:83, :188): component j is the dot with inverse row j.
RT_HD f3 xf_point(f3 p, const float4* R) {
    return mk(fmaf(p.x, R[0].x, fmaf(p.y, R[1].x, fmaf(p.z, R[2].x, R[3].x))),
              fmaf(p.x, R[0].y, fmaf(p.y, R[1].y, fmaf(p.z, R[2].y, R[3].y))),
              fmaf(p.x, R[0].z, fmaf(p.y, R[1].z, fmaf(p.z, R[2].z, R[3].z))));
}
RT_HD f3 xf_vec(f3 v, const float4* R) {
    return mk(fmaf(v.x, R[0].x, fmaf(v.y, R[1].x, v.z * R[2].x)),
              fmaf(v.x, R[0].y, fmaf(v.y, R[1].y, v.z * R[2].y)),
              fmaf(v.x, R[0].z, fmaf(v.y, R[1].z, v.z * R[2].z)));
}
RT_HD f3 xf_normal(f3 v, const float4* I) {
    return mk(fmaf(v.x, I[0].x, fmaf(v.y, I[0].y, v.z * I[0].z)),
              fmaf(v.x, I[1].x, fmaf(v.y, I[1].y, v.z * I[1].z)),
              fmaf(v.x, I[2].x, fmaf(v.y, I[2].y, v.z * I[2].z)));
}

struct alignas(64) GNode {
    float4 a, b, c;
    int4 d;
};
struct alignas(64) GEntry {
    float4 a, b, c, d;
};

// 4-wide node (128 B, one L2 line) of the any-hit BVH: the binary tree with
// every internal node's subtree collapsed to up to four children (the
// largest-area internal child is expanded first).  Child boxes in SoA order;
// ch = node index or leaf reference (same encoding as the binary tree); an
// unused slot holds a point box at 1e30 and an empty leaf.  A BVH entry's
// any-hit root is in c.x.
struct alignas(128) GNode4 {
    float4 mnx, mny, mnz, mxx, mxy, mxz;
    int4 ch, pad;
};

// Material kinds on the device (rt_material_kind) and derived flags.
enum MatFlags : int32_t {
    kMatThinWalled = 1,
    kMatHasTransColor = 2,
    kMatSpecTransmissive = 4,   // IsSpecularTransmissive
    kMatDielectricIor = 8,      // TryGetDielectricIor returns true
    kMatEmissiveProxy = 16,
    kMatTextured = 32,          // baseColor/roughness/metallic from the scene's texture set
    kMatImage = 64,             // albedo / baseColor from an ImageTexture (rt_material.image)
};

// ImageTexture on the device (ImageTexture.cs:42-289): one header per image
// and a table of mip levels; texels are linear RGB (w unused), level by level.
struct GImage {
    int32_t level0;             // first entry of the level table
    int32_t mips;               // level count, 1 + floor(log2(max(w, h)))
    float scale_u, scale_v;     // uv_scale
};
struct GLevel {
    int32_t off, w, h, pad;     // first texel, size
};

// Device material record (8 × 16 B).  Disney ShadingParams and the
// per-material constants of ComputeLobeWeights / ComputeF0 /
// ResolveTransmission are precomputed on the host (DisneyBsdf.cs:213-293,
// 1016-1049, 1097-1139, 1943-1959); they depend only on the material.
struct alignas(16) GMat {
    // m0: kind, flags, light index (emitter→light map, -1 none), pad
    int4 m0;
    // m1: base colour (or albedo / emission colour×intensity), metallic
    float4 m1;
    // m2: F0.xyz, roughness
    float4 m2;
    // m3: alpha, ax, ay, ior
    float4 m3;
    // m4: lobe probabilities pD, pS, pT, pSh
    float4 m4;
    // m5: pMs, pDt, pC, spec_trans
    float4 m5;
    // m6: sheen, sheen_tint, sheen_roughness, diff_trans
    float4 m6;
    // m7: clearcoat, cc_alpha, cc_f0, aniso rotation
    float4 m7;
    // m8: transmission tint.xyz, alpha_iso
    float4 m8;
    // m9: sigma_a.xyz, e_avg(alpha_iso)
    float4 m9;
    // m10: thin film thickness (nm, 0 = none), film ior (ShadingParams,
    // DisneyBsdf.cs:287-291)
    float4 m10;
};

// Light record (4 × 16 B)
struct alignas(16) GLight {
    // l0: kind, shadow samples, sqrt grid, geometry sub-kind (0 tri, 1 quad, 2 sphere)
    int4 l0;
    // l1: colour × (intensity where the estimator uses it) / emission, is_delta flag as float
    float4 l1;
    // l2: position / corner / v0 / direction / centre, radius or area
    float4 l2;
    // l3: u / e1, soft radius or cos(angular radius)
    float4 l3;
    // l4: v / e2, solid angle of the sun disc
    float4 l4;
    // l5: normal, area (area / geometry lights)
    float4 l5;
};

enum GeoSub : int32_t { kGeoTri = 0, kGeoQuad = 1, kGeoSphere = 2 };

}  // namespace rt
