// traverse.h — closest-hit traversal of the flattened BVH on gfx950.
//
// Visit order and every comparison reproduce the reference's recursive
// IHittable.Hit chain (HittableList.cs:19-38 → BvhNode.cs:157-206 →
// Mesh.cs:110-116 → primitive Hit), so the closest hit — including the
// "later primitive wins on equal t" tie rule — is bit-identical:
//   * the top-level list and every leaf are scanned in order with the
//     running closest t as tMax (accept t == tMax);
//   * at an internal node both child boxes are slab-tested against the
//     current closest; the near child (dir[axis] >= 0 → left) is entered
//     first and the far child is pushed with its entry distance tEnter;
//   * a popped far child is visited iff tEnter <= closest — exactly the
//     reference's far-box test `tExit >= tEnter` evaluated after the near
//     subtree, because its tExit term already passed at push time;
//   * a nested BVH (a Mesh, or the world BVH root) inside a leaf is entered
//     after a box test, with the rest of the leaf pushed as a continuation.
//
// The per-lane stack lives in LDS (the wavefront's traversal stack), entries
// (ref, tEnter) interleaved by lane so a wave's pushes/pops are conflict-free.
#pragma once
#include "rt_internal.h"

namespace rt {

struct Tri {
    f3 v0, e1, e2;
};

RT_HD f3 f3of(float4 v) { return mk(v.x, v.y, v.z); }
RT_HD int32_t tag_of(float4 a) { return __float_as_int(a.w); }

// Triangle.Hit (Triangle.cs:49-89)
__device__ __forceinline__ bool hit_triangle(f3 o, f3 d, f3 v0, f3 e1, f3 e2, float tmin, float tmax, float& t, float& u,
                                             float& v) {
    f3 h = cross(d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < 1e-7f) return false;
    float f = 1.0f / a;
    f3 s = o - v0;
    u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    f3 q = cross(s, e1);
    v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * dot(e2, q);
    if (t < tmin || t > tmax) return false;
    return true;
}

// Sphere.Hit (Sphere.cs:20-40)
__device__ __forceinline__ bool hit_sphere(f3 o, f3 d, f3 c, float r, float tmin, float tmax, float& t) {
    f3 oc = o - c;
    float a = length_sq(d);
    float hb = dot(oc, d);
    float cc = length_sq(oc) - r * r;
    float disc = hb * hb - a * cc;
    if (disc < 0.0f) return false;
    float sd = sqrtf(disc);
    float root = (-hb - sd) / a;
    if (root < tmin || root > tmax) {
        root = (-hb + sd) / a;
        if (root < tmin || root > tmax) return false;
    }
    t = root;
    return true;
}

// Quad.Hit (Quad.cs:318-362) [+ BackFaceCulledHittable when culled]
__device__ __forceinline__ bool hit_quad(f3 o, f3 d, const GEntry& e, const float4* aux, bool culled, float tmin,
                                         float tmax, float& t, float& u, float& v) {
    f3 n = f3of(e.b);
    float denom = dot(n, d);
    if (fabsf(denom) < 1e-8f) return false;
    float D = e.d.x;
    t = (D - dot(n, o)) / denom;
    if (t < tmin || t > tmax) return false;
    f3 q = f3of(e.a), w = f3of(e.c);
    int ai = __float_as_int(e.d.y);
    f3 U = f3of(aux[4 * ai]), V = f3of(aux[4 * ai + 1]);
    f3 ip = o + t * d;
    f3 pv = ip - q;
    u = dot(w, cross(pv, V));
    v = dot(w, cross(U, pv));
    if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f) return false;
    if (culled && !(dot(d, n) < 0.0f)) return false;
    return true;
}

// InfinitePlane.Hit (InfinitePlane.cs:28-40)
__device__ __forceinline__ bool hit_plane(f3 o, f3 d, f3 p, f3 n, float tmin, float tmax, float& t) {
    float denom = dot(n, d);
    if (fabsf(denom) < 1e-6f) return false;
    t = dot(p - o, n) / denom;
    if (t < tmin || t > tmax) return false;
    return true;
}

// AABB.Hit (AABB.cs:28-52).  The reference's Vector3.Min/Max and
// MathF.Max/Min propagate NaN (a NaN slab distance comes from an
// axis-parallel ray whose origin lies on a slab plane: 0 * inf), so such a
// box fails `tExit >= tEnter`; gfx950's v_minimum3 / v_maximum3 are those
// IEEE 754-2019 operations.  The reference's explicit empty-box rejection
// (Min > Max on an axis) is folded into the data: the host stores an empty
// box as NaN bounds, which fail the same comparison.
__device__ __forceinline__ float min_nan(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ float max_nan(float a, float b) { return __builtin_elementwise_maximum(a, b); }

__device__ __forceinline__ bool slab(f3 mn_, f3 mx_, f3 o, f3 inv, float tmin, float tmax, float& te) {
    float t0x = (mn_.x - o.x) * inv.x, t0y = (mn_.y - o.y) * inv.y, t0z = (mn_.z - o.z) * inv.z;
    float t1x = (mx_.x - o.x) * inv.x, t1y = (mx_.y - o.y) * inv.y, t1z = (mx_.z - o.z) * inv.z;
    float tsx = min_nan(t0x, t1x), tsy = min_nan(t0y, t1y), tsz = min_nan(t0z, t1z);
    float tlx = max_nan(t0x, t1x), tly = max_nan(t0y, t1y), tlz = max_nan(t0z, t1z);
    te = max_nan(tmin, max_nan(tsx, max_nan(tsy, tsz)));
    float tx = min_nan(tmax, min_nan(tlx, min_nan(tly, tlz)));
    return tx >= te;
}

struct HitOut {
    int32_t entry;   // entry index, -1 = miss
    float t, u, v;
    int4 chain;      // wrapper entries above the hit, outermost first (-1: none)
};

// The ray of the frame a hit was found in: the world ray taken through the
// transforms of the wrapper chain, outermost first (ToLocalRay,
// Transform.cs:140-155).
__device__ __forceinline__ void chain_ray(const GEntry* __restrict__ entries, const float4* __restrict__ xf,
                                          int4 chain, f3& o, f3& d) {
    const int32_t c[4] = {chain.x, chain.y, chain.z, chain.w};
#pragma unroll
    for (int k = 0; k < kMaxWrap; k++) {
        if (c[k] < 0) break;
        const int32_t xi = __float_as_int(entries[c[k]].c.x);
        if (xi >= 0) {
            const float4* I = xf + 8 * xi + 4;
            o = xf_point(o, I);
            d = xf_vec(d, I);
        }
    }
}

// Reference-traversal work counters (instrumented kernels only):
// node = BvhNode.Hit calls (box tests); tri = Hit calls on 64-B primitive
// records (triangles, quads, planes); other = Hit calls on spheres (16 B).
struct TravCount {
    uint32_t node, tri, other;
};

// Tests one primitive entry (a, b, c already loaded; d is fetched only for
// quads); updates `h` and `closest` on a hit and returns whether it hit.
template <bool COUNT>
__device__ __forceinline__ bool test_prim(const GEntry* __restrict__ entries, float4 a, float4 b, float4 c,
                                          uint32_t idx, int32_t kind, int32_t flags, f3 o, f3 d, float tmin,
                                          const float4* aux, float& closest, HitOut& h, TravCount& tc) {
    if (COUNT) {
        if (kind == kSphere) tc.other++;
        else tc.tri++;
    }
    float t, u = 0.0f, v = 0.0f;
    bool hit;
    if (kind == kTri) hit = hit_triangle(o, d, f3of(a), f3of(b), f3of(c), tmin, closest, t, u, v);
    else if (kind == kSphere) hit = hit_sphere(o, d, f3of(a), b.x, tmin, closest, t);
    else if (kind == kQuad) {
        GEntry e;
        e.a = a;
        e.b = b;
        e.c = c;
        e.d = entries[idx].d;
        hit = hit_quad(o, d, e, aux, (flags & kEntryCulled) != 0, tmin, closest, t, u, v);
    } else hit = hit_plane(o, d, f3of(a), f3of(b), tmin, closest, t);
    if (hit) {
        closest = t;
        h.entry = (int32_t)idx;
        h.t = t;
        h.u = u;
        h.v = v;
    }
    return hit;
}

// Stack layout: stk[k * kStride + lane] (uint2: ref, tEnter bits).
// ANY = true returns at the first primitive hit inside [tmin, tmax]: the
// answer to "is there any hit" equals the closest-hit answer's existence,
// because until the first hit both traversals visit the same nodes with
// the same tMax.
// XF: the scene has wrapper entries (Transform / Instance /
// CameraInvisibleHittable, rt_node).  Entering one pushes a frame marker,
// records the entry in the chain and maps the ray into the child's frame
// (a wrapper has no box test of its own: Transform.Hit and Instance.Hit call
// the child directly); popping the marker returns to the parent frame by
// re-deriving its ray from the world ray through the remaining chain, so the
// ray of every frame is the same float sequence the reference's nested Hit
// calls compute.  A hit carries the chain it was found under.
template <int SD, int STRIDE, bool COUNT, bool ANY = false, bool XF = false>
__device__ __forceinline__ HitOut trace_closest(const GNode* __restrict__ nodes, const GEntry* __restrict__ entries,
                                                const float4* __restrict__ aux, uint32_t top_start, uint32_t top_count,
                                                f3 o, f3 d, float tmin, float tmax, uint2* stk,
                                                uint32_t* overflow, TravCount& tc,
                                                const float4* __restrict__ xf = nullptr) {
    HitOut h;
    h.entry = -1;
    h.t = 0.0f;
    h.u = 0.0f;
    h.v = 0.0f;
    h.chain = make_int4(-1, -1, -1, -1);
    const f3 ow = o, dw = d;
    int4 chain = make_int4(-1, -1, -1, -1);
    int depth = 0;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float closest = tmax;
    // enters wrapper entry `e` (its c.x = transform); returns the child's leaf reference
    auto enter = [&](uint32_t e, float4 eb, float4 ec, int& sp) -> bool {
        if (depth >= kMaxWrap || sp >= SD) {
            *overflow = 1u;
            return false;
        }
        stk[sp * STRIDE] = make_uint2(kFrameRef, 0u);
        sp++;
        if (depth == 0) chain.x = (int32_t)e;
        else if (depth == 1) chain.y = (int32_t)e;
        else if (depth == 2) chain.z = (int32_t)e;
        else chain.w = (int32_t)e;
        depth++;
        const int32_t xi = __float_as_int(ec.x);
        if (xi >= 0) {
            const float4* I = xf + 8 * xi + 4;
            o = xf_point(o, I);
            d = xf_vec(d, I);
            inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        }
        return true;
    };
    auto leave = [&]() {
        depth--;
        if (depth == 0) chain.x = -1;
        else if (depth == 1) chain.y = -1;
        else if (depth == 2) chain.z = -1;
        else chain.w = -1;
        o = ow;
        d = dw;
        chain_ray(entries, xf, chain, o, d);
        inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    };
    for (uint32_t ti = 0; ti < top_count; ti++) {
        uint32_t eidx = top_start + ti;
        float4 ea = entries[eidx].a, eb = entries[eidx].b, ec = entries[eidx].c;
        int32_t tag = tag_of(ea);
        int32_t kind = tag & 0xff, flags = tag >> 8;
        if (kind != kBvh && !(XF && kind == kWrap)) {
            if (test_prim<COUNT>(entries, ea, eb, ec, eidx, kind, flags, o, d, tmin, aux, closest, h, tc)) {
                if (XF) h.chain = chain;
                if (ANY) return h;
            }
            continue;
        }
        int sp = 0;
        uint32_t cur;
        if (XF && kind == kWrap) {
            if (!enter(eidx, eb, ec, sp)) return h;
            cur = leaf_ref((uint32_t)__float_as_int(eb.w), 1u);
        } else {
            float te;
            if (COUNT && !(flags & kEntryNoBox)) tc.node++;
            if (!(flags & kEntryNoBox) && !slab(f3of(ea), f3of(eb), o, inv, tmin, closest, te)) continue;
            cur = (uint32_t)__float_as_int(eb.w);
        }
        for (;;) {
            if (ref_is_leaf(cur)) {
                uint32_t start = leaf_start(cur), cnt = leaf_count(cur);
                bool descended = false;
                for (uint32_t k = 0; k < cnt; k++) {
                    uint32_t idx = start + k;
                    float4 la = entries[idx].a, lb = entries[idx].b, lc = entries[idx].c;
                    int32_t ltag = tag_of(la);
                    int32_t lk = ltag & 0xff, lf = ltag >> 8;
                    if (lk != kBvh && !(XF && lk == kWrap)) {
                        if (test_prim<COUNT>(entries, la, lb, lc, idx, lk, lf, o, d, tmin, aux, closest, h, tc)) {
                            if (XF) h.chain = chain;
                            if (ANY) return h;
                        }
                        continue;
                    }
                    if (!(XF && lk == kWrap)) {
                        float lte;
                        if (COUNT && !(lf & kEntryNoBox)) tc.node++;
                        if (!(lf & kEntryNoBox) && !slab(f3of(la), f3of(lb), o, inv, tmin, closest, lte)) continue;
                    }
                    if (k + 1 < cnt) {
                        if (sp >= SD) { *overflow = 1u; return h; }
                        stk[sp * STRIDE] = make_uint2(leaf_ref(idx + 1, cnt - k - 1), __float_as_uint(-FLT_MAX));
                        sp++;
                    }
                    if (XF && lk == kWrap) {
                        if (!enter(idx, lb, lc, sp)) return h;
                        cur = leaf_ref((uint32_t)__float_as_int(lb.w), 1u);
                    } else cur = (uint32_t)__float_as_int(lb.w);
                    descended = true;
                    break;
                }
                if (descended) continue;
            } else {
                const GNode nd = nodes[cur];
                if (COUNT) tc.node += 2;
                f3 c0mn = mk(nd.a.x, nd.a.y, nd.a.z), c0mx = mk(nd.a.w, nd.b.x, nd.b.y);
                f3 c1mn = mk(nd.b.z, nd.b.w, nd.c.x), c1mx = mk(nd.c.y, nd.c.z, nd.c.w);
                float te0, te1;
                bool h0 = slab(c0mn, c0mx, o, inv, tmin, closest, te0);
                bool h1 = slab(c1mn, c1mx, o, inv, tmin, closest, te1);
                int axis = nd.d.z;
                float da = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
                bool left_first = da >= 0.0f;
                uint32_t rn = (uint32_t)(left_first ? nd.d.x : nd.d.y);
                uint32_t rf = (uint32_t)(left_first ? nd.d.y : nd.d.x);
                bool hn = left_first ? h0 : h1, hf = left_first ? h1 : h0;
                float tf = left_first ? te1 : te0;
                if (hn) {
                    if (hf) {
                        if (sp >= SD) { *overflow = 1u; return h; }
                        stk[sp * STRIDE] = make_uint2(rf, __float_as_uint(tf));
                        sp++;
                    }
                    cur = rn;
                    continue;
                }
                if (hf) {
                    cur = rf;
                    continue;
                }
            }
            // pop
            bool found = false;
            while (sp > 0) {
                sp--;
                uint2 s = stk[sp * STRIDE];
                if (XF && s.x == kFrameRef) {
                    leave();
                    continue;
                }
                if (__uint_as_float(s.y) <= closest) {
                    cur = s.x;
                    found = true;
                    break;
                }
            }
            if (!found) break;
        }
    }
    return h;
}

// Closest-hit traversal that also reports which accepted hit last wrote
// HitRecord.DpDu/DpDv on the reference's shared record.  BvhNode.Hit and
// Mesh.Hit pass one record down the whole chain and every accepted primitive
// hit overwrites the fields it sets: Triangle (edges), Quad (U, V) and
// Sphere (scaled tangent frame) set DpDu/DpDv, SmoothTriangle and
// InfinitePlane do not, so after a traversal whose closest hit is one of the
// latter the record still holds the last earlier writer's values.  A
// HittableList (the world list, a Mesh of <= 4 faces) collects its members
// in a fresh record and copies it out only when a member hit (Triangle.cs:83,
// Quad.cs:79, Sphere.cs:68-70, HittableList.cs:19-38, BvhNode.cs:157-206).
// Returns the closest hit as trace_closest does, with h_dp = the entry of the
// last writer (-1: none) and dp_t its hit distance.  `smooth` marks smooth
// triangles (4 float4 per primitive, w of the fourth != 0).
template <int SD, int STRIDE>
__device__ __forceinline__ HitOut trace_closest_dp(const GNode* __restrict__ nodes, const GEntry* __restrict__ entries,
                                                   const float4* __restrict__ aux, const float4* __restrict__ smooth,
                                                   uint32_t top_start, uint32_t top_count, f3 o, f3 d, float tmin,
                                                   float tmax, uint2* stk, uint32_t* overflow, int32_t& dp,
                                                   float& dp_t) {
    HitOut h;
    h.entry = -1;
    h.t = 0.0f;
    h.u = 0.0f;
    h.v = 0.0f;
    dp = -1;
    dp_t = 0.0f;
    TravCount tc{0u, 0u, 0u};
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float closest = tmax;
    auto writer = [&](uint32_t idx, int32_t kind, float4 b) {
        if (kind == kQuad || kind == kSphere) return true;
        if (kind != kTri) return false;
        return !(smooth && smooth[4 * (size_t)__float_as_int(b.w) + 3].w != 0.0f);
    };
    for (uint32_t ti = 0; ti < top_count; ti++) {
        uint32_t eidx = top_start + ti;
        float4 ea = entries[eidx].a, eb = entries[eidx].b, ec = entries[eidx].c;
        int32_t tag = tag_of(ea);
        int32_t kind = tag & 0xff, flags = tag >> 8;
        if (kind != kBvh) {
            if (test_prim<false>(entries, ea, eb, ec, eidx, kind, flags, o, d, tmin, aux, closest, h, tc) &&
                writer(eidx, kind, eb)) {
                dp = (int32_t)eidx;
                dp_t = h.t;
            }
            continue;
        }
        float te;
        if (!(flags & kEntryNoBox) && !slab(f3of(ea), f3of(eb), o, inv, tmin, closest, te)) continue;
        uint32_t cur = (uint32_t)__float_as_int(eb.w);
        bool list_leaf = (flags & kEntryNoBox) != 0;
        int sp = 0;
        for (;;) {
            if (ref_is_leaf(cur)) {
                uint32_t start = leaf_start(cur), cnt = leaf_count(cur);
                bool descended = false;
                // a <= 4-face Mesh is a HittableList: its own record
                int32_t outer_dp = dp;
                float outer_t = dp_t;
                bool any = false;
                if (list_leaf) dp = -1;
                for (uint32_t k = 0; k < cnt; k++) {
                    uint32_t idx = start + k;
                    float4 la = entries[idx].a, lb = entries[idx].b, lc = entries[idx].c;
                    int32_t ltag = tag_of(la);
                    int32_t lk = ltag & 0xff, lf = ltag >> 8;
                    if (lk != kBvh) {
                        if (test_prim<false>(entries, la, lb, lc, idx, lk, lf, o, d, tmin, aux, closest, h, tc)) {
                            any = true;
                            if (writer(idx, lk, lb)) {
                                dp = (int32_t)idx;
                                dp_t = h.t;
                            }
                        }
                        continue;
                    }
                    float lte;
                    if (!(lf & kEntryNoBox) && !slab(f3of(la), f3of(lb), o, inv, tmin, closest, lte)) continue;
                    if (k + 1 < cnt) {
                        if (sp >= SD) { *overflow = 1u; return h; }
                        stk[sp * STRIDE] = make_uint2(leaf_ref(idx + 1, cnt - k - 1), __float_as_uint(-FLT_MAX));
                        sp++;
                    }
                    cur = (uint32_t)__float_as_int(lb.w);
                    list_leaf = (lf & kEntryNoBox) != 0;
                    descended = true;
                    break;
                }
                if (list_leaf && !descended) {
                    if (!any) {
                        dp = outer_dp;
                        dp_t = outer_t;
                    }
                    list_leaf = false;
                }
                if (descended) continue;
            } else {
                const GNode nd = nodes[cur];
                f3 c0mn = mk(nd.a.x, nd.a.y, nd.a.z), c0mx = mk(nd.a.w, nd.b.x, nd.b.y);
                f3 c1mn = mk(nd.b.z, nd.b.w, nd.c.x), c1mx = mk(nd.c.y, nd.c.z, nd.c.w);
                float te0, te1;
                bool h0 = slab(c0mn, c0mx, o, inv, tmin, closest, te0);
                bool h1 = slab(c1mn, c1mx, o, inv, tmin, closest, te1);
                int axis = nd.d.z;
                float da = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
                bool left_first = da >= 0.0f;
                uint32_t rn = (uint32_t)(left_first ? nd.d.x : nd.d.y);
                uint32_t rf = (uint32_t)(left_first ? nd.d.y : nd.d.x);
                bool hn = left_first ? h0 : h1, hf = left_first ? h1 : h0;
                float tf = left_first ? te1 : te0;
                if (hn) {
                    if (hf) {
                        if (sp >= SD) { *overflow = 1u; return h; }
                        stk[sp * STRIDE] = make_uint2(rf, __float_as_uint(tf));
                        sp++;
                    }
                    cur = rn;
                    continue;
                }
                if (hf) {
                    cur = rf;
                    continue;
                }
            }
            // pop
            bool found = false;
            while (sp > 0) {
                sp--;
                uint2 s = stk[sp * STRIDE];
                if (__uint_as_float(s.y) <= closest) {
                    cur = s.x;
                    found = true;
                    break;
                }
            }
            if (!found) break;
        }
    }
    return h;
}

// Box test of the any-hit traversal: the same test as slab() without the
// entry distance.  Unused child slots hold a point box far outside the scene
// whose reference is an empty leaf.
__device__ __forceinline__ bool slab_any(f3 mn_, f3 mx_, f3 o, f3 inv, float tmin, float tmax) {
    float t0x = (mn_.x - o.x) * inv.x, t0y = (mn_.y - o.y) * inv.y, t0z = (mn_.z - o.z) * inv.z;
    float t1x = (mx_.x - o.x) * inv.x, t1y = (mx_.y - o.y) * inv.y, t1z = (mx_.z - o.z) * inv.z;
    float te = max_nan(tmin, max_nan(min_nan(t0x, t1x), max_nan(min_nan(t0y, t1y), min_nan(t0z, t1z))));
    float tx = min_nan(tmax, min_nan(max_nan(t0x, t1x), min_nan(max_nan(t0y, t1y), max_nan(t0z, t1z))));
    return tx >= te;
}

// Any-hit query on the 4-wide BVH (GNode4): is there a primitive hit with
// t in [tmin, tmax]?  The answer equals the closest-hit traversal's
// existence of a hit (every primitive whose boxes the segment enters is
// tested; the visit order does not matter for a yes/no answer), with a
// quarter of the dependent node fetches.  Hit children are visited nearest
// first (a 5-exchange sorting network on their entry distances), which
// finds occluders near the shading point sooner.
template <int SD, int STRIDE, bool COUNT = false>
__device__ __forceinline__ bool trace_any4(const GNode4* __restrict__ nodes4, const GEntry* __restrict__ entries,
                                           const float4* __restrict__ aux, uint32_t top_start, uint32_t top_count,
                                           f3 o, f3 d, float tmin, float tmax, uint2* stk, uint32_t* overflow,
                                           TravCount& tc) {
    HitOut h;
    h.entry = -1;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float closest = tmax;
    for (uint32_t ti = 0; ti < top_count; ti++) {
        uint32_t eidx = top_start + ti;
        float4 ea = entries[eidx].a, eb = entries[eidx].b, ec = entries[eidx].c;
        int32_t tag = tag_of(ea);
        int32_t kind = tag & 0xff, flags = tag >> 8;
        if (kind != kBvh) {
            if (test_prim<COUNT>(entries, ea, eb, ec, eidx, kind, flags, o, d, tmin, aux, closest, h, tc)) return true;
            continue;
        }
        float te;
        if (COUNT && !(flags & kEntryNoBox)) tc.node++;
        if (!(flags & kEntryNoBox) && !slab(f3of(ea), f3of(eb), o, inv, tmin, closest, te)) continue;
        uint32_t cur = (uint32_t)__float_as_int(ec.x);
        int sp = 0;
        for (;;) {
            if (ref_is_leaf(cur)) {
                uint32_t start = leaf_start(cur), cnt = leaf_count(cur);
                bool descended = false;
                for (uint32_t k = 0; k < cnt; k++) {
                    uint32_t idx = start + k;
                    float4 la = entries[idx].a, lb = entries[idx].b, lc = entries[idx].c;
                    int32_t ltag = tag_of(la);
                    int32_t lk = ltag & 0xff, lf = ltag >> 8;
                    if (lk != kBvh) {
                        if (test_prim<COUNT>(entries, la, lb, lc, idx, lk, lf, o, d, tmin, aux, closest, h, tc))
                            return true;
                        continue;
                    }
                    float lte;
                    if (COUNT && !(lf & kEntryNoBox)) tc.node++;
                    if (!(lf & kEntryNoBox) && !slab(f3of(la), f3of(lb), o, inv, tmin, closest, lte)) continue;
                    if (k + 1 < cnt) {
                        if (sp >= SD) { *overflow = 1u; return false; }
                        stk[sp * STRIDE] = make_uint2(leaf_ref(idx + 1, cnt - k - 1), 0u);
                        sp++;
                    }
                    cur = (uint32_t)__float_as_int(lc.x);
                    descended = true;
                    break;
                }
                if (descended) continue;
            } else {
                const GNode4 nd = nodes4[cur];
                if (COUNT) tc.node++;   // one 128-B node fetch
                const float bx0[4] = {nd.mnx.x, nd.mnx.y, nd.mnx.z, nd.mnx.w};
                const float by0[4] = {nd.mny.x, nd.mny.y, nd.mny.z, nd.mny.w};
                const float bz0[4] = {nd.mnz.x, nd.mnz.y, nd.mnz.z, nd.mnz.w};
                const float bx1[4] = {nd.mxx.x, nd.mxx.y, nd.mxx.z, nd.mxx.w};
                const float by1[4] = {nd.mxy.x, nd.mxy.y, nd.mxy.z, nd.mxy.w};
                const float bz1[4] = {nd.mxz.x, nd.mxz.y, nd.mxz.z, nd.mxz.w};
                const uint32_t cref[4] = {(uint32_t)nd.ch.x, (uint32_t)nd.ch.y, (uint32_t)nd.ch.z, (uint32_t)nd.ch.w};
                // hit children nearest first: entered in order of entry
                // distance (a miss sorts last), the others pushed farthest first
                float key[4];
                uint32_t ref[4];
                int nh = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float te;
                    const bool hk = slab(mk(bx0[k], by0[k], bz0[k]), mk(bx1[k], by1[k], bz1[k]), o, inv, tmin, closest, te);
                    key[k] = hk ? te : __int_as_float(0x7f800000);
                    ref[k] = cref[k];
                    nh += hk ? 1 : 0;
                }
                auto cswap = [&](int a, int b) {
                    const bool sw = key[b] < key[a];
                    const float ka = key[a], kb = key[b];
                    const uint32_t ra = ref[a], rb = ref[b];
                    key[a] = sw ? kb : ka;
                    key[b] = sw ? ka : kb;
                    ref[a] = sw ? rb : ra;
                    ref[b] = sw ? ra : rb;
                };
                cswap(0, 1);
                cswap(2, 3);
                cswap(0, 2);
                cswap(1, 3);
                cswap(1, 2);
                if (nh > 0) {
                    // the pushes overflow exactly when the one-by-one pushes would
                    if (sp + nh - 1 > SD) { *overflow = 1u; return false; }
                    if (nh > 3) { stk[sp * STRIDE] = make_uint2(ref[3], 0u); sp++; }
                    if (nh > 2) { stk[sp * STRIDE] = make_uint2(ref[2], 0u); sp++; }
                    if (nh > 1) { stk[sp * STRIDE] = make_uint2(ref[1], 0u); sp++; }
                    cur = ref[0];
                    continue;
                }
            }
            if (sp == 0) break;
            sp--;
            cur = stk[sp * STRIDE].x;
        }
    }
    return false;
}

// HitRecord.U/V of a reported hit: barycentrics for triangles, (alpha,
// beta) for quads (both returned by the test), Sphere.GetSphereUV
// (Sphere.cs:260-269) and InfinitePlane's tiled planar UV
// (InfinitePlane.cs:45-55).
__device__ __forceinline__ void report_uv(const GEntry* __restrict__ entries, const HitOut& h, f3 o, f3 d, float& u,
                                          float& v) {
    u = h.u;
    v = h.v;
    if (h.entry < 0) return;
    float4 a = entries[h.entry].a, b = entries[h.entry].b;
    int kind = tag_of(a) & 0xff;
    if (kind == kSphere) {
        f3 p = o + h.t * d;
        f3 n = (p - f3of(a)) / b.x;
        float theta = acosf(-n.y);
        float phi = atan2f(-n.z, n.x) + kPi;
        u = phi / (2.0f * kPi);
        v = theta / kPi;
    } else if (kind == kPlane) {
        f3 n = f3of(b);
        f3 ua = fabsf(n.y) < 0.99f ? normalize(cross(n, mk(0.0f, 1.0f, 0.0f))) : mk(1.0f, 0.0f, 0.0f);
        f3 va = cross(n, ua);
        f3 lp = (o + h.t * d) - f3of(a);
        u = dot(lp, ua);
        v = dot(lp, va);
        u -= floorf(u);
        v -= floorf(v);
    }
}

}  // namespace rt
