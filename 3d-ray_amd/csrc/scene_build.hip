// scene_build.hip — host side of rt_scene_create: world assembly, binned-SAH
// BVH, flattening to the 64-B node/entry layout, material/light tables,
// energy LUT and the Renderer constructor's scene analysis.
//
// The BVH follows BvhNode.cs:91-363 step for step (16 bins per axis,
// traversal cost 0.5, strict '<' keeps the first best split, in-place swap
// partition `idx <= bestBin`, median fallback, fat leaves of <= 4): the tree
// must be the reference's tree, because the bit-exact primary-hit gate
// depends on the visit order it induces.  Large spans are built on worker
// threads like the reference's Parallel.Invoke (BvhNode.cs:132-142); the
// ranges are disjoint so the result does not depend on scheduling.
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <cstring>
#include <functional>
#include <future>
#include <string>
#include <memory>
#include <thread>

#include "rt_scene.h"

namespace rt {

static Box box_empty() { return Box{splat(FLT_MAX), splat(-FLT_MAX)}; }
static Box box_union(const Box& a, const Box& b) { return Box{vmin(a.mn, b.mn), vmax(a.mx, b.mx)}; }
static float box_area(const Box& b) {
    f3 d = b.mx - b.mn;
    if (d.x < 0.0f || d.y < 0.0f || d.z < 0.0f) return 0.0f;
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
// The C ABI structs have fixed sizes (C# [StructLayout(Sequential)] and
// ctypes mirror them field for field).
static_assert(sizeof(rt_primitive) == 64, "rt_primitive layout");
static_assert(sizeof(rt_material) == 128, "rt_material layout");
static_assert(sizeof(rt_light) == 80, "rt_light layout");
static_assert(sizeof(rt_camera) == 56, "rt_camera layout");
static_assert(sizeof(rt_medium) == 32, "rt_medium layout");
static_assert(sizeof(rt_scene_desc) == 248, "rt_scene_desc layout");
static_assert(sizeof(rt_node) == 160, "rt_node layout");
static_assert(sizeof(rt_image) == 24, "rt_image layout");
static_assert(sizeof(rt_settings) == 96, "rt_settings layout");
static_assert(sizeof(rt_film) == 88, "rt_film layout");
static_assert(sizeof(rt_render_stats) == 152, "rt_render_stats layout");
static_assert(sizeof(rt_scene_info) == 96, "rt_scene_info layout");
static_assert(sizeof(rt_batch_job) == 32, "rt_batch_job layout");
static_assert(sizeof(rt_exr_channel) == 24, "rt_exr_channel layout");
static_assert(sizeof(rt_denoise_options) == 32, "rt_denoise_options layout");

static f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// Tangent frame of a triangle's UV mapping, E1 = dUV1.x T + dUV1.y B and
// E2 = dUV2.x T + dUV2.y B solved for T, B; normalised edges when the mapping
// is degenerate (SmoothTriangle.cs:98-129).  uv = u0 v0 u1 v1 u2 v2.
static void uv_frame(f3 e1, f3 e2, const float* uv, f3& T, f3& B) {
    const float d1x = uv[2] - uv[0], d1y = uv[3] - uv[1];
    const float d2x = uv[4] - uv[0], d2y = uv[5] - uv[1];
    const float det = d1x * d2y - d2x * d1y;
    if (fabsf(det) > 1e-8f) {
        const float inv = 1.0f / det;
        f3 t = inv * (d2y * e1 - d1y * e2);
        f3 b = inv * (-d2x * e1 + d1x * e2);
        const float tl = length(t), bl = length(b);
        if (tl > 1e-8f && bl > 1e-8f) {
            T = t / tl;
            B = b / bl;
            return;
        }
    }
    T = normalize(e1);
    B = normalize(e2);
}
// ImageTexture construction (ImageTexture.cs:60-138): 8-bit sRGB decoded
// through the gamma-2.2 table, level 0 in linear RGB, then 2x2 box levels
// (floor halving, the last row/column reused at odd edges, sum in texel
// order times 0.25) down to 1 + floor(log2(max(w, h))) levels.
static int build_images(const rt_scene_desc* d, HostScene& hs, std::string& err) {
    const int64_t ni = d->num_images;
    if (ni < 0 || (ni > 0 && !d->images)) { err = "bad image table"; return RT_ERR_INVALID_ARGUMENT; }
    for (int64_t i = 0; i < d->num_materials; i++) {
        const rt_material& m = d->materials[i];
        if (!(m.flags & RT_MATF_IMAGE)) continue;
        if (m.kind != RT_MAT_LAMBERTIAN && m.kind != RT_MAT_DISNEY) {
            err = "image textures are supported as Lambertian albedo and Disney baseColor only";
            return RT_ERR_UNSUPPORTED;
        }
        if (m.kind == RT_MAT_DISNEY && (m.spec_trans != 0.0f || (m.flags & RT_MATF_TEXTURED))) {
            err = "an image-textured Disney material needs spec_trans = 0 and no texture set";
            return RT_ERR_UNSUPPORTED;
        }
        if (m.image < 0 || m.image >= ni) { err = "material image index out of range"; return RT_ERR_INVALID_ARGUMENT; }
    }
    // ImageTexture's sRGB decode table, built once (thread-safe static init)
    static const std::array<float, 256> lut = [] {
        std::array<float, 256> t{};
        const float inv255 = 1.0f / 255.0f;
        for (int i = 0; i < 256; i++) t[i] = powf((float)i * inv255, 2.2f);
        return t;
    }();
    int64_t total = 0;
    for (int64_t k = 0; k < ni; k++) {
        const rt_image& im = d->images[k];
        if (!im.rgba8 || im.width <= 0 || im.height <= 0 || (int64_t)im.width * im.height > (1ll << 26)) {
            err = "bad image (null pixels, empty, or more than 2^26 texels)";
            return RT_ERR_INVALID_ARGUMENT;
        }
        int w = im.width, h = im.height, big = std::max(w, h), mips = 1;
        while ((big >>= 1) > 0) mips++;
        GImage g{(int32_t)hs.img_levels.size(), mips, im.scale_u, im.scale_v};
        hs.images.push_back(g);
        int64_t off = (int64_t)hs.img_texels.size();
        for (int m = 0; m < mips; m++) {
            total += (int64_t)w * h;
            if (total > (1ll << 30)) { err = "image textures exceed 2^30 texels"; return RT_ERR_UNSUPPORTED; }
            hs.img_levels.push_back(GLevel{(int32_t)off, w, h, 0});
            off += (int64_t)w * h;
            w = w > 1 ? w / 2 : 1;
            h = h > 1 ? h / 2 : 1;
        }
        hs.img_texels.resize((size_t)off);
        const GLevel* L = &hs.img_levels[g.level0];
        float4* t0 = &hs.img_texels[L[0].off];
        for (int64_t p = 0; p < (int64_t)L[0].w * L[0].h; p++) {
            const uint8_t* px = im.rgba8 + 4 * p;
            t0[p] = make_float4(lut[px[0]], lut[px[1]], lut[px[2]], 0.0f);
        }
        for (int m = 1; m < mips; m++) {
            const GLevel pv = L[m - 1], cu = L[m];
            const float4* prev = &hs.img_texels[pv.off];
            float4* cur = &hs.img_texels[cu.off];
            for (int y = 0; y < cu.h; y++)
                for (int x = 0; x < cu.w; x++) {
                    int x0 = 2 * x, y0 = 2 * y;
                    int x1 = std::min(x0 + 1, pv.w - 1), y1 = std::min(y0 + 1, pv.h - 1);
                    float r = 0.0f, gg = 0.0f, b = 0.0f;
                    const int idx[4] = {y0 * pv.w + x0, y0 * pv.w + x1, y1 * pv.w + x0, y1 * pv.w + x1};
                    for (int q = 0; q < 4; q++) {
                        r += prev[idx[q]].x;
                        gg += prev[idx[q]].y;
                        b += prev[idx[q]].z;
                    }
                    cur[y * cu.w + x] = make_float4(r * 0.25f, gg * 0.25f, b * 0.25f, 0.0f);
                }
        }
    }
    return RT_OK;
}

static int32_t fbits(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
static float ibits(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

// Primitive bounding boxes (Triangle.cs:144-149, Sphere.cs:271-275,
// Quad.cs:389-396, InfinitePlane.cs:69-73)
static Box prim_box(const rt_primitive& p) {
    const float* v = p.v;
    switch (p.kind) {
        case RT_PRIM_TRIANGLE: {
            f3 a = ld3(v), b = ld3(v + 3), c = ld3(v + 6);
            return Box{vmin(vmin(a, b), c) - splat(kEps), vmax(vmax(a, b), c) + splat(kEps)};
        }
        case RT_PRIM_SPHERE: {
            f3 c = ld3(v);
            f3 r = splat(v[3]);
            return Box{c - r, c + r};
        }
        case RT_PRIM_QUAD: {
            f3 q = ld3(v), u = ld3(v + 3), w = ld3(v + 6);
            f3 bmin = q, bmax = (q + u) + w;
            return Box{vmin(bmin, vmin(q + u, q + w)) - splat(0.0001f), vmax(bmax, vmax(q + u, q + w)) + splat(0.0001f)};
        }
        default:
            return Box{splat(-1e6f), splat(1e6f)};
    }
}

// ── binned SAH build ─────────────────────────────────────────────────────
struct BNode {
    Box box;
    int axis = 0;
    std::unique_ptr<BNode> l, r;
    int start = 0, end = 0;   // leaf item range
    bool leaf() const { return !l; }
};

static bool try_sah(std::vector<int32_t>& items, std::vector<Box>& bx, int start, int end, int& best_axis, int& mid) {
    constexpr int kBins = 16;
    int span = end - start;
    f3 cmin = splat(FLT_MAX), cmax = splat(-FLT_MAX);
    for (int i = start; i < end; i++) {
        f3 c = (bx[i].mn + bx[i].mx) * 0.5f;
        cmin = vmin(cmin, c);
        cmax = vmax(cmax, c);
    }
    f3 ext = cmax - cmin;
    float max_ext = mx(ext.x, mx(ext.y, ext.z));
    if (max_ext <= 0.0f) { best_axis = -1; mid = start; return false; }
    best_axis = -1;
    int best_bin = -1;
    float best_cost = FLT_MAX;
    int cnt[kBins], pc[kBins], sc[kBins];
    Box bb[kBins], pb[kBins], sb[kBins];
    for (int axis = 0; axis < 3; axis++) {
        float ea = get(ext, axis);
        if (ea <= 0.0f) continue;
        float cm = get(cmin, axis);
        float scale = (float)kBins / ea;
        for (int i = 0; i < kBins; i++) { cnt[i] = 0; bb[i] = box_empty(); }
        for (int i = start; i < end; i++) {
            float c = (get(bx[i].mn, axis) + get(bx[i].mx, axis)) * 0.5f;
            int idx = (int)((c - cm) * scale);
            if (idx < 0) idx = 0;
            else if (idx >= kBins) idx = kBins - 1;
            cnt[idx]++;
            bb[idx] = box_union(bb[idx], bx[i]);
        }
        int cum = 0;
        Box cb = box_empty();
        for (int i = 0; i < kBins; i++) { cum += cnt[i]; cb = box_union(cb, bb[i]); pc[i] = cum; pb[i] = cb; }
        cum = 0;
        cb = box_empty();
        for (int i = kBins - 1; i >= 0; i--) { cum += cnt[i]; cb = box_union(cb, bb[i]); sc[i] = cum; sb[i] = cb; }
        float parea = box_area(pb[kBins - 1]);
        if (parea <= 0.0f) continue;
        for (int i = 0; i < kBins - 1; i++) {
            int nl = pc[i], nr = sc[i + 1];
            if (nl == 0 || nr == 0) continue;
            float cost = 0.5f + ((float)nl * box_area(pb[i]) + (float)nr * box_area(sb[i + 1])) / parea;
            if (cost < best_cost) { best_cost = cost; best_axis = axis; best_bin = i; }
        }
    }
    if (best_axis < 0) { mid = start; return false; }
    float amin = get(cmin, best_axis);
    float ascale = (float)kBins / get(ext, best_axis);
    int wp = start;
    for (int i = start; i < end; i++) {
        float c = (get(bx[i].mn, best_axis) + get(bx[i].mx, best_axis)) * 0.5f;
        int idx = (int)((c - amin) * ascale);
        if (idx < 0) idx = 0;
        else if (idx >= kBins) idx = kBins - 1;
        if (idx <= best_bin) {
            std::swap(items[i], items[wp]);
            std::swap(bx[i], bx[wp]);
            wp++;
        }
    }
    mid = wp;
    if (mid == start || mid == end) mid = start + span / 2;
    return true;
}

static std::unique_ptr<BNode> build_node(std::vector<int32_t>& items, std::vector<Box>& bx, int start, int end, int par_depth) {
    auto n = std::make_unique<BNode>();
    int span = end - start;
    if (span <= 4) {
        Box b = bx[start];
        for (int i = 1; i < span; i++) b = box_union(b, bx[start + i]);
        n->box = b;
        n->start = start;
        n->end = end;
        return n;
    }
    int ba, mid;
    if (!try_sah(items, bx, start, end, ba, mid)) { ba = 0; mid = start + span / 2; }
    n->axis = ba;
    if (span > 8192 && par_depth < 4) {
        auto fl = std::async(std::launch::async, [&, start, mid]() { return build_node(items, bx, start, mid, par_depth + 1); });
        n->r = build_node(items, bx, mid, end, par_depth + 1);
        n->l = fl.get();
    } else {
        n->l = build_node(items, bx, start, mid, par_depth);
        n->r = build_node(items, bx, mid, end, par_depth);
    }
    n->box = box_union(n->l->box, n->r->box);
    return n;
}

// Device copy of a box: an empty box (min > max on an axis, AABB.Empty) as
// NaN bounds, which the traversal's NaN-propagating slab test rejects as
// AABB.Hit rejects an empty box (AABB.cs:36).
static Box dev_box(const Box& b) {
    if (b.mn.x > b.mx.x || b.mn.y > b.mx.y || b.mn.z > b.mx.z) {
        const float q = std::numeric_limits<float>::quiet_NaN();
        return Box{mk(q, q, q), mk(q, q, q)};
    }
    return b;
}

struct Flattener {
    HostScene& hs;
    const rt_scene_desc* d;
    std::vector<int32_t> quad_aux_of;          // primitive → quad aux index
    // object → entry template (for mesh objects: a BVH entry)
    std::vector<GEntry> obj_entry;
    int deepest = 0;                            // deepest leaf seen by the current flatten

    GEntry prim_entry(int32_t id) const {
        const rt_primitive& p = d->primitives[id];
        const float* v = p.v;
        GEntry e{};
        int32_t flags = 0;
        if (p.flags & RT_PRIMF_BACKFACE_CULLED) flags |= kEntryCulled;
        if (p.flags & RT_PRIMF_CAMERA_INVISIBLE) flags |= kEntryCamInvisible;
        int32_t tag = p.kind | (flags << 8);
        switch (p.kind) {
            case RT_PRIM_TRIANGLE: {
                f3 v0 = ld3(v), v1 = ld3(v + 3), v2 = ld3(v + 6);
                f3 e1 = v1 - v0, e2 = v2 - v0;
                f3 cr = cross(e1, e2);
                float cl = length(cr);
                f3 n = cl > 0.0f ? cr / cl : mk(0, 1, 0);
                e.a = make_float4(v0.x, v0.y, v0.z, ibits(tag));
                e.b = make_float4(e1.x, e1.y, e1.z, ibits(id));
                e.c = make_float4(e2.x, e2.y, e2.z, ibits(p.material));
                e.d = make_float4(n.x, n.y, n.z, 0.5f * cl);
                break;
            }
            case RT_PRIM_SPHERE:
                e.a = make_float4(v[0], v[1], v[2], ibits(tag));
                e.b = make_float4(v[3], 0.0f, 0.0f, ibits(id));
                e.c = make_float4(0.0f, 0.0f, 0.0f, ibits(p.material));
                break;
            case RT_PRIM_QUAD: {
                f3 q = ld3(v), u = ld3(v + 3), w = ld3(v + 6);
                f3 nn = cross(u, w);
                f3 n = normalize(nn);
                float dd = dot(n, q);
                f3 ww = nn / dot(nn, nn);
                e.a = make_float4(q.x, q.y, q.z, ibits(tag));
                e.b = make_float4(n.x, n.y, n.z, ibits(id));
                e.c = make_float4(ww.x, ww.y, ww.z, ibits(p.material));
                e.d = make_float4(dd, ibits(quad_aux_of[id]), 0.0f, 0.0f);
                break;
            }
            default: {
                f3 n = normalize(mk(v[3], v[4], v[5]));
                e.a = make_float4(v[0], v[1], v[2], ibits(tag));
                e.b = make_float4(n.x, n.y, n.z, ibits(id));
                e.c = make_float4(0.0f, 0.0f, 0.0f, ibits(p.material));
                break;
            }
        }
        return e;
    }

    // Flattens a built tree whose items are object ids (or primitive ids when
    // `prim_items`), returns the root reference.
    uint32_t flatten(const BNode* n, const std::vector<int32_t>& items, bool prim_items, int depth) {
        hs.max_depth = std::max(hs.max_depth, depth);
        deepest = std::max(deepest, depth);
        if (n->leaf()) {
            hs.num_leaves++;
            uint32_t start = (uint32_t)hs.entries.size();
            for (int i = n->start; i < n->end; i++)
                hs.entries.push_back(prim_items ? prim_entry(items[i]) : obj_entry[items[i]]);
            return leaf_ref(start, (uint32_t)(n->end - n->start));
        }
        hs.num_internal++;
        uint32_t idx = (uint32_t)hs.nodes.size();
        hs.nodes.push_back(GNode{});
        uint32_t rl = flatten(n->l.get(), items, prim_items, depth + 1);
        uint32_t rr = flatten(n->r.get(), items, prim_items, depth + 1);
        const Box a = dev_box(n->l->box);
        const Box b = dev_box(n->r->box);
        GNode& g = hs.nodes[idx];
        g.a = make_float4(a.mn.x, a.mn.y, a.mn.z, a.mx.x);
        g.b = make_float4(a.mx.y, a.mx.z, b.mn.x, b.mn.y);
        g.c = make_float4(b.mn.z, b.mx.x, b.mx.y, b.mx.z);
        g.d = make_int4((int32_t)rl, (int32_t)rr, n->axis, 0);
        return idx;
    }
};

static GEntry bvh_entry(const Box& box, uint32_t ref, int32_t flags) {
    const Box b = dev_box(box);
    GEntry e{};
    e.a = make_float4(b.mn.x, b.mn.y, b.mn.z, ibits(kBvh | (flags << 8)));
    e.b = make_float4(b.mx.x, b.mx.y, b.mx.z, ibits((int32_t)ref));
    return e;
}

// ── energy LUT (EnergyCompensationLut.cs:34-130) ─────────────────────────
static f3 vndf_sample(f3 v, float ax, float ay, float u1, float u2) {
    f3 vh = normalize(mk(ax * v.x, ay * v.y, v.z));
    float l2 = vh.x * vh.x + vh.y * vh.y;
    f3 t1 = l2 > 0.0f ? mk(-vh.y, vh.x, 0.0f) / sqrtf(l2) : mk(1, 0, 0);
    f3 t2 = cross(vh, t1);
    float r = sqrtf(u1);
    float phi = 2.0f * kPi * u2;
    float a = r * cosf(phi);
    float b = r * sinf(phi);
    float s = 0.5f * (1.0f + vh.z);
    b = (1.0f - s) * sqrtf(mx(0.0f, 1.0f - a * a)) + s * b;
    f3 nh = a * t1 + b * t2 + sqrtf(mx(0.0f, 1.0f - a * a - b * b)) * vh;
    return normalize(mk(ax * nh.x, ay * nh.y, mx(0.0f, nh.z)));
}
static float g1_aniso(f3 w, float ax, float ay) {
    float c2 = w.z * w.z;
    float lam;
    if (c2 < 1e-14f) lam = 1e10f;
    else {
        float num = (ax * w.x) * (ax * w.x) + (ay * w.y) * (ay * w.y);
        lam = 0.5f * (-1.0f + sqrtf(1.0f + num / c2));
    }
    return 1.0f / (1.0f + lam);
}

void build_energy_lut(float* E, float* Eavg) {
    auto mu_v = [](int i) { return ((float)i + 0.5f) / 32.0f; };
    for (int ai = 0; ai < 32; ai++) {
        float alpha = mx(((float)ai + 0.5f) / 32.0f, 0.001f);
        for (int mi = 0; mi < 32; mi++) {
            float mu = mu_v(mi);
            uint32_t state = ((uint32_t)mi * 0x9E3779B9u) ^ ((uint32_t)ai * 0xB5297A4Du) ^ 0xC2B2AE3Du;
            if (state == 0) state = 1;
            float st = sqrtf(mx(0.0f, 1.0f - mu * mu));
            f3 v = mk(st, 0.0f, mu);
            double sum = 0.0;
            int hits = 0;
            for (int s = 0; s < 2048; s++) {
                state ^= state << 13; state ^= state >> 17; state ^= state << 5;
                float u1 = (float)(state & 0x00FFFFFFu) / (float)0x01000000u;
                state ^= state << 13; state ^= state >> 17; state ^= state << 5;
                float u2 = (float)(state & 0x00FFFFFFu) / (float)0x01000000u;
                f3 h = vndf_sample(v, alpha, alpha, u1, u2);
                float vdh = dot(v, h);
                f3 l = 2.0f * vdh * h - v;
                if (l.z <= 0.0f) continue;
                sum += (double)g1_aniso(l, alpha, alpha);
                hits++;
            }
            double e = hits == 0 ? 0.0 : sum / 2048.0;
            E[mi * 32 + ai] = (float)(e < 0.0 ? 0.0 : (e > 1.0 ? 1.0 : e));
        }
        float avg = 0.0f;
        for (int mi = 0; mi < 32; mi++) avg += 2.0f * E[mi * 32 + ai] * mu_v(mi);
        Eavg[ai] = clampf(avg / 32.0f, 0.0f, 1.0f);
    }
}

float lut_sample_eavg(const float* Eavg, float alpha) { return eavg_lookup(Eavg, alpha); }

// ── materials (DisneyBsdf.cs:213-293, 1016-1049, 1097-1139, 1943-1959) ──
static GMat make_gmat(const rt_material& m, const float* Eavg) {
    GMat g{};
    f3 c = ld3(m.color);
    int32_t flags = 0;
    g.m0 = make_int4(m.kind, 0, -1, 0);
    switch (m.kind) {
        case RT_MAT_LAMBERTIAN:
            g.m1 = make_float4(c.x, c.y, c.z, 0.0f);
            if (m.flags & RT_MATF_IMAGE) flags |= kMatImage;
            break;
        case RT_MAT_EMISSIVE: {
            float inten = mx(m.intensity, 0.0f);
            f3 e = c * inten;
            g.m1 = make_float4(e.x, e.y, e.z, 0.0f);
            if (m.flags & RT_MATF_LIGHT_PROXY) flags |= kMatEmissiveProxy;
            break;
        }
        case RT_MAT_DIELECTRIC:
            g.m1 = make_float4(c.x, c.y, c.z, 0.0f);
            g.m3 = make_float4(0.0f, 0.0f, 0.0f, m.ior);
            flags |= kMatSpecTransmissive | kMatDielectricIor;
            break;
        default:
            return disney_gmat(m, c, m.metallic, m.roughness, Eavg);
    }
    g.m0.y = flags;
    return g;
}

// ── any-hit BVH: 4-wide collapse of every binary tree ──────────────────
static float half_area(const Box& b) {
    f3 e = b.mx - b.mn;
    return e.x * e.y + e.y * e.z + e.z * e.x;
}
static Box child_box(const GNode& n, int k) {
    Box b;
    if (k == 0) {
        b.mn = mk(n.a.x, n.a.y, n.a.z);
        b.mx = mk(n.a.w, n.b.x, n.b.y);
    } else {
        b.mn = mk(n.b.z, n.b.w, n.c.x);
        b.mx = mk(n.c.y, n.c.z, n.c.w);
    }
    return b;
}
// Returns the any-hit reference of binary reference `ref` and, in `need`,
// the stack entries a traversal below it can hold (3 siblings per level,
// one leaf continuation per nested BVH entry).
static uint32_t collapse4(HostScene& hs, uint32_t ref, int& need) {
    if (ref_is_leaf(ref)) {
        need = 0;
        uint32_t st = leaf_start(ref), cnt = leaf_count(ref);
        for (uint32_t k = 0; k < cnt; k++) {
            GEntry& e = hs.entries[st + k];
            if ((__builtin_bit_cast(int32_t, e.a.w) & 0xff) != kBvh) continue;
            int sub = 0;
            uint32_t r4 = collapse4(hs, (uint32_t)__builtin_bit_cast(int32_t, e.b.w), sub);
            e.c.x = __builtin_bit_cast(float, (int32_t)r4);
            need = std::max(need, (k + 1 < cnt ? 1 : 0) + sub);
        }
        return ref;
    }
    struct C {
        uint32_t ref;
        Box box;
    };
    std::vector<C> ch;
    const GNode n = hs.nodes[ref];
    ch.push_back({(uint32_t)n.d.x, child_box(n, 0)});
    ch.push_back({(uint32_t)n.d.y, child_box(n, 1)});
    while (ch.size() < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int i = 0; i < (int)ch.size(); i++)
            if (!ref_is_leaf(ch[i].ref) && half_area(ch[i].box) > ba) {
                ba = half_area(ch[i].box);
                best = i;
            }
        if (best < 0) break;
        const GNode m = hs.nodes[ch[best].ref];
        ch[best] = {(uint32_t)m.d.x, child_box(m, 0)};
        ch.insert(ch.begin() + best + 1, C{(uint32_t)m.d.y, child_box(m, 1)});
    }
    const uint32_t idx = (uint32_t)hs.nodes4.size();
    hs.nodes4.push_back(GNode4{});
    float mnx[4], mny[4], mnz[4], mxx[4], mxy[4], mxz[4];
    int32_t cr[4];
    int deepest = 0;
    for (int k = 0; k < 4; k++) {
        if (k < (int)ch.size()) {
            int sub = 0;
            cr[k] = (int32_t)collapse4(hs, ch[k].ref, sub);
            deepest = std::max(deepest, sub);
            mnx[k] = ch[k].box.mn.x; mny[k] = ch[k].box.mn.y; mnz[k] = ch[k].box.mn.z;
            mxx[k] = ch[k].box.mx.x; mxy[k] = ch[k].box.mx.y; mxz[k] = ch[k].box.mx.z;
        } else {
            // unused slot: a point box far outside any scene, empty leaf
            cr[k] = (int32_t)leaf_ref(0, 0);
            mnx[k] = mny[k] = mnz[k] = 1e30f;
            mxx[k] = mxy[k] = mxz[k] = 1e30f;
        }
    }
    GNode4& g = hs.nodes4[idx];
    g.mnx = make_float4(mnx[0], mnx[1], mnx[2], mnx[3]);
    g.mny = make_float4(mny[0], mny[1], mny[2], mny[3]);
    g.mnz = make_float4(mnz[0], mnz[1], mnz[2], mnz[3]);
    g.mxx = make_float4(mxx[0], mxx[1], mxx[2], mxx[3]);
    g.mxy = make_float4(mxy[0], mxy[1], mxy[2], mxy[3]);
    g.mxz = make_float4(mxz[0], mxz[1], mxz[2], mxz[3]);
    g.ch = make_int4(cr[0], cr[1], cr[2], cr[3]);
    g.pad = make_int4(0, 0, 0, 0);
    need = (int)ch.size() - 1 + deepest;
    return idx;
}
static void build_any_bvh(HostScene& hs) {
    hs.nodes4.clear();
    int need = 0;
    for (uint32_t t = 0; t < hs.top_count; t++) {
        GEntry& e = hs.entries[hs.top_start + t];
        if ((__builtin_bit_cast(int32_t, e.a.w) & 0xff) != kBvh) continue;
        int sub = 0;
        uint32_t r4 = collapse4(hs, (uint32_t)__builtin_bit_cast(int32_t, e.b.w), sub);
        e.c.x = __builtin_bit_cast(float, (int32_t)r4);
        need = std::max(need, sub);
    }
    hs.any_stack = need + 1;
}

// Flat objects of the primitive list (each primitive, or a mesh's faces).
struct Flat {
    const std::vector<int32_t>& first;
    const std::vector<int32_t>& count;
    const std::vector<Box>& box;
    const std::vector<char>& plane;
    const std::vector<int>& need;
};

// Object hierarchy (rt_node; Group.cs:20-80, Transform.cs:74-240,
// Instance.cs:54-84, CameraInvisibleHittable): one entry per node, built
// once however many parents share it.  Outputs the roots' entries, boxes,
// world-list plane flags and stack needs.
static int build_hierarchy(const rt_scene_desc* d, HostScene& hs, Flattener& F, const Flat& flat,
                           std::vector<GEntry>& r_entry, std::vector<Box>& r_box, std::vector<char>& r_plane,
                           std::vector<int>& r_need, std::string& err) {
    const int64_t nn = d->num_nodes, np = d->num_primitives;
    if (nn <= 0 || !d->nodes || d->num_roots <= 0 || !d->roots || d->num_node_children < 0 ||
        (d->num_node_children > 0 && !d->node_children)) {
        err = "bad node tables";
        return RT_ERR_INVALID_ARGUMENT;
    }
    hs.hier = true;
    std::vector<int32_t> obj_of(np, -1);
    for (size_t o = 0; o < flat.first.size(); o++) obj_of[flat.first[o]] = (int32_t)o;
    struct Out {
        GEntry e;
        Box box;
        bool plane_world = false, plane_group = false;
        int wdepth = 0, need = 0;
        int32_t standalone = -1;     // index of a copy of `e` that a wrapper can name
    };
    std::vector<Out> out(nn);
    std::vector<int> state(nn, 0);
    auto standalone = [&](int32_t n) {
        if (out[n].standalone < 0) {
            out[n].standalone = (int32_t)hs.entries.size();
            hs.entries.push_back(out[n].e);
        }
        return out[n].standalone;
    };
    auto wrap_entry = [&](int32_t child_entry, const Box& b, int32_t xf, int32_t mat, int32_t flags) {
        GEntry e{};
        e.a = make_float4(b.mn.x, b.mn.y, b.mn.z, ibits(kWrap | (flags << 8)));
        e.b = make_float4(b.mx.x, b.mx.y, b.mx.z, ibits(child_entry));
        e.c = make_float4(ibits(xf), ibits(mat), 0.0f, 0.0f);
        return e;
    };
    std::function<int(int32_t)> make = [&](int32_t n) -> int {
        if (n < 0 || n >= nn) { err = "node index out of range"; return RT_ERR_INVALID_ARGUMENT; }
        if (state[n] == 2) return RT_OK;
        if (state[n] == 1) { err = "node cycle"; return RT_ERR_INVALID_ARGUMENT; }
        state[n] = 1;
        const rt_node& nd = d->nodes[n];
        Out& o = out[n];
        const bool cam_inv = (nd.flags & RT_NODEF_CAMERA_INVISIBLE) != 0;
        switch (nd.kind) {
            case RT_NODE_PRIMITIVES: {
                int32_t ob = (nd.first >= 0 && nd.first < np) ? obj_of[nd.first] : -1;
                if (ob < 0 || flat.count[ob] != nd.count) {
                    err = "a PRIMITIVES node must name one whole object (a primitive or a mesh)";
                    return RT_ERR_INVALID_ARGUMENT;
                }
                o.e = F.obj_entry[ob];
                o.box = flat.box[ob];
                o.plane_world = o.plane_group = flat.plane[ob] != 0;
                o.need = flat.need[ob];
                break;
            }
            case RT_NODE_GROUP: {
                if (nd.count <= 0 || nd.first < 0 || (int64_t)nd.first + nd.count > d->num_node_children) {
                    err = "bad group child range";
                    return RT_ERR_INVALID_ARGUMENT;
                }
                std::vector<int32_t> fin, inf;
                bool all_planes = true;
                for (int32_t k = 0; k < nd.count; k++) {
                    const int32_t c = d->node_children[nd.first + k];
                    int rc = make(c);
                    if (rc != RT_OK) return rc;
                    (out[c].plane_group ? inf : fin).push_back(c);
                    all_planes = all_planes && out[c].plane_world;
                    o.wdepth = std::max(o.wdepth, out[c].wdepth);
                }
                std::vector<GEntry> re;
                std::vector<Box> rb;
                std::vector<int> rn;
                if (fin.size() > 4) {
                    std::vector<GEntry> ce(fin.size());
                    std::vector<int32_t> items(fin.size());
                    std::vector<Box> bx(fin.size());
                    int cneed = 0;
                    for (size_t k = 0; k < fin.size(); k++) {
                        ce[k] = out[fin[k]].e;
                        items[k] = (int32_t)k;
                        bx[k] = out[fin[k]].box;
                        cneed = std::max(cneed, out[fin[k]].need);
                    }
                    auto root = build_node(items, bx, 0, (int)items.size(), 0);
                    std::vector<GEntry> saved;
                    saved.swap(F.obj_entry);
                    F.obj_entry = ce;
                    F.deepest = 0;
                    uint32_t ref = F.flatten(root.get(), items, false, 0);
                    F.obj_entry.swap(saved);
                    re.push_back(bvh_entry(root->box, ref, 0));
                    rb.push_back(root->box);
                    rn.push_back(F.deepest + 1 + cneed);
                } else {
                    for (int32_t c : fin) { re.push_back(out[c].e); rb.push_back(out[c].box); rn.push_back(out[c].need); }
                }
                for (int32_t c : inf) { re.push_back(out[c].e); rb.push_back(out[c].box); rn.push_back(out[c].need); }
                if (re.size() == 1) {
                    o.e = re[0];
                    o.box = rb[0];
                    o.need = rn[0];
                } else {
                    if (re.size() > 15) { err = "a group's list part holds more than 15 members"; return RT_ERR_UNSUPPORTED; }
                    uint32_t start = (uint32_t)hs.entries.size();
                    Box b = rb[0];
                    int mneed = 0;
                    for (size_t k = 0; k < re.size(); k++) {
                        hs.entries.push_back(re[k]);
                        if (k) b = box_union(b, rb[k]);
                        mneed = std::max(mneed, rn[k]);
                    }
                    o.e = bvh_entry(b, leaf_ref(start, (uint32_t)re.size()), kEntryNoBox);
                    o.box = b;
                    o.need = 1 + mneed;
                }
                o.plane_world = all_planes;
                o.plane_group = false;
                break;
            }
            case RT_NODE_TRANSFORM:
            case RT_NODE_INSTANCE: {
                int rc = make(nd.child);
                if (rc != RT_OK) return rc;
                const Out& c = out[nd.child];
                int32_t xi = -1, mat = -1;
                if (nd.kind == RT_NODE_TRANSFORM) {
                    xi = (int32_t)(hs.xf.size() / 8);
                    for (int r = 0; r < 4; r++)
                        hs.xf.push_back(make_float4(nd.matrix[4 * r], nd.matrix[4 * r + 1], nd.matrix[4 * r + 2], 0.0f));
                    for (int r = 0; r < 4; r++)
                        hs.xf.push_back(make_float4(nd.inverse[4 * r], nd.inverse[4 * r + 1], nd.inverse[4 * r + 2], 0.0f));
                    for (int k = 0; k < 16; k++)
                        if (!std::isfinite(nd.matrix[k]) || !std::isfinite(nd.inverse[k])) {
                            err = "non-finite transform";
                            return RT_ERR_INVALID_ARGUMENT;
                        }
                    // ComputeWorldBox (Transform.cs:211-240)
                    const float4* M = &hs.xf[8 * xi];
                    f3 mn = splat(FLT_MAX), mxv = splat(-FLT_MAX);
                    for (int k = 0; k < 8; k++) {
                        f3 cr = mk((k & 4) ? c.box.mx.x : c.box.mn.x, (k & 2) ? c.box.mx.y : c.box.mn.y,
                                   (k & 1) ? c.box.mx.z : c.box.mn.z);
                        f3 t = xf_point(cr, M);
                        mn = mk(fminf(mn.x, t.x), fminf(mn.y, t.y), fminf(mn.z, t.z));
                        mxv = mk(fmaxf(mxv.x, t.x), fmaxf(mxv.y, t.y), fmaxf(mxv.z, t.z));
                    }
                    o.box = Box{mn, mxv};
                    o.plane_world = c.plane_world;
                    o.plane_group = c.plane_group && !cam_inv;
                } else {
                    if (nd.material >= d->num_materials) { err = "bad override material"; return RT_ERR_INVALID_ARGUMENT; }
                    mat = nd.material < 0 ? -1 : nd.material;
                    o.box = c.box;
                }
                o.e = wrap_entry(standalone(nd.child), o.box, xi, mat, cam_inv ? kEntryCamInvisible : 0);
                o.need = 1 + c.need;
                o.wdepth = 1 + c.wdepth;
                break;
            }
            default: err = "bad node kind"; return RT_ERR_INVALID_ARGUMENT;
        }
        if (cam_inv && nd.kind != RT_NODE_TRANSFORM && nd.kind != RT_NODE_INSTANCE) {
            // CameraInvisibleHittable around a primitive object or a group
            const bool pw = o.plane_world;
            const int32_t inner = (int32_t)hs.entries.size();
            hs.entries.push_back(o.e);
            o.e = wrap_entry(inner, o.box, -1, -1, kEntryCamInvisible);
            o.need += 1;
            o.wdepth += 1;
            o.plane_world = pw;
            o.plane_group = false;
        }
        if (o.wdepth > kMaxWrap) { err = "wrapper nesting deeper than 4"; return RT_ERR_UNSUPPORTED; }
        state[n] = 2;
        return RT_OK;
    };
    for (int64_t r = 0; r < d->num_roots; r++) {
        int rc = make(d->roots[r]);
        if (rc != RT_OK) return rc;
        const Out& o = out[d->roots[r]];
        r_entry.push_back(o.e);
        r_box.push_back(o.box);
        r_plane.push_back(o.plane_world ? 1 : 0);
        r_need.push_back(o.need);
    }
    // geometry lights sample world-space primitives: none may sit under a transform
    std::vector<char> seen(2 * nn, 0), under(np, 0);
    std::function<void(int32_t, bool)> mark = [&](int32_t n, bool x) {
        if (seen[2 * n + x]) return;
        seen[2 * n + x] = 1;
        const rt_node& nd = d->nodes[n];
        if (nd.kind == RT_NODE_PRIMITIVES) {
            if (x)
                for (int32_t k = 0; k < nd.count; k++) under[nd.first + k] = 1;
        } else if (nd.kind == RT_NODE_GROUP) {
            for (int32_t k = 0; k < nd.count; k++) mark(d->node_children[nd.first + k], x);
        } else mark(nd.child, x || nd.kind == RT_NODE_TRANSFORM);
    };
    for (int64_t r = 0; r < d->num_roots; r++) mark(d->roots[r], false);
    for (int64_t i = 0; i < d->num_lights; i++) {
        const rt_light& l = d->lights[i];
        if (l.kind == RT_LIGHT_GEOMETRY && l.primitive >= 0 && l.primitive < np && under[l.primitive]) {
            err = "a geometry light on a transformed primitive";
            return RT_ERR_UNSUPPORTED;
        }
    }
    for (int64_t i = 0; i < d->num_materials; i++)
        if (d->materials[i].flags & RT_MATF_IMAGE) {
            err = "image textures in a scene with an object hierarchy";
            return RT_ERR_UNSUPPORTED;
        }
    return RT_OK;
}

int build_host_scene(const rt_scene_desc* d, HostScene& hs, std::string& err) {
    if (!d) { err = "desc is null"; return RT_ERR_INVALID_ARGUMENT; }
    if (d->num_primitives < 0 || d->num_materials < 0 || d->num_lights < 0 ||
        (d->num_primitives > 0 && !d->primitives) || (d->num_materials > 0 && !d->materials) ||
        (d->num_lights > 0 && !d->lights)) {
        err = "invalid table pointer or count";
        return RT_ERR_INVALID_ARGUMENT;
    }
    const int64_t np = d->num_primitives;
    if (np >= (int64_t)kLeafStartMask) { err = "too many primitives"; return RT_ERR_UNSUPPORTED; }
    hs = HostScene();
    hs.num_prims = np;
    build_energy_lut(hs.lut_e, hs.lut_eavg);

    // materials
    hs.mats.resize(d->num_materials);
    for (int64_t i = 0; i < d->num_materials; i++) {
        int k = d->materials[i].kind;
        if (k < RT_MAT_LAMBERTIAN || k > RT_MAT_DISNEY) { err = "bad material kind"; return RT_ERR_INVALID_ARGUMENT; }
        if (k == RT_MAT_DISNEY && d->materials[i].thin_film_thickness > 0.0f) hs.thin_film = true;
        if (k == RT_MAT_DISNEY && d->materials[i].thin_film_thickness > 0.0f &&
            (d->materials[i].flags & RT_MATF_TEXTURED)) {
            // the texture adjoint's derivatives cover the Schlick Fresnel only
            err = "a thin-film Disney material cannot take the texture set";
            return RT_ERR_UNSUPPORTED;
        }
        if (d->materials[i].flags & RT_MATF_TEXTURED) {
            const rt_material& m = d->materials[i];
            if (k != RT_MAT_DISNEY) { err = "only Disney materials can be textured"; return RT_ERR_UNSUPPORTED; }
            if (m.spec_trans != 0.0f || m.diff_trans != 0.0f) {
                err = "textured Disney materials need spec_trans = 0 and diff_trans = 0";
                return RT_ERR_UNSUPPORTED;
            }
            if (!d->texels || d->tex_width <= 0 || d->tex_height <= 0 ||
                (int64_t)d->tex_width * d->tex_height > (1ll << 28)) {
                err = "textured material without a valid texture set";
                return RT_ERR_INVALID_ARGUMENT;
            }
            hs.textured = true;
        }
        hs.mats[i] = make_gmat(d->materials[i], hs.lut_eavg);
        // ShadowTransmittance is non-zero only for Dielectric and for Disney
        // with spec_trans > 0 (Dielectric.cs:46-54, DisneyBsdf.cs:570-571)
        if (k == RT_MAT_DIELECTRIC || (k == RT_MAT_DISNEY && d->materials[i].spec_trans > 0.0f))
            hs.opaque_shadows = false;
    }
    hs.raw_mats.assign(d->materials, d->materials + d->num_materials);
    int rc_img = build_images(d, hs, err);
    if (rc_img != RT_OK) return rc_img;
    bool any_image = false;
    for (int64_t i = 0; i < d->num_materials; i++) any_image |= (d->materials[i].flags & RT_MATF_IMAGE) != 0;
    if (any_image && d->texcoords && !hs.textured) {
        hs.tex_uv.resize(2 * (size_t)np);
        for (int64_t i = 0; i < np; i++) {
            const float* t = d->texcoords + 6 * i;
            hs.tex_uv[2 * i] = make_float4(t[0], t[1], t[2], t[3]);
            hs.tex_uv[2 * i + 1] = make_float4(t[4], t[5], 0.0f, 0.0f);
        }
    }
    if (hs.textured) {
        hs.tex_w = d->tex_width;
        hs.tex_h = d->tex_height;
        hs.texels.assign(d->texels, d->texels + (size_t)kTexChannels * d->tex_width * d->tex_height);
        if (d->texcoords) {
            hs.tex_uv.resize(2 * (size_t)np);
            for (int64_t i = 0; i < np; i++) {
                const float* t = d->texcoords + 6 * i;
                hs.tex_uv[2 * i] = make_float4(t[0], t[1], t[2], t[3]);
                hs.tex_uv[2 * i + 1] = make_float4(t[4], t[5], 0.0f, 0.0f);
            }
        }
    }

    // smooth triangles (SmoothTriangle.cs:68-130): unit vertex normals and
    // the tangent frame of the vertex UVs, barycentric UVs without texcoords
    if (d->normals) {
        hs.smooth.assign(4 * (size_t)np, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        bool any = false;
        for (int64_t i = 0; i < np; i++) {
            const float* n = d->normals + 9 * i;
            bool smooth = false;
            for (int k = 0; k < 9; k++) smooth |= n[k] != 0.0f;
            if (!smooth) continue;
            const rt_primitive& p = d->primitives[i];
            if (p.kind != RT_PRIM_TRIANGLE) { err = "vertex normals on a primitive that is not a triangle"; return RT_ERR_INVALID_ARGUMENT; }
            f3 nv[3];
            for (int k = 0; k < 3; k++) {
                f3 a = ld3(n + 3 * k);
                float l2 = length_sq(a);
                if (!(l2 > 0.0f) || !std::isfinite(l2)) { err = "smooth triangle with a zero or non-finite vertex normal"; return RT_ERR_INVALID_ARGUMENT; }
                nv[k] = normalize(a);
            }
            static const float kBaryUv[6] = {0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 1.0f};
            const float* uv = d->texcoords ? d->texcoords + 6 * i : kBaryUv;
            f3 v0 = ld3(p.v), e1 = ld3(p.v + 3) - v0, e2 = ld3(p.v + 6) - v0;
            f3 T, B;
            uv_frame(e1, e2, uv, T, B);
            float4* o = &hs.smooth[4 * i];
            o[0] = make_float4(nv[0].x, nv[0].y, nv[0].z, T.x);
            o[1] = make_float4(nv[1].x, nv[1].y, nv[1].z, T.y);
            o[2] = make_float4(nv[2].x, nv[2].y, nv[2].z, T.z);
            o[3] = make_float4(B.x, B.y, B.z, 1.0f);
            any = true;
        }
        if (!any) hs.smooth.clear();
    }

    // objects (SceneLoader object list; a mesh is one object)
    Flattener F{hs, d, {}, {}};
    F.quad_aux_of.assign(np, -1);
    std::vector<Box> pbox(np);
    std::vector<int32_t> obj_first, obj_count;   // primitive range per object
    std::vector<int32_t> obj_mesh;
    int last_mesh = -1;
    for (int64_t i = 0; i < np; i++) {
        const rt_primitive& p = d->primitives[i];
        if (p.kind < RT_PRIM_TRIANGLE || p.kind > RT_PRIM_PLANE) { err = "bad primitive kind"; return RT_ERR_INVALID_ARGUMENT; }
        if (p.material < 0 || p.material >= d->num_materials) { err = "bad material index"; return RT_ERR_INVALID_ARGUMENT; }
        if (p.mesh >= 0 && p.kind != RT_PRIM_TRIANGLE) { err = "mesh faces must be triangles"; return RT_ERR_INVALID_ARGUMENT; }
        pbox[i] = prim_box(p);
        if (p.kind == RT_PRIM_QUAD) {
            f3 u = ld3(p.v + 3), w = ld3(p.v + 6);
            F.quad_aux_of[i] = (int32_t)(hs.quad_aux.size() / 4);
            f3 tn = normalize(u), bn = normalize(w);
            hs.quad_aux.push_back(make_float4(u.x, u.y, u.z, length(cross(u, w))));
            hs.quad_aux.push_back(make_float4(w.x, w.y, w.z, 0.0f));
            hs.quad_aux.push_back(make_float4(tn.x, tn.y, tn.z, 0.0f));
            hs.quad_aux.push_back(make_float4(bn.x, bn.y, bn.z, 0.0f));
        }
        if (p.mesh >= 0 && p.mesh == last_mesh) {
            obj_count.back()++;
        } else {
            obj_first.push_back((int32_t)i);
            obj_count.push_back(1);
            obj_mesh.push_back(p.mesh);
            last_mesh = p.mesh;
        }
    }
    const int64_t nobj = (int64_t)obj_first.size();
    hs.num_objects = nobj;
    std::vector<Box> obox(nobj);
    std::vector<char> is_plane(nobj, 0);
    std::vector<int> obj_need(nobj, 0);        // traversal stack entries below the object's entry
    F.obj_entry.resize(nobj);
    for (int64_t o = 0; o < nobj; o++) {
        int32_t f = obj_first[o];
        if (obj_mesh[o] < 0) {
            obox[o] = pbox[f];
            is_plane[o] = d->primitives[f].kind == RT_PRIM_PLANE;
            F.obj_entry[o] = F.prim_entry(f);
            continue;
        }
        // Mesh (Mesh.cs:75-116): nested BVH over its faces when > 4, else a list
        int32_t cnt = obj_count[o];
        std::vector<int32_t> items(cnt);
        std::vector<Box> bx(cnt);
        for (int k = 0; k < cnt; k++) { items[k] = f + k; bx[k] = pbox[f + k]; }
        if (cnt > 4) {
            auto root = build_node(items, bx, 0, cnt, 0);
            F.deepest = 0;
            uint32_t ref = F.flatten(root.get(), items, true, 0);
            hs.mesh_depth = std::max(hs.mesh_depth, F.deepest);
            obj_need[o] = F.deepest;
            hs.mesh_refs.push_back(ref);
            obox[o] = root->box;
            F.obj_entry[o] = bvh_entry(root->box, ref, 0);
        } else {
            uint32_t start = (uint32_t)hs.entries.size();
            hs.mesh_depth = std::max(hs.mesh_depth, 0);
            hs.mesh_refs.push_back(leaf_ref(start, (uint32_t)cnt));
            Box b = bx[0];
            for (int k = 0; k < cnt; k++) { hs.entries.push_back(F.prim_entry(items[k])); if (k) b = box_union(b, bx[k]); }
            obox[o] = b;
            F.obj_entry[o] = bvh_entry(b, leaf_ref(start, (uint32_t)cnt), kEntryNoBox);
        }
    }

    // World object list: the flat objects, or the roots of the hierarchy
    std::vector<GEntry> wl_entry = F.obj_entry;
    std::vector<Box> wl_box = obox;
    std::vector<char> wl_plane = is_plane;
    std::vector<int> wl_need = obj_need;
    if (d->num_nodes != 0 || d->num_roots != 0) {
        Flat flat{obj_first, obj_count, obox, is_plane, obj_need};
        wl_entry.clear();
        wl_box.clear();
        wl_plane.clear();
        wl_need.clear();
        int rc = build_hierarchy(d, hs, F, flat, wl_entry, wl_box, wl_plane, wl_need, err);
        if (rc != RT_OK) return rc;
        F.obj_entry = wl_entry;
        hs.num_objects = (int64_t)wl_entry.size();
    }
    const int64_t nw = (int64_t)wl_entry.size();

    // World (SceneLoader.cs:460-487)
    std::vector<GEntry> top;
    std::vector<Box> top_boxes;
    int world_need = 0;
    if (nw > 4) {
        std::vector<int32_t> fin;
        std::vector<Box> fbx;
        std::vector<int32_t> inf;
        int fin_need = 0;
        for (int64_t o = 0; o < nw; o++) {
            if (wl_plane[o]) {
                inf.push_back((int32_t)o);
                world_need = std::max(world_need, wl_need[o]);
            } else {
                fin.push_back((int32_t)o);
                fbx.push_back(wl_box[o]);
                fin_need = std::max(fin_need, wl_need[o]);
            }
        }
        if (!fin.empty()) {
            auto root = build_node(fin, fbx, 0, (int)fin.size(), 0);
            F.deepest = 0;
            uint32_t ref = F.flatten(root.get(), fin, false, 0);
            hs.world_depth = F.deepest;
            hs.world_root = ref;
            hs.world_box = root->box;
            top.push_back(bvh_entry(root->box, ref, 0));
            top_boxes.push_back(root->box);
            hs.has_bvh = true;
            world_need = std::max(world_need, F.deepest + 1 + fin_need);
        }
        for (int32_t o : inf) { top.push_back(wl_entry[o]); top_boxes.push_back(wl_box[o]); }
    } else {
        for (int64_t o = 0; o < nw; o++) {
            top.push_back(wl_entry[o]);
            top_boxes.push_back(wl_box[o]);
            world_need = std::max(world_need, wl_need[o]);
        }
    }
    hs.top_start = (uint32_t)hs.entries.size();
    hs.top_count = (uint32_t)top.size();
    for (auto& e : top) hs.entries.push_back(e);
    if (hs.entries.size() >= kLeafStartMask) { err = "scene too large"; return RT_ERR_UNSUPPORTED; }
    hs.hier_stack = world_need;
    if (!hs.hier) build_any_bvh(hs);

    // Camera basis (Camera.cs:26-51)
    {
        const rt_camera& c = d->camera;
        f3 lf = ld3(c.look_from), la = ld3(c.look_at), vup = ld3(c.vup);
        float theta = c.vfov_deg * kPi / 180.0f;
        float h = tanf(theta / 2.0f);
        float vh = 2.0f * h, vw = c.aspect * vh;
        f3 w = normalize(lf - la);
        if (length_sq(cross(w, vup)) < 1e-8f) vup = fabsf(w.x) < 0.9f ? mk(1, 0, 0) : mk(0, 0, 1);
        f3 u = normalize(cross(vup, w));
        f3 v = cross(w, u);
        hs.cam_origin = lf;
        hs.cam_horiz = c.focus_dist * vw * u;
        hs.cam_vert = c.focus_dist * vh * v;
        hs.cam_llc = lf - hs.cam_horiz / 2.0f - hs.cam_vert / 2.0f - c.focus_dist * w;
        hs.cam_u = u;
        hs.cam_v = v;
        hs.lens_radius = c.aperture / 2.0f;
    }
    hs.sky = ld3(d->sky_color);
    if (d->global_medium.enabled) {
        hs.medium = true;
        hs.sig_a = vmax(ld3(d->global_medium.sigma_a), splat(0.0f));
        hs.sig_s = vmax(ld3(d->global_medium.sigma_s), splat(0.0f));
        hs.sig_t = hs.sig_a + hs.sig_s;
    }

    // Finite scene bounds (Renderer.cs:476-502)
    {
        Box acc = box_empty();
        bool any = false;
        for (const Box& b : top_boxes) {
            f3 dd = b.mx - b.mn;
            if (dd.x < 1e5f && dd.y < 1e5f && dd.z < 1e5f) { acc = any ? box_union(acc, b) : b; any = true; }
        }
        if (!any) {
            Box raw = box_empty();
            for (size_t k = 0; k < top_boxes.size(); k++) raw = k ? box_union(raw, top_boxes[k]) : top_boxes[k];
            acc = Box{vmax(raw.mn, splat(-1e3f)), vmin(raw.mx, splat(1e3f))};
        }
        hs.bounds = acc;
    }

    // Lights
    float flux = 0.0f;
    hs.lights.resize(d->num_lights);
    hs.light_prim.assign(d->num_lights, -1);
    hs.light_power.assign(d->num_lights, 0.0f);
    for (int64_t i = 0; i < d->num_lights; i++) {
        const rt_light& l = d->lights[i];
        GLight& g = hs.lights[i];
        g = GLight{};
        f3 col = ld3(l.color);
        float power = 0.0f;
        int ss = 1;
        switch (l.kind) {
            case RT_LIGHT_SPHERE: {
                float r = mx(l.radius, kEps);
                ss = std::max(1, l.shadow_samples);
                g.l1 = make_float4(col.x, col.y, col.z, l.intensity);
                g.l2 = make_float4(l.position[0], l.position[1], l.position[2], r);
                power = 4.0f * kPi * luminance(col) * l.intensity;
                break;
            }
            case RT_LIGHT_AREA: {
                f3 u = ld3(l.u), v = ld3(l.v);
                f3 cr = cross(u, v);
                float area = length(cr);
                f3 n = cr / area;
                ss = std::max(1, l.shadow_samples);
                g.l1 = make_float4(col.x, col.y, col.z, l.intensity);
                g.l2 = make_float4(l.position[0], l.position[1], l.position[2], area);
                g.l3 = make_float4(u.x, u.y, u.z, mx(0.0f, l.soft_radius));
                g.l4 = make_float4(v.x, v.y, v.z, 0.0f);
                g.l5 = make_float4(n.x, n.y, n.z, area);
                power = kPi * area * luminance(col) * l.intensity;
                break;
            }
            case RT_LIGHT_DIRECTIONAL: {
                f3 dir = normalize(ld3(l.position));
                float ang = mx(0.0f, l.radius);
                bool disc = ang > 0.0f;
                float rad = ang * kPi / 180.0f;
                float cosa = disc ? cosf(rad) : 1.0f;
                float solid = disc ? 2.0f * kPi * (1.0f - cosa) : 0.0f;
                ss = l.shadow_samples > 0 ? l.shadow_samples : (disc ? 4 : 1);
                g.l1 = make_float4(col.x, col.y, col.z, l.intensity);
                g.l2 = make_float4(dir.x, dir.y, dir.z, ang);
                g.l3 = make_float4(cosa, solid, 0.0f, 0.0f);
                f3 e = hs.bounds.mx - hs.bounds.mn;
                float rr = 0.5f * length(e);
                power = luminance(col) * l.intensity * (kPi * rr * rr);
                break;
            }
            case RT_LIGHT_GEOMETRY: {
                if (l.primitive < 0 || l.primitive >= np) { err = "geometry light primitive out of range"; return RT_ERR_INVALID_ARGUMENT; }
                const rt_primitive& p = d->primitives[l.primitive];
                const rt_material& pm = d->materials[p.material];
                if (pm.kind != RT_MAT_EMISSIVE || p.kind == RT_PRIM_PLANE || p.mesh >= 0) {
                    err = "geometry light needs a standalone emissive triangle, quad or sphere";
                    return RT_ERR_UNSUPPORTED;
                }
                if (!hs.smooth.empty() && hs.smooth[4 * (size_t)l.primitive + 3].w != 0.0f) {
                    err = "a smooth triangle cannot be a geometry light";
                    return RT_ERR_UNSUPPORTED;
                }
                ss = std::max(1, l.shadow_samples);
                f3 em = ld3(pm.color) * mx(pm.intensity, 0.0f);
                g.l1 = make_float4(em.x, em.y, em.z, 0.0f);
                float area;
                const float* v = p.v;
                if (p.kind == RT_PRIM_TRIANGLE) {
                    f3 v0 = ld3(v), e1 = ld3(v + 3) - v0, e2 = ld3(v + 6) - v0;
                    f3 cr = cross(e1, e2);
                    float cl = length(cr);
                    f3 n = cl > 0.0f ? cr / cl : mk(0, 1, 0);
                    area = 0.5f * cl;
                    g.l0.w = kGeoTri;
                    g.l2 = make_float4(v0.x, v0.y, v0.z, area);
                    g.l3 = make_float4(e1.x, e1.y, e1.z, mx(0.0f, l.soft_radius));
                    g.l4 = make_float4(e2.x, e2.y, e2.z, 0.0f);
                    g.l5 = make_float4(n.x, n.y, n.z, area);
                } else if (p.kind == RT_PRIM_QUAD) {
                    f3 q = ld3(v), u = ld3(v + 3), w = ld3(v + 6);
                    f3 nn = cross(u, w);
                    f3 n = normalize(nn);
                    area = length(nn);
                    g.l0.w = kGeoQuad;
                    g.l2 = make_float4(q.x, q.y, q.z, area);
                    g.l3 = make_float4(u.x, u.y, u.z, mx(0.0f, l.soft_radius));
                    g.l4 = make_float4(w.x, w.y, w.z, 0.0f);
                    g.l5 = make_float4(n.x, n.y, n.z, area);
                } else {
                    area = 4.0f * kPi * v[3] * v[3];
                    g.l0.w = kGeoSphere;
                    g.l2 = make_float4(v[0], v[1], v[2], v[3]);
                    g.l5 = make_float4(0, 0, 0, area);
                }
                power = kPi * area * luminance(em);
                hs.light_prim[i] = l.primitive;
                break;
            }
            case RT_LIGHT_ENVIRONMENT: {
                // EnvironmentLight over the flat sky: it samples only when the
                // sky can (FlatSky.HasImportanceSampling: luminance > 1e-6);
                // power pi * mean luminance * the bounds' mean cross-section
                // (EnvironmentLight.ApproximatePower, EnvironmentLight.cs:49-64)
                const f3 sky = hs.sky;
                const bool sampleable = luminance(sky) > 1e-6f;
                ss = std::max(1, l.shadow_samples);
                g.l1 = make_float4(sky.x, sky.y, sky.z, 1.0f);
                g.l0.w = sampleable ? 1 : 0;
                if (sampleable) {
                    const f3 e = hs.bounds.mx - hs.bounds.mn;
                    const float cross_section = 0.5f * (e.x * e.y + e.y * e.z + e.x * e.z);
                    power = (kPi * luminance(sky)) * cross_section;
                }
                if (hs.env_light < 0) hs.env_light = (int)i;
                break;
            }
            default:
                err = "bad light kind";
                return RT_ERR_INVALID_ARGUMENT;
        }
        int sq = (int)ceilf(sqrtf((float)ss));
        g.l0.x = l.kind;
        g.l0.y = ss;
        g.l0.z = sq;
        flux += power;
        hs.light_power[i] = power;
        hs.max_shadow_per_vertex += ss;
        hs.max_light_samples = std::max(hs.max_light_samples, ss);
    }
    // emitter → light map (Renderer.cs:409-426): geometry lights (last write
    // wins), then light proxies
    for (int64_t i = 0; i < d->num_lights; i++)
        if (d->lights[i].kind == RT_LIGHT_GEOMETRY) hs.mats[d->primitives[d->lights[i].primitive].material].m0.z = (int32_t)i;
    for (int64_t i = 0; i < d->num_lights; i++) {
        const rt_light& l = d->lights[i];
        if ((l.kind == RT_LIGHT_SPHERE || l.kind == RT_LIGHT_AREA) && l.material >= 0 && l.material < d->num_materials)
            hs.mats[l.material].m0.z = (int32_t)i;
    }

    // Russian-roulette classifier (Renderer.cs:361-378)
    f3 ext = hs.bounds.mx - hs.bounds.mn;
    float R = mx(0.5f * length(ext), 1e-3f);
    hs.total_flux = flux;
    hs.scene_radius = R;
    hs.mean_irradiance = flux / (4.0f * kPi * R * R);
    bool ind = hs.mean_irradiance < 0.5f;
    hs.rr_min_bounces = ind ? 6 : 3;
    hs.rr_min_survival = ind ? 0.40f : 0.10f;
    return RT_OK;
}

}  // namespace rt

// ── diagnostics: serialise the flattened world BVH ───────────────────────
// Pre-order: internal node → (0, axis, 6 box floats as bits); leaf → (1,
// count, 6 box floats, entry ids...) with an entry id = primitive id, or
// -(mesh ordinal + 2) for a nested mesh.  Same format as the CPU oracle's
// dump of the reference tree.
extern "C" int64_t rt_debug_bvh_dump(const rt_scene_desc* desc, int32_t* out, int64_t cap) {
    using namespace rt;
    HostScene hs;
    std::string err;
    if (build_host_scene(desc, hs, err) != RT_OK) return -1;
    if (!hs.has_bvh) return 0;
    std::vector<int32_t> v;
    auto box = [&](const Box& b) {
        float f[6] = {b.mn.x, b.mn.y, b.mn.z, b.mx.x, b.mx.y, b.mx.z};
        for (float x : f) v.push_back(fbits(x));
    };
    std::function<void(uint32_t, const Box&)> rec = [&](uint32_t ref, const Box& b) {
        if (ref_is_leaf(ref)) {
            uint32_t s = leaf_start(ref), c = leaf_count(ref);
            v.push_back(1);
            v.push_back((int32_t)c);
            box(b);
            for (uint32_t k = 0; k < c; k++) {
                const GEntry& e = hs.entries[s + k];
                int32_t kind = fbits(e.a.w) & 0xff;
                if (kind != kBvh) v.push_back(fbits(e.b.w));
                else {
                    uint32_t r = (uint32_t)fbits(e.b.w);
                    int32_t id = -1;
                    for (size_t m = 0; m < hs.mesh_refs.size(); m++)
                        if (hs.mesh_refs[m] == r) id = -(int32_t)m - 2;
                    v.push_back(id);
                }
            }
            return;
        }
        const GNode& n = hs.nodes[ref];
        v.push_back(0);
        v.push_back(n.d.z);
        box(b);
        Box b0{mk(n.a.x, n.a.y, n.a.z), mk(n.a.w, n.b.x, n.b.y)};
        Box b1{mk(n.b.z, n.b.w, n.c.x), mk(n.c.y, n.c.z, n.c.w)};
        rec((uint32_t)n.d.x, b0);
        rec((uint32_t)n.d.y, b1);
    };
    rec(hs.world_root, hs.world_box);
    int64_t cnt = (int64_t)v.size();
    if (out && cap >= cnt) std::memcpy(out, v.data(), (size_t)cnt * 4);
    return cnt;
}
