// oracle.cpp — CPU restatement of the reference's render hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it.  It follows the reference C# source function by function (file:line
// cited at each piece) with the reference's recursive control flow, so that
// a difference between it and the wavefront GPU design shows up as a test
// failure rather than being shared by both.
//
// Each function is pinned by the tests under tests/ that exercise it; the
// functions whose bits depend on .NET's Vector3 reduction order or MathF
// transcendentals are statistically pinned only.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>
#include <algorithm>
#include <functional>
#include <unordered_map>
#include <string>
#include <cmath>

#include "../include/rt3d.h"
#include "orc_core.h"

namespace orc {

// ═════════════════════════════════════════════════════════════════════════
// Sampler context (Sampler.cs:75-118).  Thread-local like the reference's
// [ThreadStatic] Ctx.  "fixed" mode feeds a caller-supplied list of uniforms
// (unit tests of BSDF/light sampling).
// ═════════════════════════════════════════════════════════════════════════
struct SamplerCtx {
    bool active = false;
    bool prng = false;          // PRNG sampler mode (prng_sample) instead of Sobol
    uint32_t seed = 0, index = 0, dim = 0;
    const float* fixed = nullptr;
    int nfixed = 0, fixed_pos = 0;
    uint64_t rng = 0x853c49e6748fea9bull;
};
static thread_local SamplerCtx t_smp;

static float rng_next(uint64_t& s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t x = (uint32_t)(((s >> 18u) ^ s) >> 27u);
    uint32_t rot = (uint32_t)(s >> 59u);
    x = (x >> rot) | (x << ((32u - rot) & 31u));
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}

// Sampler.Sample1D / MathUtils.RandomFloat (Sampler.cs:106-118)
static float random_float() {
    SamplerCtx& c = t_smp;
    if (c.fixed) {
        if (c.fixed_pos < c.nfixed) return c.fixed[c.fixed_pos++];
        c.fixed_pos++;
        return 0.5f;
    }
    if (!c.active) return rng_next(c.rng);
    uint32_t d = c.dim++;
    return c.prng ? prng_sample(c.index, d, c.seed) : sobol_sample(c.index, d, c.seed);
}

// MathUtils.RandomUnitVector (MathUtils.cs:93-101)
static V3 random_unit_vector() {
    float u = random_float();
    float v = random_float();
    float z = 1.0f - 2.0f * u;
    float r = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
    float phi = 2.0f * kPi * v;
    return V3(r * std::cos(phi), r * std::sin(phi), z);
}
// MathUtils.RandomInUnitDisk (MathUtils.cs:137-155)
static V3 random_in_unit_disk() {
    float a = 2.0f * random_float() - 1.0f;
    float b = 2.0f * random_float() - 1.0f;
    if (a == 0.0f && b == 0.0f) return V3(0.0f);
    float r, phi;
    if (a * a > b * b) { r = a; phi = (kPi / 4.0f) * (b / a); }
    else { r = b; phi = (kPi / 2.0f) - (kPi / 4.0f) * (a / b); }
    return V3(r * std::cos(phi), r * std::sin(phi), 0.0f);
}

// Traversal instrumentation (node box tests / primitive tests), used to
// count algorithmic bytes on the reference tree.
static thread_local int64_t t_node_tests = 0;
static thread_local int64_t t_prim_tests = 0;

struct Material;

// HitRecord.cs:9-163 (the fields the path reads)
struct HitRecord {
    V3 point, normal, shading_ng, tangent, bitangent;
    V3 dpdu, dpdv;                 // HitRecord.DpDu/DpDv: left as they are by primitives that do not set them
    float fp[4] = {0, 0, 0, 0};    // FilterFootprint UV partials (dudx, dvdx, dudy, dvdy); zero = none
    float t = 0, u = 0, v = 0;
    bool front_face = false;
    const Material* mat = nullptr;
    float relative_eta = 0.0f;
    int prim = -1;
    bool camera_invisible = false;
    void set_face_normal(const Ray& r, V3 outward) {
        front_face = dot(r.d, outward) < 0.0f;
        normal = front_face ? outward : -outward;
    }
};

// ═════════════════════════════════════════════════════════════════════════
// Geometry
// ═════════════════════════════════════════════════════════════════════════
struct Hittable {
    virtual ~Hittable() {}
    virtual bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const = 0;
    virtual AABB box() const = 0;
    // SceneLoader.IsInfinitePlane (SceneLoader.cs:3640-3650): the world
    // list's test, which sees through Transform, CameraInvisible and a Group
    // of planes
    virtual bool is_plane() const { return false; }
    // Group.IsInfinitePlane (Group.cs:88-94): only InfinitePlane and a
    // Transform of one
    virtual bool is_plane_in_group() const { return false; }
};

// Triangle.cs:35-89,144-149
struct Triangle : Hittable {
    V3 v0, v1, v2, e1, e2, n, tan, bitan;
    float area;
    const Material* mat;
    int prim;
    Triangle(V3 a, V3 b, V3 c, const Material* m, int id) : v0(a), v1(b), v2(c), mat(m), prim(id) {
        e1 = v1 - v0;
        e2 = v2 - v0;
        V3 cr = cross(e1, e2);
        float cl = length(cr);
        n = cl > 0.0f ? cr / cl : V3(0, 1, 0);
        area = 0.5f * cl;
        tan = length_sq(e1) > 0.0f ? normalize(e1) : V3(1, 0, 0);
        bitan = length_sq(e2) > 0.0f ? normalize(e2) : V3(0, 0, 1);
    }
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        t_prim_tests++;
        V3 h = cross(r.d, e2);
        float a = dot(e1, h);
        if (std::fabs(a) < 1e-7f) return false;
        float f = 1.0f / a;
        V3 s = r.o - v0;
        float u = f * dot(s, h);
        if (u < 0.0f || u > 1.0f) return false;
        V3 q = cross(s, e1);
        float v = f * dot(r.d, q);
        if (v < 0.0f || u + v > 1.0f) return false;
        float t = f * dot(e2, q);
        if (t < tmin || t > tmax) return false;
        rec.t = t;
        rec.point = r.at(t);
        rec.set_face_normal(r, n);
        rec.u = u;
        rec.v = v;
        rec.tangent = tan;
        rec.bitangent = bitan;
        rec.dpdu = e1;                  // Triangle.cs:83-84
        rec.dpdv = e2;
        rec.mat = mat;
        rec.prim = prim;
        rec.camera_invisible = false;
        return true;
    }
    AABB box() const override {
        V3 mn = vmin(vmin(v0, v1), v2) - V3(kEpsilon);
        V3 mx = vmax(vmax(v0, v1), v2) + V3(kEpsilon);
        return AABB(mn, mx);
    }
    // Triangle.SampleStratified/SampleAt (Triangle.cs:120-140)
    void sample_stratified(int idx, int sq, V3& p, V3& nrm, float& ar) const {
        float inv = 1.0f / (float)sq;
        int su = idx % sq, sv = idx / sq;
        float xi1 = ((float)su + random_float()) * inv;
        float xi2 = ((float)sv + random_float()) * inv;
        float r1 = std::sqrt(xi1);
        float u = 1.0f - r1;
        float v = xi2 * r1;
        p = v0 + u * e1 + v * e2;
        nrm = n;
        ar = area;
    }
};

// SmoothTriangle.cs:68-130 (constructor), 147-197 (Hit).  The Möller-Trumbore
// test is the flat triangle's; the hit carries the interpolated vertex normal,
// oriented by the face normal's front test, and the frame of the UV mapping.
// rec.u/v stay the barycentrics: the texture lookup interpolates the vertex
// UVs from them exactly as SmoothTriangle.Hit fills HitRecord.U/V.
struct SmoothTriangle : Triangle {
    V3 n0, n1, n2;
    SmoothTriangle(V3 a, V3 b, V3 c, V3 na, V3 nb, V3 nc, const float uv[6], const Material* m, int id)
        : Triangle(a, b, c, m, id) {
        n0 = normalize(na);
        n1 = normalize(nb);
        n2 = normalize(nc);
        float du1x = uv[2] - uv[0], du1y = uv[3] - uv[1];
        float du2x = uv[4] - uv[0], du2y = uv[5] - uv[1];
        float det = du1x * du2y - du2x * du1y;
        bool done = false;
        if (std::fabs(det) > 1e-8f) {
            float inv = 1.0f / det;
            V3 t = inv * (du2y * e1 - du1y * e2);
            V3 bb = inv * (-du2x * e1 + du1x * e2);
            float tl = length(t), bl = length(bb);
            if (tl > 1e-8f && bl > 1e-8f) {
                tan = t / tl;
                bitan = bb / bl;
                done = true;
            }
        }
        if (!done) {
            tan = normalize(e1);
            bitan = normalize(e2);
        }
    }
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        V3 du = rec.dpdu, dv = rec.dpdv;   // SmoothTriangle.Hit does not set DpDu/DpDv
        if (!Triangle::hit(r, tmin, tmax, rec)) return false;
        rec.dpdu = du;
        rec.dpdv = dv;
        float w0 = 1.0f - rec.u - rec.v;
        V3 sn = normalize(w0 * n0 + rec.u * n1 + rec.v * n2);
        rec.normal = rec.front_face ? sn : -sn;
        return true;
    }
};

static void sphere_uv(V3 p, float& u, float& v) {
    float theta = std::acos(-p.y);
    float phi = std::atan2(-p.z, p.x) + kPi;
    u = phi / (2.0f * kPi);
    v = theta / kPi;
}

// Frisvad ONB + local→world (Sphere.cs:237-256, SphereLight.cs:284-305)
static V3 local_to_world(V3 local, V3 w) {
    V3 u, v;
    if (w.z < -0.999f) { u = V3(0, -1, 0); v = V3(-1, 0, 0); }
    else {
        float a = 1.0f / (1.0f + w.z);
        float b = -w.x * w.y * a;
        u = V3(1.0f - w.x * w.x * a, b, -w.x);
        v = V3(b, 1.0f - w.y * w.y * a, -w.y);
    }
    V3 res = local.x * u + local.y * v + local.z * w;
    float l2 = length_sq(res);
    return l2 > 1e-8f ? res / std::sqrt(l2) : w;
}

// Sphere.cs:20-75,139-229,271-275
struct Sphere : Hittable {
    V3 c;
    float r;
    const Material* mat;
    int prim;
    Sphere(V3 center, float radius, const Material* m, int id) : c(center), r(radius), mat(m), prim(id) {}
    bool hit(const Ray& ray, float tmin, float tmax, HitRecord& rec) const override {
        t_prim_tests++;
        V3 oc = ray.o - c;
        float a = length_sq(ray.d);
        float hb = dot(oc, ray.d);
        float cc = length_sq(oc) - r * r;
        float disc = hb * hb - a * cc;
        if (disc < 0.0f) return false;
        float sd = std::sqrt(disc);
        float root = (-hb - sd) / a;
        if (root < tmin || root > tmax) {
            root = (-hb + sd) / a;
            if (root < tmin || root > tmax) return false;
        }
        rec.t = root;
        rec.point = ray.at(root);
        V3 outward = (rec.point - c) / r;
        rec.set_face_normal(ray, outward);
        sphere_uv(outward, rec.u, rec.v);
        V3 td = cross(V3(0, 1, 0), outward);
        if (length_sq(td) < 1e-4f) td = V3(1, 0, 0);
        rec.tangent = normalize(td);
        rec.bitangent = normalize(cross(outward, rec.tangent));
        // Sphere.cs:65-70: dP/du and dP/dv of phi = 2 pi u, theta = pi v
        float st = std::sqrt(std::fmax(0.0f, 1.0f - outward.y * outward.y));
        rec.dpdu = rec.tangent * (2.0f * kPi * r * st);
        rec.dpdv = rec.bitangent * (kPi * r);
        rec.mat = mat;
        rec.prim = prim;
        rec.camera_invisible = false;
        return true;
    }
    AABB box() const override { return AABB(c - V3(r), c + V3(r)); }
    // Sphere.SampleSolidAngleCore (Sphere.cs:139-203)
    void sample_solid_angle(V3 from, float xi1, float xi2, V3& p, V3& nrm, float& pdf) const {
        V3 to_c = c - from;
        float d2 = length_sq(to_c);
        float r2 = r * r;
        if (d2 <= r2) {
            V3 pu = random_unit_vector();
            p = c + pu * r;
            nrm = pu;
            pdf = 1.0f / (4.0f * kPi);
            return;
        }
        float dist = std::sqrt(d2);
        V3 w = to_c / dist;
        float s2 = r2 / d2;
        float ctm = std::sqrt(std::fmax(0.0f, 1.0f - s2));
        float ct = 1.0f - xi1 * (1.0f - ctm);
        float st = std::sqrt(std::fmax(0.0f, 1.0f - ct * ct));
        float phi = 2.0f * kPi * xi2;
        V3 local(st * std::cos(phi), st * std::sin(phi), ct);
        V3 dw = local_to_world(local, w);
        V3 oc = from - c;
        float b = dot(oc, dw);
        float cc = length_sq(oc) - r2;
        float disc = b * b - cc;
        float t;
        if (disc <= 0.0f) t = -b;
        else {
            float sd = std::sqrt(disc);
            t = -b - sd;
            if (t < kEpsilon) t = -b + sd;
        }
        p = from + t * dw;
        nrm = normalize(p - c);
        float sa = 2.0f * kPi * (1.0f - ctm);
        pdf = sa > kEpsilon ? 1.0f / sa : 0.0f;
    }
    // Sphere.SolidAnglePdf (Sphere.cs:206-229)
    float solid_angle_pdf(V3 from, V3 wi) const {
        V3 to_c = c - from;
        float d2 = length_sq(to_c);
        float r2 = r * r;
        if (d2 <= r2) return 1.0f / (4.0f * kPi);
        float s2 = r2 / d2;
        float ctm = std::sqrt(std::fmax(0.0f, 1.0f - s2));
        float dist = std::sqrt(d2);
        V3 w = to_c / dist;
        float wl = length(wi);
        if (wl < kEpsilon) return 0.0f;
        float ct = dot(wi, w) / wl;
        if (ct < ctm) return 0.0f;
        float sa = 2.0f * kPi * (1.0f - ctm);
        return sa > kEpsilon ? 1.0f / sa : 0.0f;
    }
};

// Quad.cs:302-396 (+ BackFaceCulledHittable.cs:42-50 via `culled`)
struct Quad : Hittable {
    V3 q, u, v, n, w, tan, bitan;
    float d, area;
    const Material* mat;
    int prim;
    bool culled;
    Quad(V3 Q, V3 U, V3 V, const Material* m, int id, bool cull) : q(Q), u(U), v(V), mat(m), prim(id), culled(cull) {
        V3 nn = cross(u, v);
        n = normalize(nn);
        d = dot(n, q);
        w = nn / dot(nn, nn);
        tan = normalize(u);
        bitan = normalize(v);
        area = length(nn);
    }
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        t_prim_tests++;
        float denom = dot(n, r.d);
        if (std::fabs(denom) < 1e-8f) return false;
        float t = (d - dot(n, r.o)) / denom;
        if (t < tmin || t > tmax) return false;
        V3 ip = r.at(t);
        V3 pv = ip - q;
        float alpha = dot(w, cross(pv, v));
        float beta = dot(w, cross(u, pv));
        if (alpha < 0.0f || alpha > 1.0f || beta < 0.0f || beta > 1.0f) return false;
        bool front = dot(r.d, n) < 0.0f;
        if (culled && !front) return false;
        rec.u = alpha;
        rec.v = beta;
        rec.t = t;
        rec.point = ip;
        rec.set_face_normal(r, n);
        rec.tangent = tan;
        rec.bitangent = bitan;
        rec.dpdu = u;                   // Quad.cs:79-80
        rec.dpdv = v;
        rec.mat = mat;
        rec.prim = prim;
        rec.camera_invisible = false;
        return true;
    }
    AABB box() const override {
        V3 bmin = q, bmax = (q + u) + v;
        V3 mn = vmin(bmin, vmin(q + u, q + v)) - V3(0.0001f);
        V3 mx = vmax(bmax, vmax(q + u, q + v)) + V3(0.0001f);
        return AABB(mn, mx);
    }
    void sample_stratified(int idx, int sq, V3& p, V3& nrm, float& ar) const {
        float inv = 1.0f / (float)sq;
        int su = idx % sq, sv = idx / sq;
        float a = ((float)su + random_float()) * inv;
        float b = ((float)sv + random_float()) * inv;
        p = (q + a * u) + b * v;
        nrm = n;
        ar = area;
    }
};

// InfinitePlane.cs:18-80
struct Plane : Hittable {
    V3 p, n, ua, va;
    const Material* mat;
    int prim;
    Plane(V3 point, V3 normal, const Material* m, int id) : p(point), mat(m), prim(id) {
        n = normalize(normal);
        ua = std::fabs(n.y) < 0.99f ? normalize(cross(n, V3(0, 1, 0))) : V3(1, 0, 0);
        va = cross(n, ua);
    }
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        t_prim_tests++;
        float denom = dot(n, r.d);
        if (std::fabs(denom) < 1e-6f) return false;
        float t = dot(p - r.o, n) / denom;
        if (t < tmin || t > tmax) return false;
        rec.t = t;
        rec.point = r.at(t);
        rec.set_face_normal(r, n);
        V3 lp = rec.point - p;
        rec.u = dot(lp, ua);
        rec.v = dot(lp, va);
        rec.u -= std::floor(rec.u);
        rec.v -= std::floor(rec.v);
        rec.tangent = ua;
        rec.bitangent = va;
        rec.mat = mat;
        rec.prim = prim;
        rec.camera_invisible = false;
        return true;
    }
    AABB box() const override { return AABB(V3(-1e6f), V3(1e6f)); }
    bool is_plane() const override { return true; }
    bool is_plane_in_group() const override { return true; }
};

// CameraInvisibleHittable: stamps the camera bit of the visibility mask.
struct CameraInvisible : Hittable {
    Hittable* inner;
    explicit CameraInvisible(Hittable* h) : inner(h) {}
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        if (!inner->hit(r, tmin, tmax, rec)) return false;
        rec.camera_invisible = true;
        return true;
    }
    AABB box() const override { return inner->box(); }
    bool is_plane() const override { return inner->is_plane(); }
};

// HittableList.cs:19-38
struct HittableList : Hittable {
    std::vector<Hittable*> objs;
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        HitRecord tmp;
        bool any = false;
        float closest = tmax;
        for (Hittable* o : objs) {
            if (o->hit(r, tmin, closest, tmp)) {
                any = true;
                closest = tmp.t;
                rec = tmp;
            }
        }
        return any;
    }
    AABB box() const override {
        if (objs.empty()) return AABB::empty();
        AABB b = objs[0]->box();
        for (size_t i = 1; i < objs.size(); i++) b = surround(b, objs[i]->box());
        return b;
    }
};

// BvhNode.cs:45-371
struct BvhNode : Hittable {
    BvhNode* left = nullptr;
    BvhNode* right = nullptr;
    std::vector<Hittable*> prims;
    AABB bbox;
    int axis = 0;

    static constexpr int kMaxLeaf = 4, kBins = 16;
    static constexpr float kTrav = 0.5f;

    static bool try_sah(std::vector<Hittable*>& o, std::vector<AABB>& bx, int start, int end, int& best_axis, int& mid) {
        int span = end - start;
        V3 cmin(FLT_MAX), cmax(-FLT_MAX);
        for (int i = start; i < end; i++) {
            V3 c = (bx[i].mn + bx[i].mx) * 0.5f;
            cmin = vmin(cmin, c);
            cmax = vmax(cmax, c);
        }
        V3 ext = cmax - cmin;
        float max_ext = fmax_p(ext.x, fmax_p(ext.y, ext.z));
        if (max_ext <= 0.0f) { best_axis = -1; mid = start; return false; }
        best_axis = -1;
        int best_bin = -1;
        float best_cost = FLT_MAX;
        int cnt[kBins], pc[kBins], sc[kBins];
        AABB bb[kBins], pb[kBins], sb[kBins];
        for (int axis = 0; axis < 3; axis++) {
            float ea = ext[axis];
            if (ea <= 0.0f) continue;
            float cm = cmin[axis];
            float scale = (float)kBins / ea;
            for (int i = 0; i < kBins; i++) { cnt[i] = 0; bb[i] = AABB::empty(); }
            for (int i = start; i < end; i++) {
                float c = ((bx[i].mn + bx[i].mx) * 0.5f)[axis];
                int idx = (int)((c - cm) * scale);
                if (idx < 0) idx = 0;
                else if (idx >= kBins) idx = kBins - 1;
                cnt[idx]++;
                bb[idx] = surround(bb[idx], bx[i]);
            }
            int cum = 0;
            AABB cb = AABB::empty();
            for (int i = 0; i < kBins; i++) { cum += cnt[i]; cb = surround(cb, bb[i]); pc[i] = cum; pb[i] = cb; }
            cum = 0;
            cb = AABB::empty();
            for (int i = kBins - 1; i >= 0; i--) { cum += cnt[i]; cb = surround(cb, bb[i]); sc[i] = cum; sb[i] = cb; }
            float parea = pb[kBins - 1].area();
            if (parea <= 0.0f) continue;
            for (int i = 0; i < kBins - 1; i++) {
                int nl = pc[i], nr = sc[i + 1];
                if (nl == 0 || nr == 0) continue;
                float cost = kTrav + ((float)nl * pb[i].area() + (float)nr * sb[i + 1].area()) / parea;
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_bin = i; }
            }
        }
        if (best_axis < 0) { mid = start; return false; }
        float amin = cmin[best_axis];
        float ascale = (float)kBins / ext[best_axis];
        int wp = start;
        for (int i = start; i < end; i++) {
            float c = ((bx[i].mn + bx[i].mx) * 0.5f)[best_axis];
            int idx = (int)((c - amin) * ascale);
            if (idx < 0) idx = 0;
            else if (idx >= kBins) idx = kBins - 1;
            if (idx <= best_bin) {
                std::swap(o[i], o[wp]);
                std::swap(bx[i], bx[wp]);
                wp++;
            }
        }
        mid = wp;
        if (mid == start || mid == end) mid = start + span / 2;
        return true;
    }

    BvhNode(std::vector<Hittable*>& o, std::vector<AABB>& bx, int start, int end) {
        int span = end - start;
        if (span <= kMaxLeaf) {
            AABB b = bx[start];
            prims.push_back(o[start]);
            for (int i = 1; i < span; i++) { prims.push_back(o[start + i]); b = surround(b, bx[start + i]); }
            bbox = b;
            axis = 0;
            return;
        }
        int ba, mid;
        if (!try_sah(o, bx, start, end, ba, mid)) { ba = 0; mid = start + span / 2; }
        axis = ba;
        left = new BvhNode(o, bx, start, mid);
        right = new BvhNode(o, bx, mid, end);
        bbox = surround(left->bbox, right->bbox);
    }
    ~BvhNode() override { delete left; delete right; }

    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        t_node_tests++;
        if (!bbox.hit(r, tmin, tmax)) return false;
        if (!prims.empty()) {
            bool any = false;
            float closest = tmax;
            for (Hittable* p : prims) {
                if (p->hit(r, tmin, closest, rec)) { any = true; closest = rec.t; }
            }
            return any;
        }
        float da = r.d[axis];
        const BvhNode* nr = da >= 0.0f ? left : right;
        const BvhNode* fr = da >= 0.0f ? right : left;
        bool hn = nr->hit(r, tmin, tmax, rec);
        bool hf = fr->hit(r, tmin, hn ? rec.t : tmax, rec);
        return hn || hf;
    }
    AABB box() const override { return bbox; }

    void stats(int depth, int64_t& nodes, int64_t& leaves, int& maxd) const {
        if (!prims.empty()) { leaves++; maxd = std::max(maxd, depth); return; }
        nodes++;
        left->stats(depth + 1, nodes, leaves, maxd);
        right->stats(depth + 1, nodes, leaves, maxd);
    }
};

static Hittable* build_bvh(std::vector<Hittable*> objs) {
    std::vector<AABB> bx(objs.size());
    for (size_t i = 0; i < objs.size(); i++) bx[i] = objs[i]->box();
    return new BvhNode(objs, bx, 0, (int)objs.size());
}

// Mesh.cs:75-116 — nested BVH (or list when <= 4 faces)
struct Mesh : Hittable {
    std::vector<Hittable*> tris;
    Hittable* accel = nullptr;
    void finish() {
        if (tris.size() > 4) accel = build_bvh(tris);
        else {
            auto* l = new HittableList();
            l->objs = tris;
            accel = l;
        }
    }
    ~Mesh() override { delete accel; }
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override { return accel->hit(r, tmin, tmax, rec); }
    AABB box() const override { return accel->box(); }
};

// ── Wrappers of the object hierarchy (rt_node) ──────────────────────────
// System.Numerics.Vector3.Transform / TransformNormal with Matrix4x4 in its
// field order M11..M44 (row vectors): the third-party runtime (.NET 10,
// RayTracer.csproj TargetFramework net10.0) evaluates them in
// DirectXMath's XMVector3Transform order with multiply-adds,
// fma(x, row1, fma(y, row2, fma(z, row3, row4))); TransformNormal drops
// row4 and starts from z * row3.
static V3 xf_point(V3 p, const float* M) {
    return V3(std::fmaf(p.x, M[0], std::fmaf(p.y, M[4], std::fmaf(p.z, M[8], M[12]))),
              std::fmaf(p.x, M[1], std::fmaf(p.y, M[5], std::fmaf(p.z, M[9], M[13]))),
              std::fmaf(p.x, M[2], std::fmaf(p.y, M[6], std::fmaf(p.z, M[10], M[14]))));
}
static V3 xf_vec(V3 v, const float* M) {
    return V3(std::fmaf(v.x, M[0], std::fmaf(v.y, M[4], v.z * M[8])),
              std::fmaf(v.x, M[1], std::fmaf(v.y, M[5], v.z * M[9])),
              std::fmaf(v.x, M[2], std::fmaf(v.y, M[6], v.z * M[10])));
}

// Transform.cs:74-240: the child hit in object space (ToLocalRay), the
// record mapped back (MapHitToWorld); the world box is the child box's 8
// corners mapped by the matrix (ComputeWorldBox).
struct TransformH : Hittable {
    Hittable* inner;
    float M[16], I[16], N[16];          // matrix, inverse, normal matrix (transposed inverse)
    AABB wbox;
    TransformH(Hittable* h, const float* m, const float* inv) : inner(h) {
        for (int k = 0; k < 16; k++) { M[k] = m[k]; I[k] = inv[k]; }
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) N[4 * r + c] = inv[4 * c + r];
        AABB b = inner->box();
        V3 mn(FLT_MAX), mx(-FLT_MAX);
        for (int k = 0; k < 8; k++) {
            V3 c((k & 4) ? b.mx.x : b.mn.x, (k & 2) ? b.mx.y : b.mn.y, (k & 1) ? b.mx.z : b.mn.z);
            V3 t = xf_point(c, M);
            mn = V3(std::fmin(mn.x, t.x), std::fmin(mn.y, t.y), std::fmin(mn.z, t.z));
            mx = V3(std::fmax(mx.x, t.x), std::fmax(mx.y, t.y), std::fmax(mx.z, t.z));
        }
        wbox = AABB(mn, mx);
    }
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        Ray local(xf_point(r.o, I), xf_vec(r.d, I));
        if (!inner->hit(local, tmin, tmax, rec)) return false;
        rec.point = xf_point(rec.point, M);
        rec.normal = normalize(xf_vec(rec.normal, N));
        rec.tangent = normalize(xf_vec(rec.tangent, M));
        rec.bitangent = normalize(xf_vec(rec.bitangent, M));
        if (dot(rec.dpdu, rec.dpdu) > 0.0f) rec.dpdu = xf_vec(rec.dpdu, M);
        if (dot(rec.dpdv, rec.dpdv) > 0.0f) rec.dpdv = xf_vec(rec.dpdv, M);
        return true;
    }
    AABB box() const override { return wbox; }
    bool is_plane() const override { return inner->is_plane(); }
    bool is_plane_in_group() const override { return inner->is_plane_in_group(); }
};

// Instance.cs:54-84: the shared template's hit with the material override.
struct InstanceH : Hittable {
    Hittable* inner;
    const Material* override_mat;
    InstanceH(Hittable* h, const Material* m) : inner(h), override_mat(m) {}
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override {
        if (!inner->hit(r, tmin, tmax, rec)) return false;
        if (override_mat) rec.mat = override_mat;
        return true;
    }
    AABB box() const override { return inner->box(); }
};

// Group.cs:20-80: a BVH over the finite children when more than 4, the
// infinite planes after it; one member is hit directly.
struct GroupH : Hittable {
    std::vector<Hittable*> children;
    Hittable* root = nullptr;
    HittableList* list = nullptr;
    Hittable* bvh = nullptr;
    explicit GroupH(std::vector<Hittable*> ch) : children(std::move(ch)) {
        std::vector<Hittable*> fin, inf, roots;
        for (Hittable* c : children) (c->is_plane_in_group() ? inf : fin).push_back(c);
        if (fin.size() > 4) {
            bvh = build_bvh(fin);
            roots.push_back(bvh);
        } else roots.insert(roots.end(), fin.begin(), fin.end());
        roots.insert(roots.end(), inf.begin(), inf.end());
        if (roots.size() == 1) root = roots[0];
        else {
            list = new HittableList();
            list->objs = roots;
            root = list;
        }
    }
    ~GroupH() override { delete bvh; delete list; }
    bool hit(const Ray& r, float tmin, float tmax, HitRecord& rec) const override { return root->hit(r, tmin, tmax, rec); }
    AABB box() const override { return root->box(); }
    bool is_plane() const override {
        if (children.empty()) return false;
        for (Hittable* c : children)
            if (!c->is_plane()) return false;
        return true;
    }
};

// ═════════════════════════════════════════════════════════════════════════
// Energy-compensation LUT (EnergyCompensationLut.cs:34-171) + microfacet
// helpers (Microfacet.cs).
// ═════════════════════════════════════════════════════════════════════════
static void build_tangent_frame(V3 n, V3& t, V3& b) {
    if (n.z < -0.999f) { t = V3(0, -1, 0); b = V3(-1, 0, 0); }
    else {
        float a = 1.0f / (1.0f + n.z);
        float bb = -n.x * n.y * a;
        t = V3(1.0f - n.x * n.x * a, bb, -n.x);
        b = V3(bb, 1.0f - n.y * n.y * a, -n.y);
    }
}
static float d_ggx_aniso(V3 h, float ax, float ay) {
    float hx = h.x / ax, hy = h.y / ay;
    float d = hx * hx + hy * hy + h.z * h.z;
    return 1.0f / std::fmax(kPi * ax * ay * d * d, 1e-20f);
}
static float lambda_ggx_aniso(V3 w, float ax, float ay) {
    float c2 = w.z * w.z;
    if (c2 < 1e-14f) return 1e10f;
    float num = (ax * w.x) * (ax * w.x) + (ay * w.y) * (ay * w.y);
    return 0.5f * (-1.0f + std::sqrt(1.0f + num / c2));
}
static float g1_ggx_aniso(V3 w, float ax, float ay) { return 1.0f / (1.0f + lambda_ggx_aniso(w, ax, ay)); }
static float smith_g1(float ndx, float alpha) {
    float a2 = alpha * alpha;
    float n2 = ndx * ndx;
    float den = ndx + std::sqrt(a2 + (1.0f - a2) * n2);
    return 2.0f * ndx / std::fmax(den, 1e-7f);
}
static V3 sample_ggx_vndf(V3 v, float ax, float ay, float u1, float u2) {
    V3 vh = normalize(V3(ax * v.x, ay * v.y, v.z));
    float l2 = vh.x * vh.x + vh.y * vh.y;
    V3 t1 = l2 > 0.0f ? V3(-vh.y, vh.x, 0.0f) / std::sqrt(l2) : V3(1, 0, 0);
    V3 t2 = cross(vh, t1);
    float r = std::sqrt(u1);
    float phi = 2.0f * kPi * u2;
    float a = r * std::cos(phi);
    float b = r * std::sin(phi);
    float s = 0.5f * (1.0f + vh.z);
    b = (1.0f - s) * std::sqrt(std::fmax(0.0f, 1.0f - a * a)) + s * b;
    V3 nh = a * t1 + b * t2 + std::sqrt(std::fmax(0.0f, 1.0f - a * a - b * b)) * vh;
    return normalize(V3(ax * nh.x, ay * nh.y, std::fmax(0.0f, nh.z)));
}

struct EnergyLut {
    float E[32 * 32];
    float Eavg[32];
    static float mu_value(int i) { return ((float)i + 0.5f) / 32.0f; }
    static float alpha_value(int i) { return std::fmax(((float)i + 0.5f) / 32.0f, 0.001f); }
    static float next_float(uint32_t& s) {
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        return (float)(s & 0x00FFFFFFu) / (float)0x01000000u;
    }
    static float integrate(float mu, float alpha, uint32_t seed) {
        float st = std::sqrt(std::fmax(0.0f, 1.0f - mu * mu));
        V3 v(st, 0.0f, mu);
        double sum = 0.0;
        uint32_t state = seed == 0 ? 1u : seed;
        int hits = 0;
        for (int s = 0; s < 2048; s++) {
            float u1 = next_float(state);
            float u2 = next_float(state);
            V3 h = sample_ggx_vndf(v, alpha, alpha, u1, u2);
            float vdh = dot(v, h);
            V3 l = 2.0f * vdh * h - v;
            if (l.z <= 0.0f) continue;
            sum += (double)g1_ggx_aniso(l, alpha, alpha);
            hits++;
        }
        if (hits == 0) return 0.0f;
        double e = sum / 2048.0;
        e = e < 0.0 ? 0.0 : (e > 1.0 ? 1.0 : e);
        return (float)e;
    }
    EnergyLut() {
        for (int ai = 0; ai < 32; ai++) {
            float alpha = alpha_value(ai);
            for (int mi = 0; mi < 32; mi++) {
                float mu = mu_value(mi);
                uint32_t seed = ((uint32_t)mi * 0x9E3779B9u) ^ ((uint32_t)ai * 0xB5297A4Du) ^ 0xC2B2AE3Du;
                E[mi * 32 + ai] = integrate(mu, alpha, seed);
            }
            float avg = 0.0f;
            for (int mi = 0; mi < 32; mi++) avg += 2.0f * E[mi * 32 + ai] * mu_value(mi);
            Eavg[ai] = clampf(avg / 32.0f, 0.0f, 1.0f);
        }
    }
    float sample_e(float mu, float alpha) const {
        mu = clampf(mu, 0.0f, 1.0f);
        alpha = clampf(alpha, 0.001f, 1.0f);
        float fmu = mu * 32.0f - 0.5f;
        float fal = alpha * 32.0f - 0.5f;
        int mi0 = std::min(std::max((int)std::floor(fmu), 0), 31);
        int ai0 = std::min(std::max((int)std::floor(fal), 0), 31);
        int mi1 = std::min(mi0 + 1, 31);
        int ai1 = std::min(ai0 + 1, 31);
        float tu = clampf(fmu - (float)mi0, 0.0f, 1.0f);
        float ta = clampf(fal - (float)ai0, 0.0f, 1.0f);
        float e00 = E[mi0 * 32 + ai0], e01 = E[mi0 * 32 + ai1];
        float e10 = E[mi1 * 32 + ai0], e11 = E[mi1 * 32 + ai1];
        float e0 = e00 + (e01 - e00) * ta;
        float e1 = e10 + (e11 - e10) * ta;
        return e0 + (e1 - e0) * tu;
    }
    float sample_eavg(float alpha) const {
        alpha = clampf(alpha, 0.001f, 1.0f);
        float fal = alpha * 32.0f - 0.5f;
        int ai0 = std::min(std::max((int)std::floor(fal), 0), 31);
        int ai1 = std::min(ai0 + 1, 31);
        float ta = clampf(fal - (float)ai0, 0.0f, 1.0f);
        return Eavg[ai0] + (Eavg[ai1] - Eavg[ai0]) * ta;
    }
};
static const EnergyLut& energy_lut() {
    static EnergyLut lut;
    return lut;
}

// SheenCharlie.cs
static float charlie_d(float ndh, float alpha) {
    alpha = std::fmax(alpha, 0.04f);
    float inv = 1.0f / alpha;
    float c2 = ndh * ndh;
    float s2 = std::fmax(1.0f - c2, 0.0f);
    float sp = std::pow(s2, 0.5f * inv);
    return (2.0f + inv) * sp / (2.0f * kPi);
}
static float charlie_lambda_log(float c, float alpha) {
    float a, b, cc, d, e;
    if (alpha < 0.25f) { a = 25.3245f; b = 3.32435f; cc = 0.16801f; d = -1.27393f; e = -4.85967f; }
    else { a = 21.5473f; b = 3.82987f; cc = 0.19823f; d = -1.97760f; e = -4.32054f; }
    return a / (1.0f + b * std::pow(c, cc)) + d * c + e;
}
static float charlie_lambda(float c, float alpha) {
    c = clampf(c, 1e-4f, 1.0f);
    if (alpha <= 0.25f - 0.05f) return std::exp(charlie_lambda_log(c, 0.0f));
    if (alpha >= 0.25f + 0.05f) return std::exp(charlie_lambda_log(c, 1.0f));
    float t = (alpha - (0.25f - 0.05f)) / (2.0f * 0.05f);
    float lo = std::exp(charlie_lambda_log(c, 0.0f));
    float hi = std::exp(charlie_lambda_log(c, 1.0f));
    return lo + (hi - lo) * t;
}
static float charlie_brdf(float ndv, float ndl, float ndh, float alpha) {
    float den = 4.0f * std::fmax(ndv, 1e-6f) * std::fmax(ndl, 1e-6f);
    float g = 1.0f / ((1.0f + charlie_lambda(ndv, alpha)) * (1.0f + charlie_lambda(ndl, alpha)));
    return charlie_d(ndh, alpha) * g / den;
}

// ═════════════════════════════════════════════════════════════════════════
// Materials (IMaterial.cs)
// ═════════════════════════════════════════════════════════════════════════
struct BsdfSample {
    V3 wo, f;
    float pdf = 0;
    bool is_delta = false;
    bool has_next_absorption = false;
    V3 next_absorption;
    int transition = 0;  // 0 none, 1 enter, 2 exit
    bool caustic_caster = true;
};

struct Material {
    virtual ~Material() {}
    virtual bool needs_direct_lighting() const { return true; }
    virtual bool is_delta_scatter() const { return false; }
    virtual V3 evaluate_direct(V3, V3, V3, const HitRecord&) const { return V3(0.0f); }
    virtual V3 shadow_transmittance(V3, const HitRecord&) const { return V3(0.0f); }
    virtual V3 aov_albedo(const HitRecord&) const { return V3(1.0f); }
    virtual V3 shadow_absorption(const HitRecord&) const { return V3(0.0f); }
    virtual bool is_specular_transmissive(const HitRecord&) const { return false; }
    virtual bool try_dielectric_ior(const HitRecord&, float& ior) const { ior = 1.0f; return false; }
    virtual V3 evaluate(V3, V3, const HitRecord&) const { return V3(0.0f); }
    virtual float pdf(V3, V3, const HitRecord&) const { return 0.0f; }
    virtual bool sample(V3, const HitRecord&, BsdfSample&) const { return false; }
    virtual V3 emit(bool) const { return V3(0.0f); }
    virtual bool scatter(const Ray&, const HitRecord&, V3& att, Ray& scattered) const = 0;
    virtual bool is_emissive() const { return false; }
    virtual bool is_light_proxy() const { return false; }
    // textured Disney: the material record whose baseColor / roughness /
    // metallic the texture set replaces per hit (null: not textured)
    virtual const rt_material* texture_raw() const { return nullptr; }
    // ImageTexture albedo / baseColor: the image index and the material
    // record whose colour it replaces per hit (-1: none)
    int image = -1;
    rt_material image_raw{};
};

// Lambertian.cs:37-96
struct Lambertian : Material {
    V3 albedo;
    explicit Lambertian(V3 a) : albedo(a) {}
    V3 aov_albedo(const HitRecord&) const override { return albedo; }
    V3 evaluate_direct(V3 to_light, V3, V3 normal, const HitRecord&) const override {
        float ndl = std::fmax(dot(normal, to_light), 0.0f);
        if (ndl <= 0.0f) return V3(0.0f);
        return albedo * (ndl / kPi);
    }
    bool scatter(const Ray&, const HitRecord& rec, V3& att, Ray& scattered) const override {
        V3 dir = rec.normal + random_unit_vector();
        if (near_zero(dir)) dir = rec.normal;
        scattered = Ray(rec.point, dir);
        att = albedo;
        return true;
    }
    V3 evaluate(V3, V3 l, const HitRecord& rec) const override {
        if (dot(rec.normal, l) <= 0.0f) return V3(0.0f);
        return albedo * (1.0f / kPi);
    }
    float pdf(V3, V3 l, const HitRecord& rec) const override {
        float c = dot(rec.normal, l);
        return c > 0.0f ? c * (1.0f / kPi) : 0.0f;
    }
    bool sample(V3, const HitRecord& rec, BsdfSample& s) const override {
        V3 n = rec.normal;
        V3 dir = n + random_unit_vector();
        if (near_zero(dir)) dir = n;
        V3 wo = normalize(dir);
        float ndw = dot(n, wo);
        if (ndw <= 0.0f) return false;
        s = BsdfSample();
        s.wo = wo;
        s.f = albedo * (1.0f / kPi);
        s.pdf = ndw * (1.0f / kPi);
        s.is_delta = false;
        return true;
    }
};

// Emissive.cs:56-104
struct Emissive : Material {
    V3 color;
    float intensity;
    bool proxy;
    Emissive(V3 c, float i, bool p) : color(c), intensity(std::fmax(i, 0.0f)), proxy(p) {}
    bool needs_direct_lighting() const override { return false; }
    V3 aov_albedo(const HitRecord&) const override { return color; }
    bool scatter(const Ray&, const HitRecord&, V3& att, Ray&) const override { att = V3(0.0f); return false; }
    V3 emit(bool front) const override {
        if (!front) return V3(0.0f);
        return color * intensity;
    }
    V3 emission_at() const { return color * intensity; }
    bool is_emissive() const override { return true; }
    bool is_light_proxy() const override { return proxy; }
};

// Dielectric.cs:46-105
struct Dielectric : Material {
    float ior;
    V3 albedo;
    Dielectric(float i, V3 a) : ior(i), albedo(a) {}
    bool is_delta_scatter() const override { return true; }
    V3 aov_albedo(const HitRecord&) const override { return albedo; }
    V3 shadow_transmittance(V3 wi, const HitRecord& rec) const override {
        float ct = std::fmin(std::fabs(dot(wi, rec.normal)), 1.0f);
        float fr = fresnel_dielectric(ct, 1.0f / ior);
        return (1.0f - fr) * albedo;
    }
    bool is_specular_transmissive(const HitRecord&) const override { return true; }
    bool try_dielectric_ior(const HitRecord&, float& i) const override { i = ior; return true; }
    bool scatter(const Ray& rin, const HitRecord& rec, V3& att, Ray& scattered) const override {
        att = albedo;
        float ri = rec.relative_eta > 0.0f ? rec.relative_eta : (rec.front_face ? (1.0f / ior) : ior);
        V3 ud = normalize(rin.d);
        V3 n = rec.normal;
        float ct = std::fmin(dot(-ud, n), 1.0f);
        float st = std::sqrt(std::fmax(0.0f, 1.0f - ct * ct));
        bool cannot = ri * st > 1.0f;
        V3 dir;
        if (cannot || schlick(ct, ri) > random_float()) dir = reflect(ud, n);
        else dir = refract(ud, n, ri);
        V3 on = dot(dir, rec.normal) >= 0.0f ? rec.normal : -rec.normal;
        scattered = Ray(offset_origin(rec.point, on), dir);
        return true;
    }
};

// DisneyBsdf.cs
// ThinFilm.Evaluate (ThinFilm.cs:40-136) and its helpers (:150-217):
// Belcour-Barla 2017 iridescent Fresnel, RGB through the Gaussian fit of the
// CIE sensitivities.
static float tf_schlick(float f0, float cos_t) {
    float x = std::fmin(std::fmax(1.0f - cos_t, 0.0f), 1.0f);
    float x5 = x * x;
    x5 *= x5 * x;
    return f0 + (1.0f - f0) * x5;
}
static V3 tf_sensitivity(float opd_nm, V3 shift) {
    float phase = 2.0f * kPi * opd_nm * 1e-9f;
    V3 val(5.4856e-13f, 4.4201e-13f, 5.2481e-13f);
    V3 pos(1.6810e+06f, 1.7953e+06f, 2.2084e+06f);
    V3 var(4.3278e+09f, 9.3046e+09f, 6.6121e+09f);
    V3 s2v(std::sqrt(2.0f * kPi * var.x), std::sqrt(2.0f * kPi * var.y), std::sqrt(2.0f * kPi * var.z));
    V3 cs(std::cos(pos.x * phase + shift.x), std::cos(pos.y * phase + shift.y), std::cos(pos.z * phase + shift.z));
    float ph2 = phase * phase;
    V3 ex(std::exp(-ph2 * var.x), std::exp(-ph2 * var.y), std::exp(-ph2 * var.z));
    V3 xyz = val * s2v * cs * ex;
    xyz.x += 9.7470e-14f * std::sqrt(2.0f * kPi * 4.5282e+09f) * std::cos(2.2399e+06f * phase + shift.x) *
             std::exp(-ph2 * 4.5282e+09f);
    xyz = xyz / 1.0685e-7f;
    float r = 3.2404542f * xyz.x - 1.5371385f * xyz.y - 0.4985314f * xyz.z;
    float g = -0.9692660f * xyz.x + 1.8760108f * xyz.y + 0.0415560f * xyz.z;
    float b = 0.0556434f * xyz.x - 0.2040259f * xyz.y + 1.0572252f * xyz.z;
    return V3(r, g, b);
}
static V3 thin_film_fresnel(float cos1, float film_ior, float thick_nm, V3 base_f0) {
    const float outside = 1.0f;
    float thickness = std::fmax(thick_nm, 0.0f);
    float t = std::fmin(std::fmax((thickness - 0.0f) / (30.0f - 0.0f), 0.0f), 1.0f);
    float sm = t * t * (3.0f - 2.0f * t);
    float eta2 = std::fmax(outside + (film_ior - outside) * sm, 1.0001f);
    float c1 = std::fmin(std::fmax(cos1, 0.0f), 1.0f);
    float er = outside / eta2;
    float s2sq = er * er * (1.0f - c1 * c1);
    if (s2sq >= 1.0f) return V3(1.0f);
    float cos2 = std::sqrt(1.0f - s2sq);
    float q12 = (eta2 - outside) / (eta2 + outside);
    float r12 = tf_schlick(q12 * q12, c1);
    float t121 = 1.0f - r12;
    float phi12 = (eta2 < outside) ? kPi : 0.0f;
    float phi21 = kPi - phi12;
    V3 bf(std::fmin(std::fmax(base_f0.x, 0.0f), 0.9999f), std::fmin(std::fmax(base_f0.y, 0.0f), 0.9999f),
          std::fmin(std::fmax(base_f0.z, 0.0f), 0.9999f));
    V3 sq(std::sqrt(bf.x), std::sqrt(bf.y), std::sqrt(bf.z));
    V3 bior = (V3(1.0f) + sq) / (V3(1.0f) - sq);
    V3 q23 = (bior - V3(eta2)) / (bior + V3(eta2));
    V3 r023 = q23 * q23;
    V3 r23(tf_schlick(r023.x, cos2), tf_schlick(r023.y, cos2), tf_schlick(r023.z, cos2));
    V3 phi23(bior.x < eta2 ? kPi : 0.0f, bior.y < eta2 ? kPi : 0.0f, bior.z < eta2 ? kPi : 0.0f);
    float opd = 2.0f * eta2 * thickness * cos2;
    V3 phi = V3(phi21) + phi23;
    V3 R = V3(r12) * r23;
    V3 R123(std::fmin(std::fmax(R.x, 1e-5f), 0.9999f), std::fmin(std::fmax(R.y, 1e-5f), 0.9999f),
            std::fmin(std::fmax(R.z, 1e-5f), 0.9999f));
    V3 r123(std::sqrt(R123.x), std::sqrt(R123.y), std::sqrt(R123.z));
    V3 Rs = (t121 * t121) * r23 / (V3(1.0f) - R123);
    V3 I = V3(r12) + Rs;
    V3 Cm = Rs - V3(t121);
    for (int mm = 1; mm <= 2; mm++) {
        Cm = Cm * r123;
        V3 Sm = 2.0f * tf_sensitivity((float)mm * opd, (float)mm * phi);
        I = I + Cm * Sm;
    }
    return vmax(I, V3(0.0f));
}

struct Disney : Material {
    V3 base;
    float metallic_, roughness_, specular_, specular_tint_, sheen_, sheen_tint_, sheen_roughness_;
    float clearcoat_, clearcoat_gloss_, spec_trans_, ior_, aniso_, aniso_rot_, diff_trans_;
    float coat_ior_, coat_roughness_, tf_thick_, tf_ior_, trans_depth_;
    bool has_trans_color, has_coat_rough, thin_walled;
    V3 trans_color;

    // ShadingParams (DisneyBsdf.cs:213-293)
    struct SP {
        float metallic, roughness, specular, specular_tint, sheen, sheen_tint, sheen_roughness;
        float clearcoat, clearcoat_gloss, spec_trans, ior, aniso, aniso_rot, diff_trans;
        float alpha, ax, ay, cc_alpha, cc_f0, tf_thick, tf_ior;
    } sp;

    rt_material raw{};
    const rt_material* texture_raw() const override { return (raw.flags & RT_MATF_TEXTURED) ? &raw : nullptr; }
    explicit Disney(const rt_material& m) {
        raw = m;
        base = V3(m.color[0], m.color[1], m.color[2]);
        metallic_ = m.metallic; roughness_ = m.roughness; specular_ = m.specular;
        specular_tint_ = m.specular_tint; sheen_ = m.sheen; sheen_tint_ = m.sheen_tint;
        sheen_roughness_ = m.sheen_roughness; clearcoat_ = m.clearcoat; clearcoat_gloss_ = m.clearcoat_gloss;
        spec_trans_ = m.spec_trans; ior_ = m.ior; aniso_ = m.anisotropic; aniso_rot_ = m.anisotropic_rotation;
        diff_trans_ = m.diff_trans; coat_ior_ = m.coat_ior; coat_roughness_ = m.coat_roughness;
        tf_thick_ = m.thin_film_thickness; tf_ior_ = m.thin_film_ior; trans_depth_ = m.transmission_depth;
        has_trans_color = (m.flags & RT_MATF_TRANSMISSION_COLOR) != 0;
        has_coat_rough = (m.flags & RT_MATF_COAT_ROUGHNESS) != 0;
        thin_walled = (m.flags & RT_MATF_THIN_WALLED) != 0;
        trans_color = V3(m.transmission_color[0], m.transmission_color[1], m.transmission_color[2]);
        SP& s = sp;
        s.metallic = clampf(metallic_, 0, 1);
        s.roughness = clampf(roughness_, 0, 1);
        s.specular = clampf(specular_, 0, 2);
        s.specular_tint = clampf(specular_tint_, 0, 1);
        s.sheen = clampf(sheen_, 0, 1);
        s.sheen_tint = clampf(sheen_tint_, 0, 1);
        s.sheen_roughness = clampf(sheen_roughness_, 0.04f, 1);
        s.clearcoat = clampf(clearcoat_, 0, 1);
        s.clearcoat_gloss = clampf(clearcoat_gloss_, 0, 1);
        s.spec_trans = clampf(spec_trans_, 0, 1);
        s.ior = std::fmax(ior_, 1.0001f);
        s.aniso = clampf(aniso_, 0, 1);
        s.aniso_rot = aniso_rot_ - std::floor(aniso_rot_);
        s.diff_trans = clampf(diff_trans_, 0, 1);
        s.alpha = std::fmax(s.roughness * s.roughness, 0.001f);
        float aspect = std::sqrt(1.0f - 0.9f * s.aniso);
        s.ax = std::fmax(s.alpha / aspect, 0.001f);
        s.ay = std::fmax(s.alpha * aspect, 0.001f);
        float cr = has_coat_rough ? coat_roughness_ : -1.0f;
        s.cc_alpha = cr >= 0.0f ? std::fmax(cr * cr, 0.001f) : lerpf(0.1f, 0.001f, s.clearcoat_gloss);
        float ec = std::fmax(coat_ior_, 1.0001f);
        float r0 = (ec - 1.0f) / (ec + 1.0f);
        s.cc_f0 = r0 * r0;
        s.tf_thick = std::fmax(tf_thick_, 0.0f);
        s.tf_ior = std::fmax(tf_ior_, 1.0001f);
    }

    static constexpr float kDeltaAlpha = 2.5e-3f;
    static float schlick_weight(float c) {
        float x = clampf(1.0f - c, 0.0f, 1.0f);
        float x2 = x * x;
        return x2 * x2 * x;
    }
    V3 compute_f0() const {
        float r = (sp.ior - 1.0f) / (sp.ior + 1.0f);
        float f0d = r * r;
        float scaled = f0d * 2.0f * sp.specular;
        float lum = luminance(base);
        V3 tint = lum > 0.0f ? base / lum : V3(1.0f);
        V3 diel = lerp(V3(scaled), scaled * tint, sp.specular_tint);
        return lerp(diel, base, sp.metallic);
    }
    // EvalFresnel (DisneyBsdf.cs:1984-1989)
    V3 eval_fresnel(float c, V3 f0) const {
        if (sp.tf_thick > 0.0f) return thin_film_fresnel(c, sp.tf_ior, sp.tf_thick, f0);
        float w = schlick_weight(c);
        return f0 + (V3(1.0f) - f0) * w;
    }
    V3 multiscatter_weight(V3 f0, float ndv, float ndl) const {
        const EnergyLut& L = energy_lut();
        float ai = std::sqrt(sp.ax * sp.ay);
        float eo = L.sample_e(ndv, ai), ei = L.sample_e(ndl, ai), ea = L.sample_eavg(ai);
        float om = std::fmax(1.0f - ea, 1e-6f);
        V3 favg = f0 + (V3(1.0f) - f0) * (1.0f / 21.0f);
        V3 den = V3(1.0f) - favg * (1.0f - ea);
        den = V3(std::fmax(den.x, 1e-6f), std::fmax(den.y, 1e-6f), std::fmax(den.z, 1e-6f));
        V3 fms = favg * favg * ea / den;
        return fms * ((1.0f - eo) * (1.0f - ei) / om);
    }
    V3 eval_multiscatter(V3 f0, float ndv, float ndl) const { return multiscatter_weight(f0, ndv, ndl) * (1.0f / kPi); }

    struct Frame {
        V3 t, b, n;
        V3 to_local(V3 w) const { return V3(dot(w, t), dot(w, b), dot(w, n)); }
        V3 to_world(V3 w) const { return w.x * t + w.y * b + w.z * n; }
    };
    // BuildShadingFrame (DisneyBsdf.cs:1825-1866)
    Frame shading_frame(const HitRecord& rec) const {
        V3 n = rec.normal, t, b;
        if (length_sq(rec.tangent) > 1e-10f && length_sq(rec.bitangent) > 1e-10f) {
            t = rec.tangent - dot(rec.tangent, n) * n;
            float tl = length_sq(t);
            if (tl > 1e-10f) { t = t / std::sqrt(tl); b = cross(n, t); }
            else build_tangent_frame(n, t, b);
        } else build_tangent_frame(n, t, b);
        if (sp.aniso_rot != 0.0f) {
            float ang = 2.0f * kPi * sp.aniso_rot;
            float c = std::cos(ang), s = std::sin(ang);
            V3 tr = c * t + s * b;
            V3 br = -s * t + c * b;
            t = tr; b = br;
        }
        return Frame{t, b, n};
    }
    Frame clearcoat_frame(const HitRecord& rec) const {
        Frame f;
        f.n = rec.normal;
        build_tangent_frame(f.n, f.t, f.b);
        return f;
    }

    struct Lobes {
        float d, s, t, c, sh, ms, dt, total;
        float pd() const { return d / total; }
        float ps() const { return s / total; }
        float pt() const { return t / total; }
        float pc() const { return c / total; }
        float psh() const { return sh / total; }
        float pms() const { return ms / total; }
        float pdt() const { return dt / total; }
    };
    // ComputeLobeWeights (DisneyBsdf.cs:1097-1139)
    Lobes lobe_weights() const {
        float dall = (1.0f - sp.metallic) * (1.0f - sp.spec_trans);
        float dw = dall * (1.0f - sp.diff_trans);
        float dtw = dall * sp.diff_trans;
        float sf0;
        if (sp.metallic > 0.5f) sf0 = luminance(base);
        else {
            float r = (sp.ior - 1.0f) / (sp.ior + 1.0f);
            sf0 = r * r * 2.0f * sp.specular;
        }
        float floor_ = 0.1f * (1.0f - sp.spec_trans * 0.9f);
        float sw = std::fmax(floor_, lerpf(sf0, 1.0f, sp.metallic));
        float tw = (1.0f - sp.metallic) * sp.spec_trans;
        float cw = sp.clearcoat * 0.04f;
        float shw = sp.sheen * 0.25f * dall;
        float ai = std::sqrt(sp.ax * sp.ay);
        float ea = energy_lut().sample_eavg(ai);
        float msw = sw * (1.0f - ea);
        Lobes L{dw, sw, tw, cw, shw, msw, dtw, 0};
        float sum = dw + sw + tw + cw + shw + msw + dtw;
        L.total = sum < 1e-6f ? 1.0f : sum;
        return L;
    }
    // ResolveTransmission (DisneyBsdf.cs:1016-1049)
    void resolve_transmission(V3& tint, V3& sigma) const {
        if (!has_trans_color) {
            tint = V3(std::sqrt(std::fmax(base.x, 0.0f)), std::sqrt(std::fmax(base.y, 0.0f)), std::sqrt(std::fmax(base.z, 0.0f)));
            sigma = V3(0.0f);
            return;
        }
        float td = std::fmax(trans_depth_, 0.0f);
        if (td <= 0.0f) { tint = trans_color; sigma = V3(0.0f); return; }
        V3 c(std::fmax(trans_color.x, 1e-6f), std::fmax(trans_color.y, 1e-6f), std::fmax(trans_color.z, 1e-6f));
        float inv = 1.0f / td;
        sigma = V3(-std::log(c.x) * inv, -std::log(c.y) * inv, -std::log(c.z) * inv);
        tint = V3(1.0f);
    }

    V3 aov_albedo(const HitRecord&) const override { return base; }

    // EvaluateDirect (DisneyBsdf.cs:418-535)
    V3 evaluate_direct(V3 to_light, V3 to_eye, V3 normal, const HitRecord& rec) const override {
        float ndl = std::fmax(dot(normal, to_light), 0.0f);
        if (ndl <= 0.0f) return V3(0.0f);
        float ndv = std::fmax(dot(normal, to_eye), 0.001f);
        V3 h = normalize(to_light + to_eye);
        float ndh = std::fmax(dot(normal, h), 0.0f);
        float vdh = std::fmax(dot(to_eye, h), 0.0f);
        float ldh = std::fmax(dot(to_light, h), 0.0f);
        float dw = (1.0f - sp.metallic) * (1.0f - sp.spec_trans) * (1.0f - sp.diff_trans);
        V3 diffuse(0.0f);
        if (dw > 0.0f) {
            float fd90 = 0.5f + 2.0f * sp.roughness * ldh * ldh;
            float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
            float fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
            diffuse = base * fd * (dw * ndl / kPi);
        }
        V3 f0 = compute_f0();
        V3 spec(0.0f);
        if (sp.alpha > kDeltaAlpha) {
            Frame fr = shading_frame(rec);
            V3 vl = fr.to_local(to_eye), ll = fr.to_local(to_light), hl = fr.to_local(h);
            float D = d_ggx_aniso(hl, sp.ax, sp.ay);
            float G = g1_ggx_aniso(vl, sp.ax, sp.ay) * g1_ggx_aniso(ll, sp.ax, sp.ay);
            V3 F = eval_fresnel(vdh, f0);
            spec = D * G * F / std::fmax(4.0f * ndv, 1e-6f);
        }
        V3 cc(0.0f);
        if (sp.clearcoat > 0.0f && sp.cc_alpha > kDeltaAlpha) {
            V3 cn = rec.normal;
            float cndv = std::fmax(dot(cn, to_eye), 0.001f);
            float cndl = dot(cn, to_light);
            if (cndl > 0.0f) {
                float cndh = std::fmax(dot(cn, h), 0.0f);
                float cvdh = std::fmax(dot(to_eye, h), 0.0f);
                float a2 = sp.cc_alpha * sp.cc_alpha;
                float den = cndh * cndh * (a2 - 1.0f) + 1.0f;
                float cD = a2 / (kPi * den * den);
                float cG = smith_g1(cndv, sp.cc_alpha) * smith_g1(cndl, sp.cc_alpha);
                float cF = sp.cc_f0 + (1.0f - sp.cc_f0) * schlick_weight(cvdh);
                cc = V3(sp.clearcoat * 0.25f * cD * cG * cF / std::fmax(4.0f * cndv, 1e-6f));
            }
        }
        V3 sheen(0.0f);
        if (sp.sheen > 0.0f && dw > 0.0f) {
            float lum = luminance(base);
            V3 tint = lum > 0.0f ? base / lum : V3(1.0f);
            V3 scol = lerp(V3(1.0f), tint, sp.sheen_tint);
            float sb = charlie_brdf(ndv, ndl, ndh, sp.sheen_roughness);
            sheen = sp.sheen * sb * scol * ndl;
        }
        V3 ms = eval_multiscatter(f0, ndv, ndl) * ndl;
        return diffuse + sheen + spec + ms + cc;
    }

    // ShadowTransmittance (DisneyBsdf.cs:564-585)
    V3 shadow_transmittance(V3 wi, const HitRecord& rec) const override {
        float st = spec_trans_;
        if (st <= 0.0f) return V3(0.0f);
        float ior = std::fmax(ior_, 1.0001f);
        float eta = 1.0f / ior;
        float ct = std::fmin(std::fabs(dot(wi, rec.normal)), 1.0f);
        float fr = fresnel_dielectric(ct, eta);
        V3 tint, sig;
        resolve_transmission(tint, sig);
        return st * (1.0f - fr) * tint;
    }
    bool is_specular_transmissive(const HitRecord&) const override {
        if (thin_walled) return false;
        return spec_trans_ >= 0.5f && roughness_ <= 0.05f;
    }
    V3 shadow_absorption(const HitRecord&) const override {
        V3 tint, sig;
        resolve_transmission(tint, sig);
        return sig;
    }
    bool try_dielectric_ior(const HitRecord&, float& i) const override {
        if (thin_walled || spec_trans_ <= 0.0f) { i = 1.0f; return false; }
        i = std::fmax(ior_, 1.0001f);
        return true;
    }

    // Evaluate (DisneyBsdf.cs:649-771)
    V3 evaluate(V3 V, V3 L, const HitRecord& rec) const override {
        V3 N = rec.normal;
        float ndl = dot(N, L), ndv = dot(N, V);
        if (ndv <= 0.0f) return V3(0.0f);
        if (ndl <= 0.0f) {
            float dall = (1.0f - sp.metallic) * (1.0f - sp.spec_trans);
            if (dall <= 0.0f || sp.diff_trans <= 0.0f) return V3(0.0f);
            return base * (dall * sp.diff_trans / kPi);
        }
        V3 hr = V + L;
        float hl2 = length_sq(hr);
        if (hl2 < 1e-14f) return V3(0.0f);
        V3 H = hr / std::sqrt(hl2);
        float ndh = std::fmax(dot(N, H), 0.0f);
        float vdh = std::fmax(dot(V, H), 0.0f);
        float ldh = std::fmax(dot(L, H), 0.0f);
        float ds = (1.0f - sp.metallic) * (1.0f - sp.spec_trans) * (1.0f - sp.diff_trans);
        V3 diffuse(0.0f);
        if (ds > 0.0f) {
            float fd90 = 0.5f + 2.0f * sp.roughness * ldh * ldh;
            float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
            float fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
            diffuse = base * fd * (ds / kPi);
        }
        V3 sheen(0.0f);
        if (sp.sheen > 0.0f && ds > 0.0f) {
            float lum = luminance(base);
            V3 tint = lum > 0.0f ? base / lum : V3(1.0f);
            V3 scol = lerp(V3(1.0f), tint, sp.sheen_tint);
            float sb = charlie_brdf(ndv, ndl, ndh, sp.sheen_roughness);
            sheen = sp.sheen * sb * scol;
        }
        Frame fr = shading_frame(rec);
        V3 vl = fr.to_local(V), ll = fr.to_local(L), hlc = fr.to_local(H);
        float D = d_ggx_aniso(hlc, sp.ax, sp.ay);
        float G = g1_ggx_aniso(vl, sp.ax, sp.ay) * g1_ggx_aniso(ll, sp.ax, sp.ay);
        V3 f0 = compute_f0();
        V3 F = eval_fresnel(vdh, f0);
        V3 spec = D * G / std::fmax(4.0f * ndv * ndl, 1e-7f) * F;
        V3 cc(0.0f);
        if (sp.clearcoat > 0.0f) {
            V3 cn = rec.normal;
            float cndv = dot(cn, V), cndl = dot(cn, L);
            if (cndv > 0.0f && cndl > 0.0f) {
                float cndh = std::fmax(dot(cn, H), 0.0f);
                float cvdh = std::fmax(dot(V, H), 0.0f);
                float a2 = sp.cc_alpha * sp.cc_alpha;
                float den = cndh * cndh * (a2 - 1.0f) + 1.0f;
                float cD = a2 / (kPi * den * den);
                float cG = smith_g1(cndv, sp.cc_alpha) * smith_g1(cndl, sp.cc_alpha);
                float cF = sp.cc_f0 + (1.0f - sp.cc_f0) * schlick_weight(cvdh);
                float c = sp.clearcoat * 0.25f * cD * cG * cF / std::fmax(4.0f * cndv * cndl, 1e-7f);
                cc = V3(c);
            }
        }
        V3 ms = eval_multiscatter(f0, ndv, ndl);
        return diffuse + sheen + spec + ms + cc;
    }

    // Pdf (DisneyBsdf.cs:783-870)
    float pdf(V3 V, V3 L, const HitRecord& rec) const override {
        V3 N = rec.normal;
        float ndl = dot(N, L), ndv = dot(N, V);
        if (ndv <= 0.0f) return 0.0f;
        Lobes w = lobe_weights();
        float cos_pdf = ndl > 0.0f ? ndl / kPi : 0.0f;
        float dt_pdf = ndl < 0.0f ? -ndl / kPi : 0.0f;
        float spec_pdf = 0.0f, cc_pdf = 0.0f;
        V3 hr = V + L;
        float hl2 = length_sq(hr);
        if (hl2 >= 1e-14f) {
            V3 H = hr / std::sqrt(hl2);
            float ndh = dot(N, H);
            if (ndh > 0.0f && sp.alpha > kDeltaAlpha) {
                float sndv = std::fmax(ndv, 1e-7f);
                Frame fr = shading_frame(rec);
                V3 vl = fr.to_local(V), hlc = fr.to_local(H);
                float D = d_ggx_aniso(hlc, sp.ax, sp.ay);
                float g1v = g1_ggx_aniso(vl, sp.ax, sp.ay);
                spec_pdf = g1v * D / (4.0f * sndv);
            }
            if (sp.cc_alpha > kDeltaAlpha) {
                V3 cn = rec.normal;
                float cndv = dot(cn, V), cndh = dot(cn, H);
                if (cndv > 0.0f && cndh > 0.0f) {
                    float s = std::fmax(cndv, 1e-7f);
                    float a2 = sp.cc_alpha * sp.cc_alpha;
                    float den = cndh * cndh * (a2 - 1.0f) + 1.0f;
                    float cD = a2 / (kPi * den * den);
                    float g = smith_g1(cndv, sp.cc_alpha);
                    cc_pdf = g * cD / (4.0f * s);
                }
            }
        }
        return w.pd() * cos_pdf + w.psh() * cos_pdf + w.pms() * cos_pdf + w.ps() * spec_pdf + w.pc() * cc_pdf +
               w.pdt() * dt_pdf;
    }

    enum Lobe { kDiffuse, kSpecular, kTransmission, kSheen, kMultiscatter, kDiffTrans, kClearcoat };

    // Lobe samplers (DisneyBsdf.cs:1263-1709)
    bool scatter_diffuse(const HitRecord& rec, V3 N, V3 V, float prob, V3& att, V3& dir) const {
        V3 sd = N + random_unit_vector();
        if (near_zero(sd)) sd = N;
        sd = normalize(sd);
        dir = sd;
        V3 L = sd;
        float ndv = std::fmax(dot(N, V), 0.001f), ndl = std::fmax(dot(N, L), 0.001f);
        V3 hr = V + L;
        float hl2 = length_sq(hr);
        V3 H = hl2 > 1e-7f ? hr / std::sqrt(hl2) : N;
        float ldh = std::fmax(dot(L, H), 0.0f);
        float fd90 = 0.5f + 2.0f * sp.roughness * ldh * ldh;
        float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
        float fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
        att = base * fd;
        att *= (1.0f - sp.diff_trans);
        att /= std::fmax(prob, 0.1f);
        (void)rec;
        return true;
    }
    bool scatter_diff_trans(V3 N, float prob, V3& att, V3& dir) const {
        V3 sd = -N + random_unit_vector();
        if (near_zero(sd)) sd = -N;
        dir = normalize(sd);
        att = base * sp.diff_trans;
        att /= std::fmax(prob, 0.05f);
        return true;
    }
    bool scatter_sheen(V3 N, V3 V, float prob, V3& att, V3& dir) const {
        V3 sd = N + random_unit_vector();
        if (near_zero(sd)) sd = N;
        sd = normalize(sd);
        dir = sd;
        V3 L = sd;
        float ndv = std::fmax(dot(N, V), 1e-4f), ndl = std::fmax(dot(N, L), 1e-4f);
        V3 hr = V + L;
        float hl2 = length_sq(hr);
        V3 H = hl2 > 1e-7f ? hr / std::sqrt(hl2) : N;
        float ndh = std::fmax(dot(N, H), 0.0f);
        float sb = charlie_brdf(ndv, ndl, ndh, sp.sheen_roughness);
        float lum = luminance(base);
        V3 tint = lum > 0.0f ? base / lum : V3(1.0f);
        V3 scol = lerp(V3(1.0f), tint, sp.sheen_tint);
        att = sp.sheen * kPi * sb * scol;
        att /= std::fmax(prob, 0.1f);
        return true;
    }
    bool scatter_specular(const HitRecord& rec, V3 N, V3 V, float prob, V3& att, V3& dir) const {
        Frame fr = shading_frame(rec);
        V3 vl = fr.to_local(V);
        float u1 = random_float();
        float u2 = random_float();
        V3 hl = sample_ggx_vndf(vl, sp.ax, sp.ay, u1, u2);
        V3 H = fr.to_world(hl);
        V3 L = reflect(-V, H);
        if (dot(L, N) <= 0.0f) { att = V3(0.0f); dir = N; return false; }
        dir = L;
        float vdh = std::fmax(dot(V, H), 1e-4f);
        V3 f0 = compute_f0();
        V3 F = eval_fresnel(vdh, f0);
        V3 ll = fr.to_local(L);
        float g1l = g1_ggx_aniso(ll, sp.ax, sp.ay);
        att = F * g1l;
        att /= std::fmax(prob, 0.1f);
        return true;
    }
    bool scatter_transmission(const Ray& rin, const HitRecord& rec, V3 N, V3 V, float prob, V3& att, V3& dir, bool& is_refl) const {
        float eta = thin_walled ? (1.0f / sp.ior)
                                : (rec.relative_eta > 0.0f ? rec.relative_eta : (rec.front_face ? (1.0f / sp.ior) : sp.ior));
        V3 ud = normalize(rin.d);
        Frame fr = shading_frame(rec);
        bool rough = sp.roughness > 0.01f;
        V3 Ht;
        if (rough) {
            V3 vl = fr.to_local(V);
            float u1 = random_float();
            float u2 = random_float();
            V3 hl = sample_ggx_vndf(vl, sp.ax, sp.ay, u1, u2);
            Ht = fr.to_world(hl);
        } else Ht = N;
        if (dot(Ht, ud) > 0.0f) Ht = -Ht;
        float ct = std::fmin(dot(-ud, Ht), 1.0f);
        float st = std::sqrt(std::fmax(0.0f, 1.0f - ct * ct));
        bool cannot = !thin_walled && eta * st > 1.0f;
        V3 d;
        float frs = cannot ? 1.0f : fresnel_dielectric(ct, eta);
        if (frs > random_float()) { d = reflect(ud, Ht); is_refl = true; }
        else if (thin_walled) { d = ud; is_refl = false; }
        else { d = refract(ud, Ht, eta); is_refl = false; }
        if (near_zero(d)) { d = reflect(ud, N); is_refl = true; }
        dir = normalize(d);
        V3 col;
        if (is_refl) col = V3(1.0f);
        else { V3 tint, sig; resolve_transmission(tint, sig); col = tint; }
        att = col;
        if (rough) {
            V3 ll = fr.to_local(dir);
            ll.z = std::fabs(ll.z);
            att *= g1_ggx_aniso(ll, sp.ax, sp.ay);
        }
        att /= std::fmax(prob, 0.1f);
        return true;
    }
    bool scatter_clearcoat(const HitRecord& rec, V3 N, V3 V, float prob, V3& att, V3& dir) const {
        Frame fr = clearcoat_frame(rec);
        V3 cn = fr.n;
        if (dot(cn, V) <= 0.0f) { att = V3(0.0f); dir = N; return false; }
        V3 vl = fr.to_local(V);
        float u1 = random_float();
        float u2 = random_float();
        V3 hl = sample_ggx_vndf(vl, sp.cc_alpha, sp.cc_alpha, u1, u2);
        V3 H = fr.to_world(hl);
        V3 L = reflect(-V, H);
        if (dot(L, N) <= 0.0f || dot(L, cn) <= 0.0f) { att = V3(0.0f); dir = N; return false; }
        dir = L;
        float cndl = std::fmax(dot(cn, L), 1e-4f);
        float cvdh = std::fmax(dot(V, H), 1e-4f);
        float fres = sp.cc_f0 + (1.0f - sp.cc_f0) * schlick_weight(cvdh);
        float g1l = smith_g1(cndl, sp.cc_alpha);
        att = V3(fres * g1l);
        att /= std::fmax(prob, 0.1f);
        return true;
    }
    bool scatter_multiscatter(V3 N, V3 V, float prob, V3& att, V3& dir) const {
        V3 sd = N + random_unit_vector();
        if (near_zero(sd)) sd = N;
        sd = normalize(sd);
        dir = sd;
        float ndv = std::fmax(dot(N, V), 1e-3f), ndl = std::fmax(dot(N, sd), 1e-3f);
        V3 f0 = compute_f0();
        V3 w = multiscatter_weight(f0, ndv, ndl);
        att = w / std::fmax(prob, 0.05f);
        return true;
    }

    // ScatterInternal (DisneyBsdf.cs:1176-1248)
    bool scatter_internal(const Ray& rin, const HitRecord& rec, V3& att, V3& dir, Lobe& lobe, bool& is_refl) const {
        V3 N = rec.normal;
        V3 V = normalize(-rin.d);
        Lobes w = lobe_weights();
        float pD = w.pd(), pS = w.ps(), pT = w.pt(), pSh = w.psh(), pMs = w.pms(), pDt = w.pdt();
        float rnd = random_float();
        bool res;
        is_refl = false;
        if (rnd < pD) { lobe = kDiffuse; res = scatter_diffuse(rec, N, V, pD, att, dir); }
        else if ((rnd -= pD) < pS) { lobe = kSpecular; res = scatter_specular(rec, N, V, pS, att, dir); }
        else if ((rnd -= pS) < pT) { lobe = kTransmission; res = scatter_transmission(rin, rec, N, V, pT, att, dir, is_refl); }
        else if ((rnd -= pT) < pSh) { lobe = kSheen; res = scatter_sheen(N, V, pSh, att, dir); }
        else if ((rnd -= pSh) < pMs) { lobe = kMultiscatter; res = scatter_multiscatter(N, V, pMs, att, dir); }
        else if ((rnd -= pMs) < pDt) { lobe = kDiffTrans; res = scatter_diff_trans(N, pDt, att, dir); }
        else { lobe = kClearcoat; res = scatter_clearcoat(rec, N, V, w.pc(), att, dir); }
        if (std::isnan(att.x) || std::isinf(att.x) || std::isnan(att.y) || std::isinf(att.y) || std::isnan(att.z) ||
            std::isinf(att.z))
            att = base;
        if (raw.flags & RT_MATF_TEXTURED) last_lobe = lobe;   // per-hit instance (adjoint)
        return res;
    }
    mutable int last_lobe = -1;
    // The weight the lobe sampler returned for direction dir, without the
    // division by the lobe probability (restated from the samplers above).
    V3 lobe_att(int lobe, const HitRecord& rec, V3 V, V3 dir) const {
        V3 N = rec.normal;
        if (lobe == kDiffuse) {
            float ndv = std::fmax(dot(N, V), 0.001f), ndl = std::fmax(dot(N, dir), 0.001f);
            V3 hr = V + dir;
            float hl2 = length_sq(hr);
            V3 H = hl2 > 1e-7f ? hr / std::sqrt(hl2) : N;
            float ldh = std::fmax(dot(dir, H), 0.0f);
            float fd90 = 0.5f + 2.0f * sp.roughness * ldh * ldh;
            float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
            float fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
            return base * fd * (1.0f - sp.diff_trans);
        }
        if (lobe == kSpecular) {
            Frame fr = shading_frame(rec);
            V3 H = normalize(V + dir);
            float vdh = std::fmax(dot(V, H), 1e-4f);
            return eval_fresnel(vdh, compute_f0()) * g1_ggx_aniso(fr.to_local(dir), sp.ax, sp.ay);
        }
        if (lobe == kSheen) {
            float ndv = std::fmax(dot(N, V), 1e-4f), ndl = std::fmax(dot(N, dir), 1e-4f);
            V3 hr = V + dir;
            float hl2 = length_sq(hr);
            V3 H = hl2 > 1e-7f ? hr / std::sqrt(hl2) : N;
            float ndh = std::fmax(dot(N, H), 0.0f);
            float lum = luminance(base);
            V3 tint = lum > 0.0f ? base / lum : V3(1.0f);
            return sp.sheen * kPi * charlie_brdf(ndv, ndl, ndh, sp.sheen_roughness) * lerp(V3(1.0f), tint, sp.sheen_tint);
        }
        if (lobe == kMultiscatter) {
            float ndv = std::fmax(dot(N, V), 1e-3f), ndl = std::fmax(dot(N, dir), 1e-3f);
            return multiscatter_weight(compute_f0(), ndv, ndl);
        }
        if (lobe == kDiffTrans) return base * sp.diff_trans;
        return V3(0.0f);
    }

    bool scatter(const Ray& rin, const HitRecord& rec, V3& att, Ray& scattered) const override {
        V3 dir;
        Lobe l;
        bool refl;
        bool r = scatter_internal(rin, rec, att, dir, l, refl);
        scattered = Ray(rec.point, dir);
        return r;
    }

    // Sample (DisneyBsdf.cs:893-997)
    bool sample(V3 V, const HitRecord& rec, BsdfSample& out) const override {
        Ray incoming(rec.point, -V);
        V3 att, wo;
        Lobe lobe;
        bool refl;
        if (!scatter_internal(incoming, rec, att, wo, lobe, refl)) return false;
        float ndw = dot(rec.normal, wo);
        out = BsdfSample();
        out.wo = wo;
        if (lobe == kDiffTrans) {
            V3 fb = evaluate(V, wo, rec);
            float pb = pdf(V, wo, rec);
            if (pb <= 0.0f) { out.f = att; out.pdf = 1.0f; out.is_delta = true; return true; }
            out.f = fb; out.pdf = pb; out.is_delta = false;
            return true;
        }
        if (lobe == kTransmission) {
            out.f = att;
            out.pdf = 1.0f;
            out.is_delta = true;
            if (ndw < 0.0f && !thin_walled) {
                V3 tint, sigma;
                resolve_transmission(tint, sigma);
                bool has = sigma.x > 0.0f || sigma.y > 0.0f || sigma.z > 0.0f;
                if (rec.front_face && has) { out.has_next_absorption = true; out.next_absorption = sigma; }
                else if (!rec.front_face) { out.has_next_absorption = true; out.next_absorption = V3(0.0f); }
                out.transition = rec.front_face ? 1 : 2;
            }
            return true;
        }
        if (ndw <= 0.0f) return false;
        bool nd_spec = lobe == kSpecular && sp.alpha <= kDeltaAlpha;
        bool nd_coat = lobe == kClearcoat && sp.cc_alpha <= kDeltaAlpha;
        if (nd_spec || nd_coat) {
            out.f = att;
            out.pdf = 1.0f;
            out.is_delta = true;
            out.caustic_caster = sp.metallic >= 0.5f;
            return true;
        }
        out.f = evaluate(V, wo, rec);
        out.pdf = pdf(V, wo, rec);
        out.is_delta = false;
        return true;
    }
};

static Material* make_material(const rt_material& m) {
    V3 c(m.color[0], m.color[1], m.color[2]);
    switch (m.kind) {
        case RT_MAT_LAMBERTIAN: return new Lambertian(c);
        case RT_MAT_EMISSIVE: return new Emissive(c, m.intensity, (m.flags & RT_MATF_LIGHT_PROXY) != 0);
        case RT_MAT_DIELECTRIC: return new Dielectric(m.ior, c);
        case RT_MAT_DISNEY: return new Disney(m);
        default: return nullptr;
    }
}

// ═════════════════════════════════════════════════════════════════════════
// ImageTexture (ImageTexture.cs:42-289) and FootprintMath (FootprintMath.cs:38-119)
// ═════════════════════════════════════════════════════════════════════════
struct ImageTex {
    std::vector<std::vector<float>> mip;   // linear RGB, row-major
    std::vector<int> mw, mh;
    float su = 1.0f, sv = 1.0f;
    explicit ImageTex(const rt_image& im) : su(im.scale_u), sv(im.scale_v) {
        // gamma-2.2 decode table of the byte values
        float table[256];
        const float inv255 = 1.0f / 255.0f;
        for (int i = 0; i < 256; i++) table[i] = std::pow(i * inv255, 2.2f);
        int w = im.width, h = im.height;
        std::vector<float> l0((size_t)w * h * 3);
        for (size_t p = 0; p < (size_t)w * h; p++)
            for (int c = 0; c < 3; c++) l0[3 * p + c] = table[im.rgba8[4 * p + c]];
        int levels = 1 + (int)std::floor(std::log2((double)std::max(w, h)));
        mip.push_back(std::move(l0));
        mw.push_back(w);
        mh.push_back(h);
        for (int m = 1; m < levels; m++) {
            int pw = mw[m - 1], ph = mh[m - 1];
            int cw = pw > 1 ? pw / 2 : 1, ch = ph > 1 ? ph / 2 : 1;
            const std::vector<float>& prev = mip[m - 1];
            std::vector<float> cur((size_t)cw * ch * 3);
            for (int y = 0; y < ch; y++)
                for (int x = 0; x < cw; x++) {
                    int x0 = x * 2, y0 = y * 2, x1 = std::min(x0 + 1, pw - 1), y1 = std::min(y0 + 1, ph - 1);
                    for (int c = 0; c < 3; c++) {
                        float s = 0.0f;
                        s += prev[3 * ((size_t)y0 * pw + x0) + c];
                        s += prev[3 * ((size_t)y0 * pw + x1) + c];
                        s += prev[3 * ((size_t)y1 * pw + x0) + c];
                        s += prev[3 * ((size_t)y1 * pw + x1) + c];
                        cur[3 * ((size_t)y * cw + x) + c] = s * 0.25f;
                    }
                }
            mip.push_back(std::move(cur));
            mw.push_back(cw);
            mh.push_back(ch);
        }
    }
    int levels() const { return (int)mip.size(); }
    static int wrap(int i, int n) { i %= n; return i < 0 ? i + n : i; }
    V3 texel(int m, int x, int y) const {
        const float* p = &mip[m][3 * ((size_t)y * mw[m] + x)];
        return V3(p[0], p[1], p[2]);
    }
    V3 bilinear(float u, float v, int m) const {
        int w = mw[m], h = mh[m];
        float px = u * w - 0.5f, py = v * h - 0.5f;
        int x0 = (int)std::floor(px), y0 = (int)std::floor(py);
        float fx = px - x0, fy = py - y0;
        V3 top = lerp(texel(m, wrap(x0, w), wrap(y0, h)), texel(m, wrap(x0 + 1, w), wrap(y0, h)), fx);
        V3 bot = lerp(texel(m, wrap(x0, w), wrap(y0 + 1, h)), texel(m, wrap(x0 + 1, w), wrap(y0 + 1, h)), fx);
        return lerp(top, bot, fy);
    }
    V3 trilinear(float u, float v, float lod) const {
        int lo = (int)std::floor(lod);
        int hi = std::min(lo + 1, levels() - 1);
        float f = lod - lo;
        V3 a = bilinear(u, v, lo);
        if (f <= 0.0f) return a;
        return lerp(a, bilinear(u, v, hi), f);
    }
    V3 value(float u, float v) const {
        u *= su;
        v *= sv;
        u = u - std::floor(u);
        v = v - std::floor(v);
        return bilinear(u, 1.0f - v, 0);
    }
    // the footprint-aware Value (fp all zero: no footprint)
    V3 value(float u, float v, const float fp[4]) const {
        bool has = fp[0] != 0.0f || fp[1] != 0.0f || fp[2] != 0.0f || fp[3] != 0.0f;
        if (!has || levels() <= 1) return value(u, v);
        float dudx = fp[0] * su, dvdx = fp[1] * sv, dudy = fp[2] * su, dvdy = fp[3] * sv;
        float lx = std::sqrt(dudx * dudx + dvdx * dvdx), ly = std::sqrt(dudy * dudy + dvdy * dvdy);
        float minor = std::min(lx, ly), major = std::max(lx, ly);
        if (major <= 0.0f) return value(u, v);
        if (minor * 8 < major) minor = major / 8;
        float maxd = (float)std::max(mw[0], mh[0]);
        float lod = std::log2(std::max(minor * maxd, 1.0f));
        lod = std::min(std::max(lod, 0.0f), (float)(levels() - 1));
        int n = (int)std::ceil(major / minor);
        n = std::min(std::max(n, 1), 8);
        float mu = lx >= ly ? dudx : dudy, mv = lx >= ly ? dvdx : dvdy;
        float uc = u * su, vc = v * sv;
        V3 acc(0.0f);
        float inv = 1.0f / n;
        for (int i = 0; i < n; i++) {
            float t = (i + 0.5f) * inv - 0.5f;
            float uu = uc + t * mu, vv = vc + t * mv;
            uu = uu - std::floor(uu);
            vv = vv - std::floor(vv);
            acc += trilinear(uu, 1.0f - vv, lod);
        }
        return acc * inv;
    }
};

static bool plane_hit_point(V3 o, V3 d, V3 P, V3 N, V3& hit) {
    float den = dot(d, N);
    if (std::fabs(den) < 1e-12f) return false;
    float t = dot(P - o, N) / den;
    hit = o + t * d;
    return std::isfinite(t);
}

// FootprintMath.Compute: the UV partials of ray r's differentials at the
// hit (point P, normal N) against dpdu/dpdv; zero when none.
static void footprint(const Ray& r, V3 P, V3 N, V3 dpdu, V3 dpdv, float out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    V3 px, py;
    if (!plane_hit_point(r.dox, r.ddx, P, N, px) || !plane_hit_point(r.dox, r.ddy, P, N, py)) return;
    V3 dx = px - P, dy = py - P;
    float ax = std::fabs(N.x), ay = std::fabs(N.y), az = std::fabs(N.z);
    int i0, i1;
    if (ax > ay && ax > az) { i0 = 1; i1 = 2; }
    else if (ay > az) { i0 = 0; i1 = 2; }
    else { i0 = 0; i1 = 1; }
    float a00 = dpdu[i0], a01 = dpdv[i0], a10 = dpdu[i1], a11 = dpdv[i1];
    float det = a00 * a11 - a01 * a10;
    if (std::fabs(det) < 1e-20f) return;
    float inv = 1.0f / det;
    out[0] = (a11 * dx[i0] - a01 * dx[i1]) * inv;
    out[1] = (-a10 * dx[i0] + a00 * dx[i1]) * inv;
    out[2] = (a11 * dy[i0] - a01 * dy[i1]) * inv;
    out[3] = (-a10 * dy[i0] + a00 * dy[i1]) * inv;
}

// ═════════════════════════════════════════════════════════════════════════
// Transparent shadow rays (ShadowRay.cs:59-158)
// ═════════════════════════════════════════════════════════════════════════
static thread_local bool t_block_spec_trans = false;
static thread_local int64_t t_shadow_segments = 0;

static V3 shadow_transmittance(const Hittable* world, const Ray& ray, float tmin, float tmax) {
    V3 thr(1.0f);
    V3 o = ray.o, d = ray.d;
    float rem = tmax;
    V3 sigma(0.0f), entry;
    bool in_med = false;
    for (int i = 0; i < 8; i++) {
        HitRecord rec;
        Ray step(o, d);
        t_shadow_segments++;
        if (!world->hit(step, tmin, rem, rec)) return thr;
        if (rec.mat == nullptr) return V3(0.0f);
        if (t_block_spec_trans && rec.mat->is_specular_transmissive(rec)) return V3(0.0f);
        if (in_med) {
            float dd = length(rec.point - entry);
            thr = V3(thr.x * std::exp(-sigma.x * dd), thr.y * std::exp(-sigma.y * dd), thr.z * std::exp(-sigma.z * dd));
            in_med = false;
            if (near_zero(thr)) return V3(0.0f);
        }
        V3 t = rec.mat->shadow_transmittance(d, rec);
        if (near_zero(t)) return V3(0.0f);
        thr *= t;
        if (near_zero(thr)) return V3(0.0f);
        if (rec.front_face) {
            V3 s = rec.mat->shadow_absorption(rec);
            if (!near_zero(s)) { in_med = true; entry = rec.point; sigma = s; }
        }
        V3 on = dot(d, rec.normal) >= 0.0f ? rec.normal : -rec.normal;
        o = offset_origin(rec.point, on);
        rem -= rec.t + 2.0f * kEpsilon;
        if (rem <= tmin) return thr;
    }
    return V3(0.0f);
}

// ── CausticPhotonTracer sampling helpers (CausticPhotonTracer.cs:348-410) ──
// FaceToward: flips n to point from `from` toward the scene-bounds centre
static V3 face_toward(V3 n, V3 from, const AABB& b) {
    V3 c = 0.5f * (b.mn + b.mx);
    return dot(n, c - from) < 0.0f ? -n : n;
}
// CosineDir: normalize(n + RandomUnitVector), degenerate → n
static V3 cosine_dir(V3 n) {
    V3 d = n + random_unit_vector();
    if (length_sq(d) < 1e-8f) d = n;
    return normalize(d);
}
static void photon_onb(V3 w, V3& t, V3& b) {
    V3 a = std::fabs(w.x) > 0.9f ? V3(0, 1, 0) : V3(1, 0, 0);
    t = normalize(cross(a, w));
    b = cross(w, t);
}
// RandomDisk: r = sqrt(xi1), angle = 2 pi xi2
static void photon_disk(float& x, float& y) {
    float r = std::sqrt(random_float());
    float a = 2.0f * kPi * random_float();
    x = r * std::cos(a);
    y = r * std::sin(a);
}

// ═════════════════════════════════════════════════════════════════════════
// Lights (ILight.cs)
// ═════════════════════════════════════════════════════════════════════════
struct LightSample {
    bool in_shadow;
    V3 color, dir;
    float dist;
};

struct Light {
    int shadow_samples = 1;
    const Material* proxy = nullptr;
    virtual ~Light() {}
    virtual bool is_delta() const { return true; }
    virtual float pdf_solid_angle(V3, V3) const { return 0.0f; }
    virtual float approximate_power(const AABB&) const = 0;
    virtual LightSample illuminate(V3 p, V3 n, const Hittable* world, int si) const = 0;
    // CausticPhotonTracer.TryEmit (CausticPhotonTracer.cs:279-345): one photon
    // ray and the flux of the whole light
    virtual bool emit_photon(const AABB& b, Ray& ray, V3& flux) const = 0;
    // CausticPhotonTracer.CanEmit (CausticPhotonTracer.cs:270-275)
    virtual bool can_emit() const { return true; }
};

// SphereLight.cs:112-271
struct SphereLight : Light {
    V3 c, color;
    float r, inten;
    int sq;
    float inv_sq;
    SphereLight(V3 center, float radius, V3 col, float i, int ss, const Material* px) {
        c = center; r = std::fmax(radius, kEpsilon); color = col; inten = i;
        shadow_samples = std::max(1, ss);
        proxy = px;
        sq = (int)std::ceil(std::sqrt((float)shadow_samples));
        inv_sq = 1.0f / (float)sq;
    }
    float approximate_power(const AABB&) const override { return 4.0f * kPi * luminance(color) * inten; }
    bool is_delta() const override { return false; }
    bool emit_photon(const AABB&, Ray& ray, V3& flux) const override {
        V3 dir = random_unit_vector();
        V3 p = c + dir * r;
        ray = Ray(offset_origin(p, dir), cosine_dir(dir));
        flux = color * inten * (4.0f * kPi);
        return true;
    }
    LightSample illuminate(V3 hp, V3 sn, const Hittable* world, int si) const override {
        V3 to_c = c - hp;
        float d2 = length_sq(to_c);
        if (d2 < kEpsilon * kEpsilon) return {true, V3(0.0f), V3(0, 1, 0), 0.0f};
        float dist = std::sqrt(d2);
        V3 w = to_c / dist;
        float s2 = r * r / d2;
        float ctm = s2 >= 1.0f ? -1.0f : std::sqrt(1.0f - s2);
        float xi1, xi2;
        if (si >= 0) {
            int su = si % sq, sv = si / sq;
            xi1 = ((float)su + random_float()) * inv_sq;
            xi2 = ((float)sv + random_float()) * inv_sq;
        } else { xi1 = random_float(); xi2 = random_float(); }
        float ct = 1.0f - xi1 * (1.0f - ctm);
        float st = std::sqrt(std::fmax(0.0f, 1.0f - ct * ct));
        float phi = 2.0f * kPi * xi2;
        V3 dl(st * std::cos(phi), st * std::sin(phi), ct);
        V3 dw = local_to_world(dl, w);
        V3 oc = hp - c;
        float b = dot(oc, dw);
        float cc = length_sq(oc) - r * r;
        float disc = b * b - cc;
        if (disc < 0.0f) {
            V3 cp = c + normalize(dw) * r;
            return {true, V3(0.0f), dw, length(cp - hp)};
        }
        float sd = std::sqrt(disc);
        float t = -b - sd;
        if (t < kEpsilon) t = -b + sd;
        if (t < kEpsilon) return {true, V3(0.0f), dw, 0.0f};
        V3 sp = hp + t * dw;
        V3 ln = (sp - c) / r;
        float cl = std::fmax(0.0f, dot(-dw, ln));
        if (cl <= 0.0f) return {false, V3(0.0f), dw, t};
        V3 so = offset_origin(hp, sn);
        Ray sr(so, dw);
        float stmax = length(sp - so) - kEpsilon;
        V3 tr = shadow_transmittance(world, sr, kEpsilon, stmax);
        if (near_zero(tr)) return {true, V3(0.0f), dw, t};
        float sa = 2.0f * kPi * (1.0f - ctm);
        float att = inten * sa / (float)shadow_samples;
        return {false, color * att * tr, dw, t};
    }
    float pdf_solid_angle(V3 hp, V3 wi) const override {
        V3 to_c = c - hp;
        float d2 = length_sq(to_c);
        float r2 = r * r;
        if (d2 <= r2) return 1.0f / (4.0f * kPi);
        float s2 = r2 / d2;
        float ctm = std::sqrt(std::fmax(0.0f, 1.0f - s2));
        float wl = length(wi);
        if (wl < kEpsilon) return 0.0f;
        float dist = std::sqrt(d2);
        V3 w = to_c / dist;
        float ct = dot(wi, w) / wl;
        if (ct < ctm) return 0.0f;
        float sa = 2.0f * kPi * (1.0f - ctm);
        return sa > kEpsilon ? 1.0f / sa : 0.0f;
    }
};

// AreaLight.cs:97-226
struct AreaLight : Light {
    V3 corner, u, v, color, n;
    float inten, soft, area;
    int sq;
    float inv_sq;
    AreaLight(V3 c, V3 U, V3 V, V3 col, float i, int ss, float sr, const Material* px) {
        corner = c; u = U; v = V; color = col; inten = i;
        shadow_samples = std::max(1, ss);
        soft = std::fmax(0.0f, sr);
        proxy = px;
        V3 cr = cross(u, v);
        area = length(cr);
        n = cr / area;
        sq = (int)std::ceil(std::sqrt((float)shadow_samples));
        inv_sq = 1.0f / (float)sq;
    }
    bool is_delta() const override { return false; }
    float approximate_power(const AABB&) const override { return kPi * area * luminance(color) * inten; }
    bool emit_photon(const AABB& b, Ray& ray, V3& flux) const override {
        float r1 = random_float();
        float r2 = random_float();
        V3 p = (corner + r1 * u) + r2 * v;
        V3 nn = normalize(cross(u, v));
        nn = face_toward(nn, (corner + 0.5f * u) + 0.5f * v, b);
        float ar = length(cross(u, v));
        ray = Ray(offset_origin(p, nn), cosine_dir(nn));
        flux = color * inten * ar * kPi;
        return true;
    }
    LightSample illuminate(V3 hp, V3 sn, const Hittable* world, int si) const override {
        V3 sp;
        if (si >= 0) {
            int su = si % sq, sv = si / sq;
            float ru = ((float)su + random_float()) * inv_sq;
            float rv = ((float)sv + random_float()) * inv_sq;
            sp = (corner + ru * u) + rv * v;
        } else {
            float a = random_float();
            float b = random_float();
            sp = (corner + a * u) + b * v;
        }
        V3 tl = sp - hp;
        float d2 = length_sq(tl);
        if (d2 < kEpsilon * kEpsilon) return {true, V3(0.0f), V3(0, 1, 0), 0.0f};
        float dist = std::sqrt(d2);
        V3 dl = tl / dist;
        float cl = std::fmax(0.0f, dot(-dl, n));
        if (cl <= 0.0f) return {false, V3(0.0f), dl, dist};
        V3 so = offset_origin(hp, sn);
        Ray sr(so, dl);
        float stmax = length(sp - so) - kEpsilon;
        V3 tr = shadow_transmittance(world, sr, kEpsilon, stmax);
        if (near_zero(tr)) return {true, V3(0.0f), dl, dist};
        float ad2 = d2;
        if (soft > 0.0f) ad2 = std::fmax(ad2, soft * soft);
        float att = inten * area * cl / (ad2 * (float)shadow_samples);
        return {false, color * att * tr, dl, dist};
    }
    float pdf_solid_angle(V3 hp, V3 wi) const override {
        float denom = dot(n, wi);
        if (std::fabs(denom) < 1e-7f) return 0.0f;
        float t = dot(corner - hp, n) / denom;
        if (t <= kEpsilon) return 0.0f;
        V3 p = hp + t * wi;
        V3 lc = p - corner;
        float uu = dot(u, u), vv = dot(v, v), uv = dot(u, v);
        float det = uu * vv - uv * uv;
        if (det < 1e-12f) return 0.0f;
        float lu = dot(lc, u), lv = dot(lc, v);
        float a = (lu * vv - lv * uv) / det;
        float b = (lv * uu - lu * uv) / det;
        if (a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f) return 0.0f;
        float cl = std::fabs(denom);
        if (cl < 1e-6f) return 0.0f;
        return (t * t) / (area * cl);
    }
};

// DirectionalLight.cs:60-181
struct DirectionalLight : Light {
    V3 dir, color;
    float inten, ang_deg, cos_ang, solid;
    int sq;
    float inv_sq;
    DirectionalLight(V3 d, V3 col, float i, float ang, int ss) {
        dir = normalize(d); color = col; inten = i;
        ang_deg = std::fmax(0.0f, ang);
        bool disc = ang_deg > 0.0f;
        float rad = ang_deg * kPi / 180.0f;
        cos_ang = disc ? std::cos(rad) : 1.0f;
        solid = disc ? 2.0f * kPi * (1.0f - cos_ang) : 0.0f;
        shadow_samples = ss > 0 ? ss : (disc ? 4 : 1);
        sq = (int)std::ceil(std::sqrt((float)shadow_samples));
        inv_sq = 1.0f / (float)sq;
    }
    bool is_delta() const override { return ang_deg <= 0.0f; }
    float approximate_power(const AABB& b) const override {
        V3 e = b.mx - b.mn;
        float rad = 0.5f * length(e);
        float cs = kPi * rad * rad;
        return luminance(color) * inten * cs;
    }
    bool emit_photon(const AABB& b, Ray& ray, V3& flux) const override {
        V3 c = 0.5f * (b.mn + b.mx);
        float r = std::fmax(0.5f * length(b.mx - b.mn), 1e-3f);
        V3 w = dir;
        V3 t, bb;
        photon_onb(w, t, bb);
        float dx, dy;
        photon_disk(dx, dy);
        V3 origin = (c - w * (r * 1.5f)) + (t * dx + bb * dy) * r;
        ray = Ray(origin, w);
        flux = color * inten * (kPi * r * r);
        return true;
    }
    LightSample illuminate(V3 hp, V3 sn, const Hittable* world, int si) const override {
        V3 tl;
        if (ang_deg > 0.0f) {
            float xi1, xi2;
            if (si >= 0) {
                int su = si % sq, sv = si / sq;
                xi1 = ((float)su + random_float()) * inv_sq;
                xi2 = ((float)sv + random_float()) * inv_sq;
            } else { xi1 = random_float(); xi2 = random_float(); }
            float ct = 1.0f - xi1 * (1.0f - cos_ang);
            float st = std::sqrt(std::fmax(0.0f, 1.0f - ct * ct));
            float phi = 2.0f * kPi * xi2;
            V3 w = -dir;
            V3 ls(st * std::cos(phi), st * std::sin(phi), ct);
            tl = local_to_world(ls, w);
        } else tl = -dir;
        V3 so = offset_origin(hp, sn);
        Ray sr(so, tl);
        V3 tr = shadow_transmittance(world, sr, kEpsilon, kInfinity);
        if (near_zero(tr)) return {true, V3(0.0f), tl, kInfinity};
        float e = inten / (float)shadow_samples;
        return {false, color * e * tr, tl, kInfinity};
    }
    float pdf_solid_angle(V3, V3 wi) const override {
        if (ang_deg <= 0.0f) return 0.0f;
        if (solid <= 0.0f) return 0.0f;
        float wl = length(wi);
        if (wl < kEpsilon) return 0.0f;
        float ct = dot(wi / wl, -dir);
        if (ct < cos_ang) return 0.0f;
        return 1.0f / solid;
    }
};

// EnvironmentLight.cs:20-124 over a FlatSky (FlatSky.cs:10-40,
// SkySettings.SampleDirectly / PdfSolidAngle, SkySettings.cs:315-337)
struct EnvironmentLight : Light {
    V3 sky;
    bool sampleable;   // FlatSky.HasImportanceSampling
    EnvironmentLight(V3 s, int ss) {
        sky = s;
        shadow_samples = std::max(1, ss);
        sampleable = luminance(s) > 1e-6f;
    }
    bool is_delta() const override { return false; }
    bool can_emit() const override { return false; }
    float approximate_power(const AABB& b) const override {
        if (!sampleable) return 0.0f;
        V3 e = b.mx - b.mn;
        float cross_section = 0.5f * (e.x * e.y + e.y * e.z + e.x * e.z);
        float irradiance = kPi * luminance(sky);
        return irradiance * cross_section;
    }
    bool emit_photon(const AABB&, Ray&, V3&) const override { return false; }
    LightSample illuminate(V3 hp, V3 sn, const Hittable* world, int) const override {
        if (!sampleable) return {true, V3(0.0f), V3(0, 1, 0), 0.0f};
        V3 dir = random_unit_vector();
        const float pdf = 1.0f / (4.0f * kPi);
        float dist = kInfinity;
        if (length_sq(sn) > 0.0f && dot(sn, dir) <= 0.0f) return {true, V3(0.0f), dir, dist};
        V3 so = offset_origin(hp, sn);
        Ray sr(so, dir);
        V3 tr = shadow_transmittance(world, sr, kEpsilon, kInfinity);
        if (near_zero(tr)) return {true, V3(0.0f), dir, dist};
        return {false, (sky * tr) / (pdf * (float)shadow_samples), dir, dist};
    }
    float pdf_solid_angle(V3, V3) const override { return sampleable ? 1.0f / (4.0f * kPi) : 0.0f; }
};

// GeometryLight.cs:60-279 for TRIANGLE / QUAD (area sampling) and SPHERE
// (solid-angle sampling)
struct GeometryLight : Light {
    const Triangle* tri = nullptr;
    const Quad* quad = nullptr;
    const Sphere* sph = nullptr;
    const Emissive* em;
    float soft;
    int sq;
    V3 rep_point;
    float rep_area;
    GeometryLight(const Hittable* g, const Emissive* m, int ss, float sr) : em(m) {
        tri = dynamic_cast<const Triangle*>(g);
        quad = dynamic_cast<const Quad*>(g);
        sph = dynamic_cast<const Sphere*>(g);
        shadow_samples = std::max(1, ss);
        soft = std::fmax(0.0f, sr);
        sq = (int)std::ceil(std::sqrt((float)shadow_samples));
        AABB b = g->box();
        rep_point = 0.5f * (b.mn + b.mx);
        rep_area = tri ? tri->area : (quad ? quad->area : 4.0f * kPi * sph->r * sph->r);
    }
    bool is_delta() const override { return false; }
    float approximate_power(const AABB&) const override { return kPi * rep_area * luminance(em->emission_at()); }
    bool emit_photon(const AABB& b, Ray& ray, V3& flux) const override {
        V3 p, nn;
        float ar;
        if (sph) {
            V3 d = random_unit_vector();
            ar = 4.0f * kPi * sph->r * sph->r;
            p = sph->c + d * sph->r;
            nn = d;
        } else if (tri) {
            float xi1 = random_float(), xi2 = random_float();
            float r1 = std::sqrt(xi1);
            float uu = 1.0f - r1, vv = xi2 * r1;
            p = tri->v0 + uu * tri->e1 + vv * tri->e2;
            nn = tri->n;
            ar = tri->area;
        } else {
            float uu = random_float(), vv = random_float();
            p = (quad->q + uu * quad->u) + vv * quad->v;
            nn = quad->n;
            ar = quad->area;
        }
        V3 e = em->emission_at();
        if (e.x <= 0.0f && e.y <= 0.0f && e.z <= 0.0f) return false;
        nn = face_toward(nn, p, b);
        ray = Ray(offset_origin(p, nn), cosine_dir(nn));
        flux = e * ar * kPi;
        return true;
    }
    LightSample illuminate(V3 hp, V3 sn, const Hittable* world, int si) const override {
        if (sph) {
            float inv = 1.0f / (float)sq;
            int su = si % sq, sv = si / sq;
            float xi1 = ((float)su + random_float()) * inv;
            float xi2 = ((float)sv + random_float()) * inv;
            V3 sp, ln;
            float pdf;
            sph->sample_solid_angle(hp, xi1, xi2, sp, ln, pdf);
            if (pdf <= 0.0f) return {true, V3(0.0f), V3(0, 1, 0), 0.0f};
            V3 tl = sp - hp;
            float d2 = length_sq(tl);
            if (d2 < kEpsilon * kEpsilon) return {true, V3(0.0f), V3(0, 1, 0), 0.0f};
            float dist = std::sqrt(d2);
            V3 dl = tl / dist;
            float cl = std::fmax(0.0f, dot(-dl, ln));
            if (cl <= 0.0f) return {false, V3(0.0f), dl, dist};
            V3 so = offset_origin(hp, sn);
            Ray sr(so, dl);
            float stmax = length(sp - so) - kEpsilon;
            V3 tr = shadow_transmittance(world, sr, kEpsilon, stmax);
            if (near_zero(tr)) return {true, V3(0.0f), dl, dist};
            float att = 1.0f / (pdf * (float)shadow_samples);
            return {false, em->emission_at() * att * tr, dl, dist};
        }
        V3 sp, ln;
        float ar;
        if (tri) tri->sample_stratified(si, sq, sp, ln, ar);
        else quad->sample_stratified(si, sq, sp, ln, ar);
        V3 tl = sp - hp;
        float d2 = length_sq(tl);
        if (d2 < kEpsilon * kEpsilon) return {true, V3(0.0f), V3(0, 1, 0), 0.0f};
        float dist = std::sqrt(d2);
        V3 dl = tl / dist;
        float cl = std::fmax(0.0f, dot(-dl, ln));
        if (cl <= 0.0f) return {true, V3(0.0f), dl, dist};
        V3 so = offset_origin(hp, sn);
        Ray sr(so, dl);
        float stmax = length(sp - so) - kEpsilon;
        V3 tr = shadow_transmittance(world, sr, kEpsilon, stmax);
        if (near_zero(tr)) return {true, V3(0.0f), dl, dist};
        float ad2 = d2;
        if (soft > 0.0f) ad2 = std::fmax(ad2, soft * soft);
        float att = ar * cl / (ad2 * (float)shadow_samples);
        return {false, em->emission_at() * att * tr, dl, dist};
    }
    float pdf_solid_angle(V3 hp, V3 wi) const override {
        if (sph) return sph->solid_angle_pdf(hp, wi);
        const Hittable* h = tri ? (const Hittable*)tri : (const Hittable*)quad;
        Ray r(hp, wi);
        HitRecord rec;
        if (!h->hit(r, kEpsilon, kInfinity, rec)) return 0.0f;
        float cl = std::fabs(dot(-wi, rec.normal));
        if (cl < 1e-6f) return 0.0f;
        float d2 = rec.t * rec.t;
        if (rep_area < 1e-12f) return 0.0f;
        return d2 / (rep_area * cl);
    }
};

// LightDistribution.cs:41-135
struct LightDist {
    std::vector<float> cdf, pdf;
    LightDist(const std::vector<Light*>& lights, const AABB& b, bool uniform) {
        int n = (int)lights.size();
        cdf.assign(n, 0.0f);
        pdf.assign(n, 0.0f);
        if (n == 0) return;
        std::vector<float> pw(n, 0.0f);
        float sum = 0.0f;
        if (!uniform)
            for (int i = 0; i < n; i++) { pw[i] = std::fmax(0.0f, lights[i]->approximate_power(b)); sum += pw[i]; }
        if (sum > 0.0f) {
            float inv = 1.0f / sum, cum = 0.0f;
            for (int i = 0; i < n; i++) { pdf[i] = pw[i] * inv; cum += pdf[i]; cdf[i] = cum; }
        } else {
            float u = 1.0f / (float)n, cum = 0.0f;
            for (int i = 0; i < n; i++) { pdf[i] = u; cum += u; cdf[i] = cum; }
        }
        cdf[n - 1] = 1.0f;
    }
    void sample(float xi, int& idx, float& p) const {
        int n = (int)cdf.size();
        if (n == 0) { idx = 0; p = 1.0f; return; }
        if (n == 1) { idx = 0; p = pdf[0]; return; }
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (cdf[mid] < xi) lo = mid + 1;
            else hi = mid;
        }
        idx = lo;
        p = pdf[lo];
    }
};

// ═════════════════════════════════════════════════════════════════════════
// Scene (SceneLoader.cs:282-553 world assembly) + camera (Camera.cs:26-106)
// ═════════════════════════════════════════════════════════════════════════
struct Camera {
    V3 origin, llc, horiz, vert, u, v;
    float lens_radius;
    Camera() {}
    explicit Camera(const rt_camera& c) {
        V3 lf(c.look_from[0], c.look_from[1], c.look_from[2]);
        V3 la(c.look_at[0], c.look_at[1], c.look_at[2]);
        V3 vup(c.vup[0], c.vup[1], c.vup[2]);
        float theta = c.vfov_deg * kPi / 180.0f;
        float h = std::tan(theta / 2.0f);
        float vh = 2.0f * h;
        float vw = c.aspect * vh;
        V3 w = normalize(lf - la);
        if (length_sq(cross(w, vup)) < 1e-8f) vup = std::fabs(w.x) < 0.9f ? V3(1, 0, 0) : V3(0, 0, 1);
        u = normalize(cross(vup, w));
        v = cross(w, u);
        origin = lf;
        horiz = c.focus_dist * vw * u;
        vert = c.focus_dist * vh * v;
        llc = origin - horiz / 2.0f - vert / 2.0f - c.focus_dist * w;
        lens_radius = c.aperture / 2.0f;
    }
    // GetRayWithDifferentials without motion (Camera.cs:127-153) with
    // pixel deltas dsdx, dtdy; GetRay (Camera.cs:79-105) when texture
    // filtering is off (diff = false): no differentials, and the lens ray
    // subtracts the origin and the lens offset in turn.
    Ray get_ray(float s, float t, float dsdx = 0.0f, float dtdy = 0.0f, bool diff = true) const {
        V3 target = llc + s * horiz + t * vert;
        Ray r;
        V3 lo = origin;
        if (lens_radius == 0.0f) r = Ray(origin, target - origin);
        else {
            V3 rd = lens_radius * random_in_unit_disk();
            V3 off = u * rd.x + v * rd.y;
            lo = origin + off;
            r = diff ? Ray(lo, target - lo) : Ray(lo, target - origin - off);
        }
        if (diff) {
            r.has_diff = true;
            r.dox = lo;
            r.ddx = llc + (s + dsdx) * horiz + t * vert - lo;
            r.ddy = llc + s * horiz + (t - dtdy) * vert - lo;
        }
        return r;
    }
};

struct Scene {
    std::vector<std::unique_ptr<Material>> materials;
    std::vector<Hittable*> owned;        // every allocated Hittable (deleted at the end)
    std::vector<Hittable*> prim_objs;    // per primitive id
    std::vector<Hittable*> objects;      // SceneLoader object list
    Hittable* bvh = nullptr;             // BvhNode over finite objects (or null)
    HittableList* world = nullptr;
    std::vector<std::unique_ptr<Light>> lights;
    std::vector<Light*> light_ptrs;
    Camera camera;
    V3 sky;
    const struct EnvironmentLight* env_light = nullptr;
    bool medium = false;
    V3 sig_a, sig_s, sig_t;
    AABB bounds;
    float scene_radius = 0, mean_irr = 0, total_flux = 0;
    int rr_min_bounces = 3;
    float rr_min_survival = 0.1f;
    std::vector<std::pair<const Material*, const Light*>> emitter_to_light;
    // texture set of textured Disney materials: planar [5][tex_h][tex_w],
    // optional vertex UVs (6 per primitive)
    std::vector<float> tex, tuv;
    int tex_w = 0, tex_h = 0;
    // Texture coordinate of a hit: interpolated vertex UVs of a triangle
    // (SmoothTriangle.cs:168,186) when given, else HitRecord.U/V.
    void tex_coord(const HitRecord& rec, bool is_tri, float& u, float& v) const {
        u = rec.u;
        v = rec.v;
        if (tuv.empty() || !is_tri || rec.prim < 0) return;
        const float* t = &tuv[6 * (size_t)rec.prim];
        float w1 = rec.u, w2 = rec.v, w0 = 1.0f - w1 - w2;
        u = (w0 * t[0] + w1 * t[2]) + w2 * t[4];
        v = (w0 * t[1] + w1 * t[3]) + w2 * t[5];
    }
    // Bilinear footprint: texel centres at (i + 0.5) / width, u wraps, v clamps.
    void tex_foot(float u, float v, int idx[4], float w[4]) const {
        float x = u * (float)tex_w - 0.5f, y = v * (float)tex_h - 0.5f;
        float x0 = std::floor(x), y0 = std::floor(y);
        float fx = x - x0, fy = y - y0;
        long long ix = (long long)x0;
        int i0 = (int)(((ix % tex_w) + tex_w) % tex_w), i1 = (i0 + 1) % tex_w;
        int j0 = (int)std::min<float>(std::max<float>(y0, 0.0f), (float)(tex_h - 1));
        int j1 = (int)std::min<float>(std::max<float>(y0 + 1.0f, 0.0f), (float)(tex_h - 1));
        idx[0] = j0 * tex_w + i0; idx[1] = j0 * tex_w + i1; idx[2] = j1 * tex_w + i0; idx[3] = j1 * tex_w + i1;
        w[0] = (1.0f - fx) * (1.0f - fy); w[1] = fx * (1.0f - fy); w[2] = (1.0f - fx) * fy; w[3] = fx * fy;
    }
    float tex_at(const int idx[4], const float w[4], int ch) const {
        const float* pl = &tex[(size_t)ch * tex_w * tex_h];
        return ((w[0] * pl[idx[0]] + w[1] * pl[idx[1]]) + w[2] * pl[idx[2]]) + w[3] * pl[idx[3]];
    }
    std::vector<uint8_t> tri_flag;   // per primitive: TRIANGLE
    std::vector<std::unique_ptr<ImageTex>> images;
    bool textured = false;
    // Texture lookup of a textured hit: filtered parameters and footprint.
    struct TexHit {
        const rt_material* raw;
        float p[5];        // base r, g, b, roughness, metallic
        int idx[4];
        float w[4];
    };
    TexHit tex_hit(const rt_material& raw, const HitRecord& rec) const {
        TexHit t;
        t.raw = &raw;
        float u, v;
        bool tri = rec.prim >= 0 && rec.prim < (int)tri_flag.size() && tri_flag[rec.prim];
        tex_coord(rec, tri, u, v);
        tex_foot(u, v, t.idx, t.w);
        for (int c = 0; c < 5; c++) t.p[c] = tex_at(t.idx, t.w, c);
        return t;
    }
    // The per-hit Disney of a textured material: Disney(rt_material) with
    // the texture's baseColor / roughness / metallic, parameter j moved by dp.
    static Disney material_at(const TexHit& t, int j = -1, float dp = 0.0f) {
        float q[5];
        for (int c = 0; c < 5; c++) q[c] = t.p[c] + (c == j ? dp : 0.0f);
        rt_material m = *t.raw;
        m.color[0] = q[0];
        m.color[1] = q[1];
        m.color[2] = q[2];
        m.roughness = q[3];
        m.metallic = q[4];
        return Disney(m);
    }
    ~Scene() {
        delete world;
        for (Hittable* h : owned) delete h;
    }
    const Light* light_for(const Material* m) const {
        const Light* r = nullptr;
        for (auto& p : emitter_to_light)
            if (p.first == m) r = p.second;
        return r;
    }
};

// ═════════════════════════════════════════════════════════════════════════
// Caustic photon pass (CausticPhotonTracer.cs:60-356, PhotonMap.cs:39-317)
// ═════════════════════════════════════════════════════════════════════════
struct PhotonO {
    V3 pos, inc, pow;
};

// CausticPhotonTracer.Trace (CausticPhotonTracer.cs:134-231): follows delta
// caustic-caster bounces (and delta legacy Scatter) until the first
// non-specular hit, which deposits the photon iff at least one specular
// bounce happened.  Returns whether a photon was deposited.
static bool photon_trace(const Hittable* world, Ray ray, V3 power, PhotonO& out) {
    float emax = fmax_p(power.x, fmax_p(power.y, power.z));
    if (emax <= 0.0f) return false;
    float inv_emax = 1.0f / emax;
    int spec = 0;
    bool in_med = false;
    V3 sig(0.0f), entry(0.0f);
    for (int bounce = 0; bounce < 24; bounce++) {
        HitRecord rec;
        if (!world->hit(ray, kEpsilon, kInfinity, rec)) return false;
        if (in_med) {
            float d = length(rec.point - entry);
            power = V3(power.x * std::exp(-sig.x * d), power.y * std::exp(-sig.y * d), power.z * std::exp(-sig.z * d));
            in_med = false;
        }
        const Material* m = rec.mat;
        if (!m) return false;
        V3 view = normalize(-ray.d);
        auto enter_medium = [&](V3 wo) {
            bool transmitted = dot(wo, rec.normal) < 0.0f;
            if (!transmitted || !rec.front_face) return;
            V3 sa = m->shadow_absorption(rec);
            if (sa.x <= 0.0f && sa.y <= 0.0f && sa.z <= 0.0f) return;
            in_med = true;
            sig = sa;
            entry = rec.point;
        };
        auto deposit = [&]() {
            if (spec < 1) return false;
            if (power.x <= 0.0f && power.y <= 0.0f && power.z <= 0.0f) return false;
            out.pos = rec.point;
            out.inc = normalize(-ray.d);
            out.pow = power;
            return true;
        };
        BsdfSample bs;
        if (m->sample(view, rec, bs)) {
            if (bs.is_delta && bs.caustic_caster) {
                power *= bs.f;
                enter_medium(bs.wo);
                V3 od = dot(bs.wo, rec.normal) >= 0.0f ? rec.normal : -rec.normal;
                ray = Ray(offset_origin(rec.point, od), bs.wo);
                spec++;
            } else {
                return deposit();
            }
        } else {
            V3 att;
            Ray sc;
            if (!m->scatter(ray, rec, att, sc)) return false;
            if (!m->is_delta_scatter()) return deposit();
            power *= att;
            enter_medium(sc.d);
            ray = sc;
            spec++;
        }
        if (spec >= 4) {
            float survive = std::fmin(1.0f, fmax_p(power.x, fmax_p(power.y, power.z)) * inv_emax);
            if (survive <= 1e-6f || random_float() >= survive) return false;
            power = power / survive;
        }
    }
    return false;
}

// PhotonMap (PhotonMap.cs:39-317): Teschner-hashed uniform grid in CSR form.
// Photons are kept in global photon-index order (the reference merges
// per-thread lists in scheduling order; any order gives the same map up to
// the order of photons inside a bucket).
struct PhotonMapO {
    std::vector<PhotonO> ph;
    std::vector<int32_t> start, idx, key;
    V3 origin;
    float cell = 0, inv_cell = 0;
    int mask = 0;

    void build(float cell_size) {
        cell = std::fmax(cell_size, 1e-6f);
        inv_cell = 1.0f / cell;
        int n = (int)ph.size();
        int buckets = 1;
        while (buckets < n) buckets <<= 1;
        mask = buckets - 1;
        origin = V3(0.0f);
        if (n > 0) {
            V3 mn(std::numeric_limits<float>::infinity());
            for (auto& p : ph) mn = V3(std::fmin(mn.x, p.pos.x), std::fmin(mn.y, p.pos.y), std::fmin(mn.z, p.pos.z));
            origin = mn;
        }
        // stable counting sort into CSR buckets
        start.assign(buckets + 1, 0);
        key.resize(n);
        for (int i = 0; i < n; i++) {
            key[i] = bucket_of(ph[i].pos);
            start[key[i] + 1]++;
        }
        for (int b = 1; b <= buckets; b++) start[b] += start[b - 1];
        idx.assign(n, 0);
        std::vector<int32_t> cur(start.begin(), start.end() - 1);
        for (int i = 0; i < n; i++) idx[cur[key[i]]++] = i;
    }
    void cell_of(V3 p, int& x, int& y, int& z) const {
        x = (int)std::floor((p.x - origin.x) * inv_cell);
        y = (int)std::floor((p.y - origin.y) * inv_cell);
        z = (int)std::floor((p.z - origin.z) * inv_cell);
    }
    int bucket(int x, int y, int z) const {
        uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
        return (int)(h & (uint32_t)mask);
    }
    int bucket_of(V3 p) const {
        int x, y, z;
        cell_of(p, x, y, z);
        return bucket(x, y, z);
    }
    float cell_min_d2(V3 p, int x, int y, int z) const {
        float bx = origin.x + (float)x * cell, by = origin.y + (float)y * cell, bz = origin.z + (float)z * cell;
        float dx = std::fmax(std::fmax(bx - p.x, p.x - (bx + cell)), 0.0f);
        float dy = std::fmax(std::fmax(by - p.y, p.y - (by + cell)), 0.0f);
        float dz = std::fmax(std::fmax(bz - p.z, p.z - (bz + cell)), 0.0f);
        return (dx * dx + dy * dy) + dz * dz;
    }
    static void heap_insert(int* hi, float* hd, int& count, int k, int pi, float d2) {
        if (count < k) {
            int i = count++;
            hi[i] = pi; hd[i] = d2;
            while (i > 0) {
                int parent = (i - 1) >> 1;
                if (hd[parent] >= hd[i]) break;
                std::swap(hd[parent], hd[i]);
                std::swap(hi[parent], hi[i]);
                i = parent;
            }
        } else if (d2 < hd[0]) {
            hi[0] = pi; hd[0] = d2;
            int i = 0;
            for (;;) {
                int l = 2 * i + 1, r = 2 * i + 2, lg = i;
                if (l < count && hd[l] > hd[lg]) lg = l;
                if (r < count && hd[r] > hd[lg]) lg = r;
                if (lg == i) break;
                std::swap(hd[lg], hd[i]);
                std::swap(hi[lg], hi[i]);
                i = lg;
            }
        }
    }
    // GatherKNearest (PhotonMap.cs:209-240) with ScanShell (:243-276)
    int gather(V3 p, int k, float max_r, int* hi, float* hd, float& radius) const {
        radius = 0.0f;
        if (ph.empty() || k <= 0) return 0;
        int cx, cy, cz;
        cell_of(p, cx, cy, cz);
        float max_r2 = max_r * max_r;
        int count = 0;
        int max_rings = (int)std::ceil(max_r * inv_cell) + 1;
        for (int ring = 0; ring <= max_rings; ring++) {
            for (int z = cz - ring; z <= cz + ring; z++)
                for (int y = cy - ring; y <= cy + ring; y++)
                    for (int x = cx - ring; x <= cx + ring; x++) {
                        int cheb = std::max(std::abs(x - cx), std::max(std::abs(y - cy), std::abs(z - cz)));
                        if (cheb != ring) continue;
                        float md2 = cell_min_d2(p, x, y, z);
                        if (md2 > max_r2) continue;
                        if (count >= k && md2 >= hd[0]) continue;
                        int b = bucket(x, y, z);
                        for (int kk = start[b]; kk < start[b + 1]; kk++) {
                            int pi = idx[kk];
                            V3 pos = ph[pi].pos;
                            int px, py, pz;
                            cell_of(pos, px, py, pz);
                            if (px != x || py != y || pz != z) continue;
                            float d2 = length_sq(pos - p);
                            if (d2 > max_r2) continue;
                            heap_insert(hi, hd, count, k, pi, d2);
                        }
                    }
            float reach = (float)ring * cell;
            if (count >= k && hd[0] <= reach * reach) break;
            if (reach > max_r) break;
        }
        radius = count > 0 ? std::sqrt(hd[0]) : 0.0f;
        return count;
    }
};

// CausticPhotonTracer.Build (CausticPhotonTracer.cs:60-127): per-light photon
// budgets by luminous power, a Sobol sub-sequence per source
// (seed = si * 0x9E3779B1 ^ 0xC2B2AE35, index = photon index within the source).
static bool build_photon_map(const Scene& sc, int budget, float cell_size, int nthreads, PhotonMapO& map,
                             int64_t* emitted = nullptr) {
    map = PhotonMapO();
    if (budget <= 0) return false;
    std::vector<const Light*> src;
    std::vector<float> lum;
    float total = 0.0f;
    for (const Light* l : sc.light_ptrs) {
        if (!l->can_emit()) continue;
        float p = std::fmax(0.0f, l->approximate_power(sc.bounds));
        if (p <= 0.0f) continue;
        src.push_back(l);
        lum.push_back(p);
        total += p;
    }
    if (src.empty() || total <= 0.0f) return false;
    std::vector<int> counts(src.size()), offs(src.size() + 1, 0);
    for (size_t i = 0; i < src.size(); i++) {
        int n = (int)std::nearbyint((float)budget * (lum[i] / total));
        counts[i] = std::max(1, n);
        offs[i + 1] = offs[i] + counts[i];
    }
    const int ntot = offs[src.size()];
    if (emitted) *emitted = ntot;
    std::vector<PhotonO> slot(ntot);
    std::vector<uint8_t> ok(ntot, 0);
    std::atomic<int> next(0);
    auto worker = [&]() {
        for (;;) {
            int g0 = next.fetch_add(4096);
            if (g0 >= ntot) break;
            int g1 = std::min(ntot, g0 + 4096);
            for (int g = g0; g < g1; g++) {
                int si = (int)(std::upper_bound(offs.begin(), offs.end(), g) - offs.begin()) - 1;
                int local = g - offs[si];
                SamplerCtx& c = t_smp;
                c.active = true;
                c.seed = ((uint32_t)si * 0x9E3779B1u) ^ 0xC2B2AE35u;
                c.index = (uint32_t)local;
                c.dim = 0;
                c.fixed = nullptr;
                Ray ray;
                V3 flux;
                if (src[si]->emit_photon(sc.bounds, ray, flux)) {
                    V3 scale = V3(1.0f) / (float)counts[si];
                    ok[g] = photon_trace(sc.world, ray, flux * scale, slot[g]) ? 1 : 0;
                }
                c.active = false;
            }
        }
    };
    if (nthreads <= 1) worker();
    else {
        std::vector<std::thread> th;
        for (int k = 0; k < nthreads; k++) th.emplace_back(worker);
        for (auto& x : th) x.join();
    }
    for (int g = 0; g < ntot; g++)
        if (ok[g]) map.ph.push_back(slot[g]);
    if (map.ph.empty()) return false;
    map.build(cell_size);
    return true;
}

}  // namespace orc

namespace orc {

// IorStack.cs:33-107
struct IorStack {
    float s[8];
    int n = 0;
    float top() const { return n == 0 ? 1.0f : s[n - 1]; }
    float enclosing() const { return n >= 2 ? s[n - 2] : 1.0f; }
    void push(float v) {
        if (n >= 8) {
            for (int i = 0; i < 7; i++) s[i] = s[i + 1];
            s[7] = v;
            return;
        }
        s[n++] = v;
    }
    void pop() { if (n > 0) n--; }
};

// Renderer.AovProbe (Renderer.cs:777-815)
struct AovProbe {
    bool pending = true, have_depth = false;
    float depth = -1.0f;
    V3 alb_w = V3(1.0f), tent_alb = V3(1.0f), tent_nrm = V3(0.0f), albedo = V3(0.0f), normal = V3(0.0f);
    void commit() { albedo = alb_w * tent_alb; normal = tent_nrm; pending = false; }
    void commit_as(V3 a, V3 n) { albedo = alb_w * a; normal = n; pending = false; }
};
static V3 clamp01(V3 v) { return V3(clampf(v.x, 0, 1), clampf(v.y, 0, 1), clampf(v.z, 0, 1)); }

enum RayCat { kCatCamera = 0, kCatDiffuse, kCatGlossy, kCatTransmission };

// ═════════════════════════════════════════════════════════════════════════
// FP64 adjoint of the texture parameters (no reference counterpart).  The
// render recursion runs unchanged; aw is the weight (dLoss/dPixel / spp times
// every bounce weight down to the current vertex) of the radiance the current
// ShadeSurface returns.  At a textured hit the derivative of the BSDF with
// respect to the filtered parameters p = (base r, g, b, roughness, metallic)
// is taken by central differences of the fp32 Disney evaluation (step
// kAdjH, sampling, pdfs and MIS weights held at p), accumulated in double and
// spread over the bilinear footprint:
//   NEE:    aw · k_l · dBrdf(wl)/dp      (k_l = wn Le tr terminator / ppick)
//   bounce: aw · dF(wo)/dp |cos|/pdf / q · Li(wo)   (Li = the returned radiance)
// Delta and fallback (Scatter) bounces carry the weight but add no term.
// ═════════════════════════════════════════════════════════════════════════
constexpr float kAdjH = 1e-3f;
struct AdjState {
    double aw[3] = {0, 0, 0};
    const Scene::TexHit* hit = nullptr;   // set while a textured hit's NEE runs
    double* grad = nullptr;               // [5][tex_h][tex_w]
    int64_t plane = 0;
    void add(const Scene::TexHit& t, const double g[5]) {
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < 5; j++) grad[j * plane + t.idx[k]] += (double)t.w[k] * g[j];
    }
};
static thread_local AdjState* t_adj = nullptr;
// Scales the adjoint weight by a bounce weight for the duration of a call.
struct AdjWeight {
    AdjState* A;
    double saved[3];
    AdjWeight(AdjState* a, V3 att) : A(a) {
        if (!A) return;
        for (int c = 0; c < 3; c++) saved[c] = A->aw[c];
        A->aw[0] *= att.x;
        A->aw[1] *= att.y;
        A->aw[2] *= att.z;
    }
    void restore() {
        if (A)
            for (int c = 0; c < 3; c++) A->aw[c] = saved[c];
    }
};

struct Renderer {
    const Scene& sc;
    int max_depth, spp;
    float max_rad, ind_max_rad, exposure;
    bool power_mis;
    std::unique_ptr<LightDist> ldist;
    bool capture_aovs;
    int rr_min_b;
    float rr_min_s;
    AovProbe* probe = nullptr;   // per-thread (set by the render loop)
    // photon caustics (Renderer.cs:130-152, 380-406)
    const PhotonMapO* pmap = nullptr;
    bool caustics = false;
    float gather_r = 0.0f;
    enum { kCausticNone = 0, kCausticDiffuse = 1, kCausticSuppress = 2 };
    static int next_caustic(int st, bool delta) {
        if (!delta) return kCausticDiffuse;
        return st == kCausticNone ? kCausticNone : kCausticSuppress;
    }
    // GatherCaustics (Renderer.cs:1449-1508) + CausticOccupancyWeight (:1520-1526)
    V3 gather_caustics(const HitRecord& rec, V3 view, const Material* m) const {
        return caustic_estimate(*pmap, rec, view, m, gather_r);
    }
    static V3 caustic_estimate(const PhotonMapO& map, const HitRecord& rec, V3 view, const Material* m, float gather_r) {
        const PhotonMapO* pmap = &map;
        const int K = 40;
        int hi[K];
        float hd[K];
        float radius;
        int count = pmap->gather(rec.point, K, gather_r, hi, hd, radius);
        if (count == 0) return V3(0.0f);
        bool under = count < K;
        if (under) radius = gather_r;
        if (radius <= 0.0f) return V3(0.0f);
        float inv_r = 1.0f / radius;
        V3 sum(0.0f);
        for (int i = 0; i < count; i++) {
            const PhotonO& p = pmap->ph[hi[i]];
            if (dot(p.inc, rec.normal) <= 0.0f) continue;
            float w = 1.0f - length(p.pos - rec.point) * inv_r / 1.0f;
            if (w <= 0.0f) continue;
            sum += m->evaluate(view, p.inc, rec) * (p.pow * w);
        }
        const float norm = 1.0f - 2.0f / (3.0f * 1.0f);
        V3 est = sum * (1.0f / (norm * kPi * radius * radius));
        if (under) {
            float t = (float)count / (float)K;
            est *= t * t * (3.0f - 2.0f * t);
        }
        return est;
    }

    Renderer(const Scene& s, const rt_settings& st) : sc(s) {
        max_depth = st.max_depth;
        spp = st.samples_per_pixel;
        max_rad = st.max_sample_radiance;
        ind_max_rad = max_rad * std::fmax(0.0f, st.indirect_clamp_factor);
        exposure = std::pow(2.0f, st.exposure_ev);
        power_mis = st.mis == RT_MIS_POWER;
        capture_aovs = st.capture_aovs != 0;
        if (st.light_sampling != RT_LIGHTS_ALL && !s.light_ptrs.empty())
            ldist.reset(new LightDist(s.light_ptrs, s.bounds, st.light_sampling == RT_LIGHTS_UNIFORM));
        rr_min_b = st.rr_min_bounces >= 0 ? st.rr_min_bounces : s.rr_min_bounces;
        rr_min_s = st.rr_min_survival >= 0.0f ? st.rr_min_survival : s.rr_min_survival;
    }

    static V3 sanitize(V3 c) {
        if (std::isnan(c.x) || std::isinf(c.x)) c.x = 0.0f;
        if (std::isnan(c.y) || std::isinf(c.y)) c.y = 0.0f;
        if (std::isnan(c.z) || std::isinf(c.z)) c.z = 0.0f;
        return vmax(c, V3(0.0f));
    }
    V3 clamp_radiance(V3 c) const {
        c = sanitize(c);
        float l = luminance(c);
        if (l > max_rad) c *= max_rad / l;
        return c;
    }
    V3 clamp_indirect(V3 c) const {
        c = sanitize(c);
        float l = luminance(c);
        if (l > ind_max_rad && l > 0.0f) c *= ind_max_rad / l;
        return c;
    }
    float mis_weight(float p, float q) const {
        if (power_mis) { float p2 = p * p, q2 = q * q; return p2 / (p2 + q2 + 1e-30f); }
        return p / (p + q + 1e-30f);
    }
    // WeightEmission (Renderer.cs:1549-1565)
    V3 weight_emission(V3 raw, const Material* m, const Ray& ray, float prev_pdf, bool prev_delta) const {
        if (prev_delta) return raw;
        const Light* l = m->is_emissive() ? sc.light_for(m) : nullptr;
        if (l) {
            float pl = l->pdf_solid_angle(ray.o, ray.d);
            if (prev_pdf + pl <= 1e-20f) return V3(0.0f);
            return raw * mis_weight(prev_pdf, pl);
        }
        return raw;
    }
    static V3 beer_lambert(V3 r, V3 sig, float t) {
        if (sig.x <= 0.0f && sig.y <= 0.0f && sig.z <= 0.0f) return r;
        if (std::isinf(t) && t > 0) return V3(0.0f);
        if (t <= 0.0f) return r;
        return r * V3(sig.x > 0.0f ? std::exp(-sig.x * t) : 1.0f, sig.y > 0.0f ? std::exp(-sig.y * t) : 1.0f,
                      sig.z > 0.0f ? std::exp(-sig.z * t) : 1.0f);
    }
    V3 medium_tr(float tmax) const {
        float d = std::fmin(std::fmax(tmax, 0.0f), 1e30f);
        return V3(std::exp(-sc.sig_t.x * d), std::exp(-sc.sig_t.y * d), std::exp(-sc.sig_t.z * d));
    }
    // HomogeneousMedium.Sample (HomogeneousMedium.cs:50-76)
    bool medium_sample(float tmax, float& t, V3& beta) const { return medium_sample_of(sc.sig_t, sc.sig_s, tmax, t, beta); }
    static V3 medium_tr_of(V3 sig_t, float tmax) {
        float d = std::fmin(std::fmax(tmax, 0.0f), 1e30f);
        return V3(std::exp(-sig_t.x * d), std::exp(-sig_t.y * d), std::exp(-sig_t.z * d));
    }
    static bool medium_sample_of(V3 sig_t, V3 sig_s, float tmax, float& t, V3& beta) {
        int ch = (int)(random_float() * 3.0f);
        if (ch > 2) ch = 2;
        float st = ch == 0 ? sig_t.x : (ch == 1 ? sig_t.y : sig_t.z);
        if (st <= 0.0f) { t = tmax; beta = V3(1.0f); return false; }
        float dist = -std::log(1.0f - random_float()) / st;
        bool scat = dist < tmax;
        t = scat ? dist : tmax;
        V3 tr = medium_tr_of(sig_t, t);
        V3 dens = scat ? sig_t * tr : tr;
        float pdf = (dens.x + dens.y + dens.z) / 3.0f;
        if (pdf < 1e-20f) pdf = 1.0f;
        beta = scat ? (tr * sig_s / pdf) : (tr / pdf);
        return scat;
    }

    // SampleSky (Renderer.cs:1820-1861): the flat sky, MIS-weighed against
    // the first environment light's NEE after a non-delta bounce
    V3 sample_sky(const Ray& ray, float prev_pdf, bool prev_delta) const {
        V3 sky = sc.sky;
        if (capture_aovs && probe && probe->pending) probe->commit_as(clamp01(sky), V3(0.0f));
        const EnvironmentLight* env = sc.env_light;
        if (prev_delta || !env || !env->sampleable) return sky;
        float pl = env->pdf_solid_angle(ray.o, ray.d);
        if (prev_pdf + pl <= 1e-20f) return V3(0.0f);
        return sky * mis_weight(prev_pdf, pl);
    }

    // ComputeDirectLighting (Renderer.cs:1894-2036)
    V3 direct_lighting(const HitRecord& rec, V3 view, const Material* m) const {
        V3 result(0.0f);
        float view_term = shadow_terminator(rec.shading_ng, rec.normal, view);
        auto one = [&](const Light* light, int s, float ppick, V3& acc) {
            LightSample ls = light->illuminate(rec.point, rec.normal, sc.world, s);
            if (ls.in_shadow) return;
            V3 brdf = m ? m->evaluate_direct(ls.dir, view, rec.normal, rec)
                        : V3(std::fmax(dot(rec.normal, ls.dir), 0.0f) / kPi);
            brdf *= view_term * shadow_terminator(rec.shading_ng, rec.normal, ls.dir);
            if (brdf == V3(0.0f)) return;
            V3 tr(1.0f);
            if (sc.medium) {
                float sd = std::isinf(ls.dist) ? 1e30f : ls.dist;
                tr = medium_tr(sd);
            }
            float wn = 1.0f;
            if (!light->is_delta() && m) {
                float pb = m->pdf(view, ls.dir, rec);
                if (pb > 0.0f) {
                    float pl = light->pdf_solid_angle(rec.point, ls.dir);
                    float pn = ppick > 0.0f ? ppick * pl : pl;
                    wn = (pn + pb > 0.0f) ? mis_weight(pn, pb) : 1.0f;
                }
            }
            acc += wn * ls.color * brdf * tr;
            if (t_adj && t_adj->hit) {
                // NEE term: d(brdf)/dp at the light direction, MIS weight held fixed
                const Scene::TexHit& th = *t_adj->hit;
                float stl = view_term * shadow_terminator(rec.shading_ng, rec.normal, ls.dir);
                V3 k = wn * ls.color * tr * stl;
                double inv = ppick > 0.0f ? 1.0 / (double)ppick : 1.0;
                double g[5];
                for (int j = 0; j < 5; j++) {
                    V3 fp = Scene::material_at(th, j, kAdjH).evaluate_direct(ls.dir, view, rec.normal, rec);
                    V3 fm = Scene::material_at(th, j, -kAdjH).evaluate_direct(ls.dir, view, rec.normal, rec);
                    double h2 = 2.0 * (double)kAdjH;
                    g[j] = inv * (t_adj->aw[0] * k.x * (((double)fp.x - fm.x) / h2) +
                                  t_adj->aw[1] * k.y * (((double)fp.y - fm.y) / h2) +
                                  t_adj->aw[2] * k.z * (((double)fp.z - fm.z) / h2));
                }
                t_adj->add(th, g);
            }
        };
        if (ldist) {
            if (sc.light_ptrs.empty()) return result;
            float xi = random_float();
            int li;
            float pp;
            ldist->sample(xi, li, pp);
            const Light* light = sc.light_ptrs[li];
            V3 acc(0.0f);
            for (int s = 0; s < light->shadow_samples; s++) one(light, s, pp, acc);
            if (pp > 0.0f) result += acc / pp;
        } else {
            for (const Light* light : sc.light_ptrs) {
                V3 acc(0.0f);
                for (int s = 0; s < light->shadow_samples; s++) one(light, s, -1.0f, acc);
                result += acc;
            }
        }
        return result;
    }

    // ComputeDirectLightingMedium (Renderer.cs:2099-2206), isotropic phase
    V3 direct_lighting_medium(V3 p) const {
        V3 result(0.0f);
        V3 dn(0.0f);
        auto one = [&](const Light* light, int s, float ppick, V3& acc) {
            LightSample ls = light->illuminate(p, dn, sc.world, s);
            if (ls.in_shadow) return;
            float pv = kInv4Pi;
            float sd = std::isinf(ls.dist) ? 1e30f : ls.dist;
            V3 tr = medium_tr(sd);
            float wn = 1.0f;
            if (!light->is_delta()) {
                float pp = kInv4Pi;
                if (pp > 0.0f) {
                    float pl = light->pdf_solid_angle(p, ls.dir);
                    float pn = ppick > 0.0f ? ppick * pl : pl;
                    if (pn + pp > 0.0f) wn = mis_weight(pn, pp);
                }
            }
            acc += wn * ls.color * pv * tr;
        };
        if (ldist) {
            if (sc.light_ptrs.empty()) return result;
            float xi = random_float();
            int li;
            float pp;
            ldist->sample(xi, li, pp);
            const Light* light = sc.light_ptrs[li];
            V3 acc(0.0f);
            for (int s = 0; s < light->shadow_samples; s++) one(light, s, pp, acc);
            if (pp > 0.0f) result += acc / pp;
        } else {
            for (const Light* light : sc.light_ptrs) {
                V3 acc(0.0f);
                for (int s = 0; s < light->shadow_samples; s++) one(light, s, -1.0f, acc);
                result += acc;
            }
        }
        return result;
    }

    static int classify(V3 n, V3 wo, bool delta) {
        if (dot(n, wo) < 0.0f) return kCatTransmission;
        return delta ? kCatGlossy : kCatDiffuse;
    }

    // ShadeSurface (Renderer.cs:1008-1228)
    V3 shade_surface(const Ray& ray, HitRecord& rec, int depth, float prev_pdf, bool prev_delta, V3 absorption, V3 beta,
                     IorStack& iors, int cst) const {
        std::unique_ptr<Disney> tex_mat;
        Scene::TexHit th;
        const Scene::TexHit* thp = nullptr;
        std::unique_ptr<Material> img_mat;
        if (rec.mat && rec.mat->image >= 0) {
            // ITexture.Value(in HitRecord): HitRecord U/V and the footprint
            float u, v;
            bool tri = rec.prim >= 0 && rec.prim < (int)sc.tri_flag.size() && sc.tri_flag[rec.prim];
            sc.tex_coord(rec, tri, u, v);
            V3 c = sc.images[rec.mat->image]->value(u, v, rec.fp);
            rt_material m = rec.mat->image_raw;
            m.color[0] = c.x;
            m.color[1] = c.y;
            m.color[2] = c.z;
            img_mat.reset(make_material(m));
            rec.mat = img_mat.get();
        }
        if (rec.mat && rec.mat->texture_raw()) {
            th = sc.tex_hit(*rec.mat->texture_raw(), rec);
            thp = &th;
            tex_mat.reset(new Disney(Scene::material_at(th)));
            rec.mat = tex_mat.get();
        }
        const Material* m = rec.mat;
        rec.shading_ng = rec.normal;
        if (capture_aovs && probe && probe->pending) {
            probe->tent_nrm = rec.normal;
            probe->tent_alb = m ? clamp01(m->aov_albedo(rec)) : V3(1.0f);
        }
        V3 emitted(0.0f);
        if (m && !(caustics && cst == kCausticSuppress)) {
            V3 raw = m->emit(rec.front_face);
            if (raw.x > 0.0f || raw.y > 0.0f || raw.z > 0.0f) emitted = weight_emission(raw, m, ray, prev_pdf, prev_delta);
        }
        V3 view = normalize(-ray.d);
        float mior;
        if (m && m->try_dielectric_ior(rec, mior) && mior > 0.0f)
            rec.relative_eta = rec.front_face ? (iors.top() / mior) : (mior / iors.enclosing());
        bool needs = m ? m->needs_direct_lighting() : true;
        AdjState* A = t_adj;
        if (A) A->hit = thp;
        V3 direct = needs ? direct_lighting(rec, view, m) : V3(0.0f);
        if (A) A->hit = nullptr;
        if (caustics && needs && m && !m->is_delta_scatter()) direct += gather_caustics(rec, view, m);
        if (!m) return emitted + direct;

        BsdfSample s;
        if (m->sample(view, rec, s))
            return shade_sample_bounce(m, rec, s, depth, emitted, direct, absorption, beta, iors, cst, view, thp);

        V3 att;
        Ray scat;
        if (m->scatter(ray, rec, att, scat)) {
            if (capture_aovs && probe && probe->pending) {
                if (m->is_delta_scatter()) probe->alb_w *= clamp01(att);
                else probe->commit();
            }
            V3 nt = beta * att;
            int used = max_depth - depth;
            if (used >= rr_min_b) {
                float sp = fmax_p(nt.x, fmax_p(nt.y, nt.z));
                sp = fmax_p(sp, rr_min_s);
                sp = fmin_p(sp, 0.95f);
                if (random_float() > sp) return emitted + direct;
                float inv = 1.0f / sp;
                att *= inv;
                nt *= inv;
            }
            bool nd = m->is_delta_scatter();
            int cat = classify(rec.normal, scat.d, nd);
            IorStack ni = iors;
            float sior;
            if (m->try_dielectric_ior(rec, sior) && dot(scat.d, rec.normal) < 0.0f) {
                if (rec.front_face) ni.push(sior);
                else ni.pop();
            }
            AdjWeight aw(A, att);
            V3 ind = trace_ray(scat, depth - 1, 0.0f, nd, ni, absorption, nt, cat, next_caustic(cst, nd));
            aw.restore();
            ind = sanitize(ind);
            if (A && thp) adj_lobe_term(A, th, static_cast<const Disney*>(m)->last_lobe, rec, view, scat.d, att, ind);
            V3 contrib = att * ind;
            if (depth == max_depth) contrib = clamp_indirect(contrib);
            return emitted + direct + contrib;
        }
        return emitted + direct;
    }

    // ShadeSampleBounce (Renderer.cs:1235-1383)
    // Bounce term of a delta or fallback (Scatter) bounce: d(lobe weight)/dp
    // at the sampled direction, times the fixed 1/(lobe probability * q).
    static void adj_lobe_term(AdjState* A, const Scene::TexHit& th, int lobe, const HitRecord& rec, V3 V, V3 dir, V3 att,
                              V3 ind) {
        V3 w0 = Scene::material_at(th).lobe_att(lobe, rec, V, dir);
        double g[5];
        for (int j = 0; j < 5; j++) {
            V3 wp = Scene::material_at(th, j, kAdjH).lobe_att(lobe, rec, V, dir);
            V3 wm = Scene::material_at(th, j, -kAdjH).lobe_att(lobe, rec, V, dir);
            double acc = 0.0;
            for (int c = 0; c < 3; c++) {
                double wv = c == 0 ? w0.x : (c == 1 ? w0.y : w0.z);
                if (wv == 0.0) continue;
                double scale = (c == 0 ? att.x : (c == 1 ? att.y : att.z)) / wv;   // 1 / (prob * q)
                double dw = ((c == 0 ? wp.x : (c == 1 ? wp.y : wp.z)) - (c == 0 ? wm.x : (c == 1 ? wm.y : wm.z))) /
                            (2.0 * (double)kAdjH);
                acc += A->aw[c] * dw * scale * (c == 0 ? ind.x : (c == 1 ? ind.y : ind.z));
            }
            g[j] = acc;
        }
        A->add(th, g);
    }

    V3 shade_sample_bounce(const Material* m, const HitRecord& rec, const BsdfSample& s, int depth, V3 emitted, V3 direct,
                           V3 absorption, V3 beta, IorStack& iors, int cst, V3 view = V3(0.0f),
                           const Scene::TexHit* th = nullptr) const {
        V3 att;
        if (s.is_delta) att = s.f;
        else {
            float ndw = dot(rec.normal, s.wo);
            float a = std::fabs(ndw);
            if (a <= 0.0f || s.pdf <= 0.0f) return emitted + direct;
            att = s.f * a / s.pdf;
        }
        if (capture_aovs && probe && probe->pending) {
            if (s.is_delta) probe->alb_w *= clamp01(s.f);
            else probe->commit();
        }
        V3 nt = beta * att;
        int used = max_depth - depth;
        float rr_inv = 1.0f;
        if (used >= rr_min_b) {
            float sp = fmax_p(nt.x, fmax_p(nt.y, nt.z));
            sp = fmax_p(sp, rr_min_s);
            sp = fmin_p(sp, 0.95f);
            if (random_float() > sp) return emitted + direct;
            float inv = 1.0f / sp;
            rr_inv = inv;
            att *= inv;
            nt *= inv;
        }
        V3 od = dot(s.wo, rec.normal) >= 0.0f ? rec.normal : -rec.normal;
        Ray scat(offset_origin(rec.point, od), s.wo);
        float npdf = s.is_delta ? 0.0f : s.pdf;
        V3 nabs = s.has_next_absorption ? s.next_absorption : absorption;
        int cat = classify(rec.normal, s.wo, s.is_delta);
        V3 ind;
        AdjState* A = t_adj;
        AdjWeight aw(A, att);
        if (s.transition == 1 || s.transition == 2) {
            IorStack ni = iors;
            if (s.transition == 1) {
                float pi;
                ni.push(m->try_dielectric_ior(rec, pi) ? pi : iors.top());
            } else ni.pop();
            ind = trace_ray(scat, depth - 1, npdf, s.is_delta, ni, nabs, nt, cat, next_caustic(cst, s.is_delta));
        } else {
            ind = trace_ray(scat, depth - 1, npdf, s.is_delta, iors, nabs, nt, cat, next_caustic(cst, s.is_delta));
        }
        aw.restore();
        ind = sanitize(ind);
        if (A && th && s.is_delta) {
            const Disney* dm = static_cast<const Disney*>(m);
            adj_lobe_term(A, *th, dm->last_lobe, rec, view, s.wo, att, ind);
        }
        if (A && th && !s.is_delta) {
            // bounce term: d(att)/dp = d f(view, wo)/dp * |cos| / pdf (/ RR
            // survival), pdf and sampling held fixed
            double scale = (double)std::fabs(dot(rec.normal, s.wo)) / (double)s.pdf;
            double rr = (double)rr_inv;
            double g[5];
            for (int j = 0; j < 5; j++) {
                V3 fp = Scene::material_at(*th, j, kAdjH).evaluate(view, s.wo, rec);
                V3 fm = Scene::material_at(*th, j, -kAdjH).evaluate(view, s.wo, rec);
                double k = scale * rr / (2.0 * (double)kAdjH);
                g[j] = A->aw[0] * ((double)fp.x - fm.x) * k * ind.x + A->aw[1] * ((double)fp.y - fm.y) * k * ind.y +
                       A->aw[2] * ((double)fp.z - fm.z) * k * ind.z;
            }
            A->add(*th, g);
        }
        V3 contrib = att * ind;
        if (depth == max_depth) contrib = clamp_indirect(contrib);
        return emitted + direct + contrib;
    }

    // TraceRay (Renderer.cs:1581-1783)
    V3 trace_ray(const Ray& ray, int depth, float prev_pdf, bool prev_delta, IorStack& iors, V3 absorption, V3 beta,
                 int cat, int cst = kCausticNone) const {
        if (depth <= 0) return V3(0.0f);
        float bm = fmax_p(beta.x, fmax_p(beta.y, beta.z));
        if (bm < 1e-4f) return V3(0.0f);
        Ray cur = ray;
        HitRecord rec;
        bool hit;
        int skips = 0;
        while (true) {
            rec = HitRecord();
            hit = sc.world->hit(cur, kEpsilon, kInfinity, rec);
            if (!hit || cat != kCatCamera || !rec.camera_invisible) break;
            Ray next(offset_origin(rec.point, cur.d), cur.d);
            next.has_diff = cur.has_diff;
            next.dox = cur.dox;
            next.ddx = cur.ddx;
            next.ddy = cur.ddy;
            cur = next;
            if (++skips >= 8) break;
        }
        if (capture_aovs && hit && cat == kCatCamera && probe && probe->pending && !probe->have_depth) {
            probe->depth = rec.t * length(cur.d);
            probe->have_depth = true;
        }
        // analytic filter footprint of a ray with differentials (Renderer.cs:1665-1677)
        if (hit && cur.has_diff) {
            V3 du = length_sq(rec.dpdu) > 0.0f ? rec.dpdu : rec.tangent;
            V3 dv = length_sq(rec.dpdv) > 0.0f ? rec.dpdv : rec.bitangent;
            footprint(cur, rec.point, rec.normal, du, dv, rec.fp);
        }
        if (!sc.medium) {
            V3 r = !hit ? sample_sky(cur, prev_pdf, prev_delta)
                        : shade_surface(cur, rec, depth, prev_pdf, prev_delta, absorption, beta, iors, cst);
            return beer_lambert(r, absorption, hit ? rec.t : std::numeric_limits<float>::infinity());
        }
        float tmax = hit ? rec.t : 1e30f;
        float tmed;
        V3 mb;
        bool scat = medium_sample(tmax, tmed, mb);
        if (scat) {
            V3 p = cur.o + cur.d * tmed;
            if (capture_aovs && probe && probe->pending) {
                if (!probe->have_depth) { probe->depth = tmed * length(cur.d); probe->have_depth = true; }
                probe->commit_as(V3(1.0f), V3(0.0f));
            }
            V3 lnee = direct_lighting_medium(p);
            V3 lind(0.0f);
            V3 mt = beta * mb;
            int used = max_depth - depth;
            float iscale = 1.0f;
            bool kill = false;
            if (used >= rr_min_b) {
                float sp = fmax_p(mt.x, fmax_p(mt.y, mt.z));
                sp = fmax_p(sp, rr_min_s);
                sp = fmin_p(sp, 0.95f);
                if (random_float() > sp) kill = true;
                else iscale = 1.0f / sp;
            }
            if (!kill) {
                float u1 = random_float(), u2 = random_float();
                float z = 1.0f - 2.0f * u1;
                float rr = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
                float phi = 2.0f * kPi * u2;
                V3 wi(rr * std::cos(phi), rr * std::sin(phi), z);
                float ppdf = kInv4Pi, pval = kInv4Pi;
                float pw = ppdf > 1e-20f ? pval / ppdf : 0.0f;
                V3 nt = mt * (pw * iscale);
                lind = pw * iscale * trace_ray(Ray(p, wi), depth - 1, ppdf, false, iors, absorption, nt, kCatCamera,
                                               kCausticDiffuse);
            }
            return beer_lambert(mb * (lnee + lind), absorption, tmed);
        }
        V3 r = !hit ? mb * sample_sky(cur, prev_pdf, prev_delta)
                    : mb * shade_surface(cur, rec, depth, prev_pdf, prev_delta, absorption, beta * mb, iors, cst);
        return beer_lambert(r, absorption, hit ? rec.t : std::numeric_limits<float>::infinity());
    }

    // AcesToneMap (Renderer.cs:893-907)
    static V3 aces(V3 c) {
        c = vmax(c, V3(0.0f));
        V3 a = c * (2.51f * c + V3(0.03f));
        V3 b = c * (2.43f * c + V3(0.59f)) + V3(0.14f);
        V3 m(a.x / b.x, a.y / b.y, a.z / b.z);
        const float ig = 1.0f / 2.2f;
        return V3(std::pow(clampf(m.x, 0, 1), ig), std::pow(clampf(m.y, 0, 1), ig), std::pow(clampf(m.z, 0, 1), ig));
    }
};

// ── Scene construction ───────────────────────────────────────────────────
static std::string g_err;

// Object hierarchy (rt_node): the world list becomes the roots' hittables;
// Group / Transform / Instance / CameraInvisible wrappers per node, shared
// nodes built once (SceneLoader.cs:3259-3337 caches a template and wraps
// it in one Instance per use).  Limits as the device's: wrapper nesting at
// most 4, a Group's list part at most 15 members.
static bool build_hierarchy(const rt_scene_desc* d, Scene& sc, const std::vector<int32_t>& obj_first,
                            std::vector<char>& under_xf) {
    const int64_t nn = d->num_nodes, np = d->num_primitives;
    if (nn <= 0 || !d->nodes || d->num_roots <= 0 || !d->roots || d->num_node_children < 0 ||
        (d->num_node_children > 0 && !d->node_children)) {
        g_err = "bad node tables";
        return false;
    }
    // flat object starting at each primitive (a lone primitive or a mesh)
    std::vector<int32_t> obj_of(np, -1), obj_size(obj_first.size(), 0);
    for (size_t o = 0; o < obj_first.size(); o++) {
        obj_of[obj_first[o]] = (int32_t)o;
        int32_t end = o + 1 < obj_first.size() ? obj_first[o + 1] : (int32_t)np;
        obj_size[o] = end - obj_first[o];
    }
    std::vector<Hittable*> built(nn, nullptr);
    std::vector<int> state(nn, 0), wdepth(nn, 0);
    std::function<Hittable*(int32_t)> make = [&](int32_t n) -> Hittable* {
        if (n < 0 || n >= nn) { g_err = "node index out of range"; return nullptr; }
        if (state[n] == 2) return built[n];
        if (state[n] == 1) { g_err = "node cycle"; return nullptr; }
        state[n] = 1;
        const rt_node& nd = d->nodes[n];
        Hittable* h = nullptr;
        int depth = 0;
        bool wrapper = (nd.flags & RT_NODEF_CAMERA_INVISIBLE) != 0;
        switch (nd.kind) {
            case RT_NODE_PRIMITIVES:
                if (nd.first < 0 || nd.first >= np || obj_of[nd.first] < 0 || obj_size[obj_of[nd.first]] != nd.count) {
                    g_err = "a PRIMITIVES node must name one whole object (a primitive or a mesh)";
                    return nullptr;
                }
                h = sc.objects[obj_of[nd.first]];
                break;
            case RT_NODE_GROUP: {
                if (nd.count <= 0 || nd.first < 0 || (int64_t)nd.first + nd.count > d->num_node_children) {
                    g_err = "bad group child range";
                    return nullptr;
                }
                std::vector<Hittable*> ch;
                for (int32_t k = 0; k < nd.count; k++) {
                    int32_t c = d->node_children[nd.first + k];
                    Hittable* x = make(c);
                    if (!x) return nullptr;
                    depth = std::max(depth, wdepth[c]);
                    ch.push_back(x);
                }
                auto* g = new GroupH(ch);
                sc.owned.push_back(g);
                if (g->list && g->list->objs.size() > 15) { g_err = "group list part over 15 members"; return nullptr; }
                h = g;
                break;
            }
            case RT_NODE_TRANSFORM:
            case RT_NODE_INSTANCE: {
                Hittable* x = make(nd.child);
                if (!x) return nullptr;
                depth = wdepth[nd.child];
                wrapper = true;
                if (nd.kind == RT_NODE_TRANSFORM) h = new TransformH(x, nd.matrix, nd.inverse);
                else {
                    const Material* om = nullptr;
                    if (nd.material >= 0) {
                        if (nd.material >= (int32_t)sc.materials.size()) { g_err = "bad override material"; return nullptr; }
                        om = sc.materials[nd.material].get();
                    }
                    h = new InstanceH(x, om);
                }
                sc.owned.push_back(h);
                break;
            }
            default: g_err = "bad node kind"; return nullptr;
        }
        if (nd.flags & RT_NODEF_CAMERA_INVISIBLE) {
            h = new CameraInvisible(h);
            sc.owned.push_back(h);
        }
        wdepth[n] = depth + (wrapper ? 1 : 0);
        if (wdepth[n] > 4) { g_err = "wrapper nesting over 4"; return nullptr; }
        state[n] = 2;
        built[n] = h;
        return h;
    };
    std::vector<Hittable*> roots;
    for (int64_t r = 0; r < d->num_roots; r++) {
        Hittable* h = make(d->roots[r]);
        if (!h) return false;
        roots.push_back(h);
    }
    // primitives reachable through a TRANSFORM
    std::vector<char> seen(2 * nn, 0);
    std::function<void(int32_t, bool)> mark = [&](int32_t n, bool xf) {
        if (seen[2 * n + xf]) return;
        seen[2 * n + xf] = 1;
        const rt_node& nd = d->nodes[n];
        if (nd.kind == RT_NODE_PRIMITIVES) {
            if (xf)
                for (int32_t k = 0; k < nd.count; k++) under_xf[nd.first + k] = 1;
        } else if (nd.kind == RT_NODE_GROUP) {
            for (int32_t k = 0; k < nd.count; k++) mark(d->node_children[nd.first + k], xf);
        } else mark(nd.child, xf || nd.kind == RT_NODE_TRANSFORM);
    };
    for (int64_t r = 0; r < d->num_roots; r++) mark(d->roots[r], false);
    sc.objects = roots;
    return true;
}

static Scene* build_scene(const rt_scene_desc* d) {
    if (!d || (d->num_primitives > 0 && !d->primitives) || (d->num_materials > 0 && !d->materials) ||
        (d->num_lights > 0 && !d->lights)) {
        g_err = "null table";
        return nullptr;
    }
    auto sc = std::make_unique<Scene>();
    if (d->num_images < 0 || (d->num_images > 0 && !d->images)) { g_err = "bad image table"; return nullptr; }
    for (int64_t k = 0; k < d->num_images; k++) {
        const rt_image& im = d->images[k];
        if (!im.rgba8 || im.width <= 0 || im.height <= 0) { g_err = "bad image"; return nullptr; }
        sc->images.emplace_back(new ImageTex(im));
    }
    bool any_image = false;
    for (int64_t i = 0; i < d->num_materials; i++) {
        const rt_material& rm = d->materials[i];
        Material* m = make_material(rm);
        if (!m) { g_err = "bad material kind"; return nullptr; }
        if (rm.flags & RT_MATF_IMAGE) {
            if ((rm.kind != RT_MAT_LAMBERTIAN && rm.kind != RT_MAT_DISNEY) || rm.image < 0 || rm.image >= d->num_images ||
                (rm.kind == RT_MAT_DISNEY && (rm.spec_trans != 0.0f || (rm.flags & RT_MATF_TEXTURED)))) {
                delete m;
                g_err = "unsupported image-textured material";
                return nullptr;
            }
            m->image = rm.image;
            m->image_raw = rm;
            any_image = true;
        }
        sc->materials.emplace_back(m);
    }
    int64_t np = d->num_primitives;
    sc->prim_objs.assign(np, nullptr);
    std::vector<int32_t> obj_first;      // first primitive of each flat object
    Mesh* cur_mesh = nullptr;
    int cur_mesh_id = -1;
    for (int64_t i = 0; i < np; i++) {
        const rt_primitive& p = d->primitives[i];
        if (p.material < 0 || p.material >= d->num_materials) { g_err = "bad material index"; return nullptr; }
        const Material* m = sc->materials[p.material].get();
        const float* v = p.v;
        Hittable* h = nullptr;
        switch (p.kind) {
            case RT_PRIM_TRIANGLE: {
                const float* nv = d->normals ? d->normals + 9 * i : nullptr;
                bool smooth = false;
                for (int k = 0; nv && k < 9; k++) smooth = smooth || nv[k] != 0.0f;
                if (smooth) {
                    static const float kBary[6] = {0, 0, 1, 0, 0, 1};
                    const float* uv = d->texcoords ? d->texcoords + 6 * i : kBary;
                    h = new SmoothTriangle(V3(v[0], v[1], v[2]), V3(v[3], v[4], v[5]), V3(v[6], v[7], v[8]),
                                           V3(nv[0], nv[1], nv[2]), V3(nv[3], nv[4], nv[5]), V3(nv[6], nv[7], nv[8]),
                                           uv, m, (int)i);
                } else {
                    h = new Triangle(V3(v[0], v[1], v[2]), V3(v[3], v[4], v[5]), V3(v[6], v[7], v[8]), m, (int)i);
                }
                break;
            }
            case RT_PRIM_SPHERE: h = new Sphere(V3(v[0], v[1], v[2]), v[3], m, (int)i); break;
            case RT_PRIM_QUAD:
                h = new Quad(V3(v[0], v[1], v[2]), V3(v[3], v[4], v[5]), V3(v[6], v[7], v[8]), m, (int)i,
                             (p.flags & RT_PRIMF_BACKFACE_CULLED) != 0);
                break;
            case RT_PRIM_PLANE: h = new Plane(V3(v[0], v[1], v[2]), V3(v[3], v[4], v[5]), m, (int)i); break;
            default: g_err = "bad primitive kind"; return nullptr;
        }
        sc->owned.push_back(h);
        sc->prim_objs[i] = h;
        Hittable* obj = h;
        if (p.flags & RT_PRIMF_CAMERA_INVISIBLE) { obj = new CameraInvisible(h); sc->owned.push_back(obj); }
        if (p.mesh >= 0) {
            if (p.kind != RT_PRIM_TRIANGLE) { g_err = "mesh faces must be triangles"; return nullptr; }
            if (p.mesh != cur_mesh_id) {
                cur_mesh = new Mesh();
                cur_mesh_id = p.mesh;
                sc->owned.push_back(cur_mesh);
                sc->objects.push_back(cur_mesh);
                obj_first.push_back((int32_t)i);
            }
            cur_mesh->tris.push_back(obj);
        } else {
            cur_mesh = nullptr;
            cur_mesh_id = -1;
            sc->objects.push_back(obj);
            obj_first.push_back((int32_t)i);
        }
    }
    sc->tri_flag.assign(np, 0);
    for (int64_t i = 0; i < np; i++) sc->tri_flag[i] = d->primitives[i].kind == RT_PRIM_TRIANGLE;
    for (int64_t i = 0; i < d->num_materials; i++) {
        const rt_material& m = d->materials[i];
        if (!(m.flags & RT_MATF_TEXTURED)) continue;
        if (m.kind != RT_MAT_DISNEY || m.spec_trans != 0.0f || m.diff_trans != 0.0f) {
            g_err = "unsupported textured material";
            return nullptr;
        }
        if (!d->texels || d->tex_width <= 0 || d->tex_height <= 0) { g_err = "textured material without texture set"; return nullptr; }
        sc->textured = true;
    }
    if (sc->textured) {
        sc->tex_w = d->tex_width;
        sc->tex_h = d->tex_height;
        sc->tex.assign(d->texels, d->texels + (size_t)5 * d->tex_width * d->tex_height);
        if (d->texcoords) sc->tuv.assign(d->texcoords, d->texcoords + 6 * (size_t)np);
    }
    if (any_image && d->texcoords && sc->tuv.empty()) sc->tuv.assign(d->texcoords, d->texcoords + 6 * (size_t)np);
    for (Hittable* o : sc->objects)
        if (Mesh* me = dynamic_cast<Mesh*>(o)) me->finish();
    if (d->num_nodes > 0 || d->num_roots > 0) {
        std::vector<char> under_xf(np, 0);
        if (!build_hierarchy(d, *sc, obj_first, under_xf)) return nullptr;
        if (any_image) { g_err = "image textures in a scene with an object hierarchy"; return nullptr; }
        for (int64_t i = 0; i < d->num_lights; i++)
            if (d->lights[i].kind == RT_LIGHT_GEOMETRY && d->lights[i].primitive >= 0 &&
                d->lights[i].primitive < np && under_xf[d->lights[i].primitive]) {
                g_err = "geometry light on a transformed primitive";
                return nullptr;
            }
    }
    // World assembly (SceneLoader.cs:460-487)
    sc->world = new HittableList();
    if (sc->objects.size() > 4) {
        std::vector<Hittable*> fin, inf;
        for (Hittable* o : sc->objects) (o->is_plane() ? inf : fin).push_back(o);
        if (!fin.empty()) {
            sc->bvh = build_bvh(fin);
            sc->owned.push_back(sc->bvh);
            sc->world->objs.push_back(sc->bvh);
        }
        for (Hittable* o : inf) sc->world->objs.push_back(o);
    } else sc->world->objs = sc->objects;

    sc->camera = Camera(d->camera);
    sc->sky = V3(d->sky_color[0], d->sky_color[1], d->sky_color[2]);
    if (d->global_medium.enabled) {
        sc->medium = true;
        const rt_medium& gm = d->global_medium;
        sc->sig_a = vmax(V3(gm.sigma_a[0], gm.sigma_a[1], gm.sigma_a[2]), V3(0.0f));
        sc->sig_s = vmax(V3(gm.sigma_s[0], gm.sigma_s[1], gm.sigma_s[2]), V3(0.0f));
        sc->sig_t = sc->sig_a + sc->sig_s;
    }
    // Lights
    for (int64_t i = 0; i < d->num_lights; i++) {
        const rt_light& l = d->lights[i];
        V3 col(l.color[0], l.color[1], l.color[2]);
        V3 pos(l.position[0], l.position[1], l.position[2]);
        const Material* px = (l.material >= 0 && l.material < d->num_materials) ? sc->materials[l.material].get() : nullptr;
        Light* L = nullptr;
        switch (l.kind) {
            case RT_LIGHT_SPHERE: L = new SphereLight(pos, l.radius, col, l.intensity, l.shadow_samples, px); break;
            case RT_LIGHT_AREA:
                L = new AreaLight(pos, V3(l.u[0], l.u[1], l.u[2]), V3(l.v[0], l.v[1], l.v[2]), col, l.intensity, l.shadow_samples,
                                  l.soft_radius, px);
                break;
            case RT_LIGHT_DIRECTIONAL: L = new DirectionalLight(pos, col, l.intensity, l.radius, l.shadow_samples); break;
            case RT_LIGHT_ENVIRONMENT: {
                auto* E = new EnvironmentLight(sc->sky, l.shadow_samples);
                if (!sc->env_light) sc->env_light = E;   // lights.OfType<EnvironmentLight>().FirstOrDefault()
                L = E;
                break;
            }
            case RT_LIGHT_GEOMETRY: {
                if (l.primitive < 0 || l.primitive >= np) { g_err = "geometry light primitive"; return nullptr; }
                const Hittable* g = sc->prim_objs[l.primitive];
                const Material* gm = sc->materials[d->primitives[l.primitive].material].get();
                const Emissive* em = dynamic_cast<const Emissive*>(gm);
                if (!em || d->primitives[l.primitive].kind == RT_PRIM_PLANE) { g_err = "geometry light needs an emissive tri/quad/sphere"; return nullptr; }
                if (dynamic_cast<const SmoothTriangle*>(g)) { g_err = "smooth triangle geometry light"; return nullptr; }
                L = new GeometryLight(g, em, l.shadow_samples, l.soft_radius);
                L->proxy = nullptr;
                break;
            }
            default: g_err = "bad light kind"; return nullptr;
        }
        sc->lights.emplace_back(L);
        sc->light_ptrs.push_back(L);
    }
    // Emitter → light map (Renderer.cs:409-426): geometry lights first, then proxies
    for (size_t i = 0; i < sc->lights.size(); i++) {
        const rt_light& l = d->lights[i];
        if (l.kind == RT_LIGHT_GEOMETRY)
            sc->emitter_to_light.push_back({sc->materials[d->primitives[l.primitive].material].get(), sc->lights[i].get()});
    }
    for (size_t i = 0; i < sc->lights.size(); i++)
        if (sc->lights[i]->proxy) sc->emitter_to_light.push_back({sc->lights[i]->proxy, sc->lights[i].get()});

    // ComputeFiniteSceneBounds + classifier (Renderer.cs:361-378,476-502)
    AABB acc = AABB::empty();
    bool any = false;
    for (Hittable* o : sc->world->objs) {
        AABB b = o->box();
        V3 dd = b.mx - b.mn;
        if (dd.x < 1e5f && dd.y < 1e5f && dd.z < 1e5f) { acc = any ? surround(acc, b) : b; any = true; }
    }
    if (!any) {
        AABB raw = sc->world->box();
        acc = AABB(vmax(raw.mn, V3(-1e3f)), vmin(raw.mx, V3(1e3f)));
    }
    sc->bounds = acc;
    float flux = 0.0f;
    for (Light* l : sc->light_ptrs) flux += l->approximate_power(acc);
    V3 ext = acc.mx - acc.mn;
    float R = std::fmax(0.5f * length(ext), 1e-3f);
    float area = 4.0f * kPi * R * R;
    sc->total_flux = flux;
    sc->scene_radius = R;
    sc->mean_irr = flux / area;
    bool ind = sc->mean_irr < 0.5f;
    sc->rr_min_bounces = ind ? 6 : 3;
    sc->rr_min_survival = ind ? 0.40f : 0.10f;
    return sc.release();
}

}  // namespace orc

using namespace orc;

// ═════════════════════════════════════════════════════════════════════════
// C exports for the ctypes test harness
// ═════════════════════════════════════════════════════════════════════════
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

float orc_sobol(uint32_t index, uint32_t dim, uint32_t seed) { return sobol_sample(index, dim, seed); }

void orc_sobol_batch(const uint32_t* index, const uint32_t* dim, const uint32_t* seed, int64_t n, float* out) {
    for (int64_t i = 0; i < n; i++) out[i] = sobol_sample(index[i], dim[i], seed[i]);
}

void orc_prng_batch(const uint32_t* index, const uint32_t* dim, const uint32_t* seed, int64_t n, float* out) {
    for (int64_t i = 0; i < n; i++) out[i] = prng_sample(index[i], dim[i], seed[i]);
}

void* orc_scene_create(const rt_scene_desc* d) { return build_scene(d); }

// ImageTexture probes: the mip chain (levels concatenated, RGB; returns the
// level count, sizes in wh[2 * level]) and n lookups at (u, v) with
// footprints fp (4 per lookup, zero = none).
int orc_image_levels(const rt_image* im, float* out, int64_t cap, int32_t* wh) {
    ImageTex t(*im);
    int64_t off = 0;
    for (int m = 0; m < t.levels(); m++) {
        wh[2 * m] = t.mw[m];
        wh[2 * m + 1] = t.mh[m];
        for (float x : t.mip[m]) {
            if (off < cap) out[off] = x;
            off++;
        }
    }
    return t.levels();
}
void orc_image_sample(const rt_image* im, int64_t n, const float* uv, const float* fp, float* out) {
    ImageTex t(*im);
    for (int64_t i = 0; i < n; i++) {
        V3 c = t.value(uv[2 * i], uv[2 * i + 1], fp + 4 * i);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
}
void orc_scene_destroy(void* s) { delete (Scene*)s; }

int orc_scene_info(void* sp, rt_scene_info* info) {
    Scene* s = (Scene*)sp;
    std::memset(info, 0, sizeof(*info));
    int64_t nodes = 0, leaves = 0;
    int maxd = 0;
    auto walk = [&](Hittable* h) {
        if (BvhNode* b = dynamic_cast<BvhNode*>(h)) b->stats(0, nodes, leaves, maxd);
    };
    if (s->bvh) walk(s->bvh);
    for (Hittable* o : s->objects)
        if (Mesh* m = dynamic_cast<Mesh*>(o)) walk(m->accel);
    info->num_primitives = (int64_t)s->prim_objs.size();
    info->num_objects = (int64_t)s->objects.size();
    info->num_nodes = nodes;
    info->num_leaves = leaves;
    info->max_depth = maxd;
    info->has_bvh = s->bvh != nullptr;
    info->rr_min_bounces = s->rr_min_bounces;
    info->rr_min_survival = s->rr_min_survival;
    info->scene_radius = s->scene_radius;
    info->mean_irradiance = s->mean_irr;
    return 0;
}

// Pre-order dump of the top-level BVH: internal node → (0, axis, 6 box
// floats as bits); leaf → (1, count, 6 box floats, entries...) where an
// entry is the primitive id or -(mesh_ordinal+2) for a nested mesh.
int64_t orc_bvh_dump(void* sp, int32_t* out, int64_t cap) {
    Scene* s = (Scene*)sp;
    if (!s->bvh) return 0;
    std::vector<int32_t> v;
    std::vector<const Hittable*> meshes;
    for (Hittable* o : s->objects)
        if (dynamic_cast<Mesh*>(o)) meshes.push_back(o);
    auto box = [&](const AABB& b) {
        float f[6] = {b.mn.x, b.mn.y, b.mn.z, b.mx.x, b.mx.y, b.mx.z};
        for (float x : f) { int32_t i; std::memcpy(&i, &x, 4); v.push_back(i); }
    };
    std::function<void(const BvhNode*)> rec = [&](const BvhNode* n) {
        if (!n->prims.empty()) {
            v.push_back(1);
            v.push_back((int32_t)n->prims.size());
            box(n->bbox);
            for (const Hittable* p : n->prims) {
                const Hittable* q = p;
                if (auto* ci = dynamic_cast<const CameraInvisible*>(q)) q = ci->inner;
                int32_t id = -1;
                if (auto* t = dynamic_cast<const Triangle*>(q)) id = t->prim;
                else if (auto* sp2 = dynamic_cast<const Sphere*>(q)) id = sp2->prim;
                else if (auto* qu = dynamic_cast<const Quad*>(q)) id = qu->prim;
                else if (auto* pl = dynamic_cast<const Plane*>(q)) id = pl->prim;
                else {
                    for (size_t k = 0; k < meshes.size(); k++)
                        if (meshes[k] == q) id = -(int32_t)k - 2;
                }
                v.push_back(id);
            }
            return;
        }
        v.push_back(0);
        v.push_back(n->axis);
        box(n->bbox);
        rec(n->left);
        rec(n->right);
    };
    rec((const BvhNode*)s->bvh);
    int64_t n = (int64_t)v.size();
    if (out && cap >= n) std::memcpy(out, v.data(), (size_t)n * 4);
    return n;
}

}  // extern "C"

namespace orc {

// q with q * q == n, or 0 when n is not a positive square.
static int exact_isqrt(int n) {
    if (n <= 0) return 0;
    int q = (int)std::lround(std::sqrt((double)n));
    return (int64_t)q * q == n ? q : 0;
}

static int check_settings(const rt_settings* st, bool prng_ok = false) {
    if (!st || st->width <= 0 || st->height <= 0 || st->samples_per_pixel <= 0 || st->max_depth < 0) {
        g_err = "invalid settings";
        return RT_ERR_INVALID_ARGUMENT;
    }
    if (st->sampler != RT_SAMPLER_SOBOL && !(prng_ok && st->sampler == RT_SAMPLER_PRNG)) {
        g_err = "sampler not supported here";
        return RT_ERR_UNSUPPORTED;
    }
    if (st->sampler == RT_SAMPLER_PRNG && (st->sample_offset != 0 || st->sample_total != 0) &&
        exact_isqrt(st->sample_total) == 0) {
        g_err = "a chunk of a PRNG render needs sample_total = the whole render's q x q strata";
        return RT_ERR_INVALID_ARGUMENT;
    }
    if (st->caustics && st->caustic_photons < 0) { g_err = "negative photon budget"; return RT_ERR_INVALID_ARGUMENT; }
    return 0;
}

static inline uint32_t pixel_seed(int i, int j, uint32_t x) {
    return (((uint32_t)i * 73856093u) ^ ((uint32_t)j * 19349663u)) ^ x;
}

}  // namespace orc

extern "C" {

// world.Hit(ray, 1e-4, float.MaxValue) for a batch of rays; optional
// per-ray node-test / primitive-test counters (reference traversal order).
int orc_trace_rays(void* sp, const float* rays, int64_t n, int32_t* prim, float* t, float* u, float* v,
                   int64_t* node_tests, int64_t* prim_tests) {
    Scene* s = (Scene*)sp;
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        Ray ray(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]));
        HitRecord rec;
        t_node_tests = 0;
        t_prim_tests = 0;
        bool h = s->world->hit(ray, kEpsilon, kInfinity, rec);
        if (prim) prim[i] = h ? rec.prim : -1;
        if (t) t[i] = h ? rec.t : 0.0f;
        if (u) u[i] = h ? rec.u : 0.0f;
        if (v) v[i] = h ? rec.v : 0.0f;
        if (node_tests) node_tests[i] = t_node_tests;
        if (prim_tests) prim_tests[i] = t_prim_tests;
    }
    return 0;
}

// Closest-hit records of world-space rays (the fields MapHitToWorld maps):
// f[8 i ..] = t, point, normal, front; m[2 i ..] = primitive, material index
// (-1, -1 on a miss).
int orc_trace_records(void* sp, const float* rays, int64_t n, float* f, int32_t* m) {
    Scene* s = (Scene*)sp;
    std::unordered_map<const Material*, int32_t> mat_index;
    for (size_t k = 0; k < s->materials.size(); k++) mat_index[s->materials[k].get()] = (int32_t)k;
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        Ray ray(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]));
        HitRecord rec;
        bool h = s->world->hit(ray, kEpsilon, kInfinity, rec);
        float* o = f + 8 * i;
        if (!h) {
            for (int k = 0; k < 8; k++) o[k] = 0.0f;
            m[2 * i] = m[2 * i + 1] = -1;
            continue;
        }
        o[0] = rec.t;
        o[1] = rec.point.x; o[2] = rec.point.y; o[3] = rec.point.z;
        o[4] = rec.normal.x; o[5] = rec.normal.y; o[6] = rec.normal.z;
        o[7] = rec.front_face ? 1.0f : 0.0f;
        m[2 * i] = rec.prim;
        auto it = mat_index.find(rec.mat);
        m[2 * i + 1] = it == mat_index.end() ? -1 : it->second;
    }
    return 0;
}

// Camera rays of Renderer.cs:630-653 + the camera-visibility skip loop of
// TraceRay (Renderer.cs:1609-1644).  Output index s*W*H + y*W + x.
int orc_trace_primary(void* sp, const rt_settings* st, int32_t sbegin, int32_t scount, int32_t* prim, float* t,
                      float* u, float* v, int64_t* node_tests, int64_t* prim_tests) {
    Scene* s = (Scene*)sp;
    int rc = check_settings(st);
    if (rc) return rc;
    int W = st->width, H = st->height;
    int64_t np = (int64_t)W * H;
    SamplerCtx& c = t_smp;
    int64_t nt = 0, pt = 0;
    for (int j = 0; j < H; j++)
        for (int i = 0; i < W; i++) {
            uint32_t ps = pixel_seed(i, j, st->seed_xor);
            for (int k = 0; k < scount; k++) {
                c.active = true; c.seed = ps; c.index = (uint32_t)(sbegin + k); c.dim = 0; c.fixed = nullptr;
                float ju = random_float();
                float jv = random_float();
                float uu = ((float)i + ju) / (float)W;
                float vv = ((float)(H - j - 1) + jv) / (float)H;
                Ray ray = s->camera.get_ray(uu, vv, 0.0f, 0.0f, st->texture_filtering != 2);
                HitRecord rec;
                bool hit;
                int skips = 0;
                t_node_tests = 0;
                t_prim_tests = 0;
                while (true) {
                    rec = HitRecord();
                    hit = s->world->hit(ray, kEpsilon, kInfinity, rec);
                    if (!hit || !rec.camera_invisible) break;
                    ray = Ray(offset_origin(rec.point, ray.d), ray.d);
                    if (++skips >= 8) break;
                }
                nt += t_node_tests;
                pt += t_prim_tests;
                int64_t o = (int64_t)k * np + (int64_t)j * W + i;
                if (prim) prim[o] = hit ? rec.prim : -1;
                if (t) t[o] = hit ? rec.t : 0.0f;
                if (u) u[o] = hit ? rec.u : 0.0f;
                if (v) v[o] = hit ? rec.v : 0.0f;
            }
            c.active = false;
        }
    if (node_tests) *node_tests = nt;
    if (prim_tests) *prim_tests = pt;
    return 0;
}

// Renderer.Render (Renderer.cs:522-733), Sobol mode, 16×16 tiles over
// `nthreads` workers.  row_step > 1 renders only rows j % row_step == 0
// (bounded CPU-baseline samples); other rows are left untouched.
int orc_render(void* sp, const rt_settings* st, const rt_film* film, int32_t nthreads, int32_t row_step,
               rt_render_stats* stats) {
    Scene* s = (Scene*)sp;
    int rc = check_settings(st, true);
    if (rc) return rc;
    if (row_step < 1) row_step = 1;
    // photon pre-pass (Renderer.cs:385-406): gather-radius cap 6 % of the
    // scene radius (or the settings override), grid cell half of it
    PhotonMapO pm;
    bool caus = false;
    float gr = 0.0f;
    if (st->caustics && st->caustic_photons > 0 && !s->light_ptrs.empty() && s->textured) {
        g_err = "photon caustics with textured materials are not supported";
        return RT_ERR_UNSUPPORTED;
    }
    if (st->caustics && st->caustic_photons > 0 && !s->light_ptrs.empty()) {
        gr = st->caustic_gather_radius > 0.0f ? st->caustic_gather_radius : std::fmax(s->scene_radius * 0.06f, 1e-4f);
        caus = build_photon_map(*s, st->caustic_photons, gr * 0.5f, std::max(1, (int)nthreads), pm);
    }
    Renderer R(*s, *st);
    // PRNG mode: ceil(sqrt(spp))^2 stratified samples (Renderer.cs:556-558);
    // a chunk takes exactly its spp samples of the sample_total = q^2 grid
    const bool prng = st->sampler == RT_SAMPLER_PRNG;
    const bool chunk = st->sample_offset != 0 || st->sample_total != 0;
    const int sq = prng && chunk ? exact_isqrt(st->sample_total) : (int)std::ceil(std::sqrt((float)st->samples_per_pixel));
    const int W = st->width, H = st->height, spp = prng && !chunk ? sq * sq : st->samples_per_pixel;
    const float inv_sq = 1.0f / (float)sq;
    const int64_t np = (int64_t)W * H;
    const int T = 16;
    int tx = (W + T - 1) / T, ty = (H + T - 1) / T;
    int total = tx * ty;
    std::atomic<int> next(0);
    std::atomic<int64_t> seg_closest(0), seg_shadow(0);
    // chunk of a longer render: global sample indices [soff, soff + spp), the
    // footprint of the whole render's spp; even global indices form half A
    const int soff = st->sample_offset, spp_fp = st->sample_total > 0 ? st->sample_total : spp;
    const int nA = (soff & 1) ? spp / 2 : (spp + 1) / 2, nB = spp - nA;
    auto worker = [&]() {
        AovProbe probe;
        Renderer RR(*s, *st);
        RR.probe = &probe;
        RR.pmap = caus ? &pm : nullptr;
        RR.caustics = caus;
        RR.gather_r = gr;
        t_block_spec_trans = caus;
        t_shadow_segments = 0;
        for (;;) {
            int tile = next.fetch_add(1);
            if (tile >= total) break;
            int x0 = (tile % tx) * T, y0 = (tile / tx) * T;
            int x1 = std::min(x0 + T, W), y1 = std::min(y0 + T, H);
            // pixel deltas of the differential rays, shrunk by ceil(sqrt(spp)) (Renderer.cs:556,616-624)
            const float pd_shrink = 1.0f / (float)(int)std::ceil(std::sqrt((float)spp_fp));
            const float pd_s = pd_shrink / (float)W, pd_t = pd_shrink / (float)H;
            for (int j = y0; j < y1; j++) {
                if (j % row_step != 0) continue;
                for (int i = x0; i < x1; i++) {
                    V3 cum(0.0f), sa(0.0f), sb(0.0f), aa(0.0f), ab(0.0f), na(0.0f), nb(0.0f);
                    float da = 0, db = 0;
                    int dca = 0, dcb = 0;
                    uint32_t ps = pixel_seed(i, j, st->seed_xor);
                    for (int smp = 0; smp < spp; smp++) {
                        SamplerCtx& c = t_smp;
                        c.active = true; c.seed = ps; c.index = (uint32_t)(soff + smp); c.dim = 0; c.fixed = nullptr;
                        c.prng = prng;
                        float ju, jv;
                        if (prng) {   // stratum (sx, sy) of the sq x sq grid (Renderer.cs:677-682)
                            const int gs = soff + smp;   // global sample index
                            ju = ((float)(gs % sq) + random_float()) * inv_sq;
                            jv = ((float)(gs / sq) + random_float()) * inv_sq;
                        } else {
                            ju = random_float();
                            jv = random_float();
                        }
                        float uu = ((float)i + ju) / (float)W;
                        float vv = ((float)(H - j - 1) + jv) / (float)H;
                        Ray ray = s->camera.get_ray(uu, vv, pd_s, pd_t, st->texture_filtering != 2);
                        IorStack iors;
                        probe = AovProbe();
                        V3 L = RR.trace_ray(ray, RR.max_depth, 0.0f, true, iors, V3(0.0f), V3(1.0f), kCatCamera);
                        c.active = false;
                        c.prng = false;
                        L = RR.clamp_radiance(L);
                        cum += L;
                        bool isA = ((soff + smp) & 1) == 0;
                        if (isA) sa += L; else sb += L;
                        if (RR.capture_aovs) {
                            if (probe.pending) probe.commit();
                            if (isA) { aa += probe.albedo; na += probe.normal; if (probe.have_depth) { da += probe.depth; dca++; } }
                            else { ab += probe.albedo; nb += probe.normal; if (probe.have_depth) { db += probe.depth; dcb++; } }
                        }
                    }
                    V3 lin = cum / (float)spp;
                    int64_t px = (int64_t)j * W + i;
                    auto put3 = [&](float* p, V3 c) { if (p) { p[px] = c.x; p[np + px] = c.y; p[2 * np + px] = c.z; } };
                    if (film) {
                        put3(film->beauty, lin);
                        // a half without samples repeats the other
                        V3 mA = nA > 0 ? sa / (float)nA : sb / (float)nB;
                        V3 mB = nB > 0 ? sb / (float)nB : mA;
                        put3(film->beauty_a, mA);
                        put3(film->beauty_b, mB);
                        if (film->sample_count) film->sample_count[px] = spp;
                        put3(film->display, Renderer::aces(lin * RR.exposure));
                        if (RR.capture_aovs) {
                            V3 alA = nA > 0 ? aa / (float)nA : ab / (float)nB, alB = nB > 0 ? ab / (float)nB : alA;
                            V3 nrA = nA > 0 ? na / (float)nA : nb / (float)nB, nrB = nB > 0 ? nb / (float)nB : nrA;
                            float dA = dca > 0 ? da / (float)dca : -1.0f;
                            float dB = nB > 0 ? (dcb > 0 ? db / (float)dcb : -1.0f) : dA;
                            if (nA == 0) dA = dB;
                            put3(film->albedo_a, alA);
                            put3(film->albedo_b, alB);
                            put3(film->normal_a, nrA);
                            put3(film->normal_b, nrB);
                            if (film->depth_a) film->depth_a[px] = dA;
                            if (film->depth_b) film->depth_b[px] = dB;
                        }
                    }
                }
            }
        }
        seg_shadow += t_shadow_segments;
    };
    if (nthreads <= 1) worker();
    else {
        std::vector<std::thread> th;
        for (int k = 0; k < nthreads; k++) th.emplace_back(worker);
        for (auto& x : th) x.join();
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        int64_t rows = (H + row_step - 1) / row_step;
        stats->camera_samples = rows * (int64_t)W * spp;
        stats->rays_shadow = seg_shadow.load();
    }
    (void)seg_closest;
    return 0;
}

// HomogeneousMedium.Sample with explicit coefficients and random numbers
// rnd[0] (channel pick), rnd[1] (distance): t, beta[3], scattered.
int orc_medium_sample(const float* sig_t, const float* sig_s, float tmax, const float* rnd, float* t, float* beta,
                      int32_t* scat) {
    SamplerCtx& c = t_smp;
    c.fixed = rnd; c.nfixed = 2; c.fixed_pos = 0;
    V3 b;
    bool sc = Renderer::medium_sample_of(V3(sig_t[0], sig_t[1], sig_t[2]), V3(sig_s[0], sig_s[1], sig_s[2]), tmax, *t, b);
    c.fixed = nullptr;
    beta[0] = b.x; beta[1] = b.y; beta[2] = b.z;
    *scat = sc ? 1 : 0;
    return 0;
}

// GatherCaustics at n points of a surface with normal nrm[k] and view
// direction view[k], Lambertian BRDF of albedo `albedo`, gather-radius cap
// max_r (the photon map from orc_photon_map_from / orc_photon_build).
int orc_caustic_estimate(void* h, const float* pts, const float* nrm, const float* view, int64_t n, const float* albedo,
                         float max_r, float* out) {
    const PhotonMapO* map = (const PhotonMapO*)h;
    Lambertian lam(V3(albedo[0], albedo[1], albedo[2]));
    for (int64_t k = 0; k < n; k++) {
        HitRecord rec;
        rec.point = V3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
        rec.normal = V3(nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2]);
        rec.front_face = true;
        rec.mat = &lam;
        V3 e = Renderer::caustic_estimate(*map, rec, V3(view[3 * k], view[3 * k + 1], view[3 * k + 2]), &lam, max_r);
        out[3 * k] = e.x; out[3 * k + 1] = e.y; out[3 * k + 2] = e.z;
    }
    return 0;
}

// Filtered texture set at n texture coordinates (uv pairs) -> out[n][5].
int orc_texture_sample(void* sp, const float* uv, int64_t n, float* out) {
    Scene* s = (Scene*)sp;
    if (!s->textured) return RT_ERR_INVALID_ARGUMENT;
    for (int64_t k = 0; k < n; k++) {
        int idx[4];
        float w[4];
        s->tex_foot(uv[2 * k], uv[2 * k + 1], idx, w);
        for (int c = 0; c < 5; c++) out[5 * k + c] = s->tex_at(idx, w, c);
    }
    return 0;
}

// FP64 CPU adjoint: d_tex[5][tex_h][tex_w] = sum over pixels and samples of
// dLoss/dBeauty / spp propagated through the render recursion (see AdjState).
// Needs a textured scene without medium or transmissive materials and
// inactive clamps (max_sample_radiance * indirect_clamp_factor >= 1e30).
int orc_adjoint(void* sp, const rt_settings* st, const float* d_beauty, double* d_tex, int32_t nthreads) {
    Scene* s = (Scene*)sp;
    int rc = check_settings(st);
    if (rc) return rc;
    if (!s->textured) { g_err = "scene has no textured material"; return RT_ERR_INVALID_ARGUMENT; }
    if (s->medium || st->caustics) { g_err = "adjoint: no medium or caustics"; return RT_ERR_UNSUPPORTED; }
    if ((double)st->max_sample_radiance * (double)st->indirect_clamp_factor < 1e30) {
        g_err = "adjoint: radiance clamps must be inactive";
        return RT_ERR_UNSUPPORTED;
    }
    const int W = st->width, H = st->height, spp = st->samples_per_pixel;
    const int64_t np = (int64_t)W * H, plane = (int64_t)s->tex_w * s->tex_h;
    std::atomic<int> next(0);
    std::mutex mu;
    std::fill(d_tex, d_tex + 5 * plane, 0.0);
    auto worker = [&]() {
        std::vector<double> grad(5 * plane, 0.0);
        AdjState A;
        A.grad = grad.data();
        A.plane = plane;
        Renderer RR(*s, *st);
        const float pd_shrink = 1.0f / (float)(int)std::ceil(std::sqrt((float)spp));
        const float pd_s = pd_shrink / (float)W, pd_t = pd_shrink / (float)H;
        for (;;) {
            int j = next.fetch_add(1);
            if (j >= H) break;
            for (int i = 0; i < W; i++) {
                int64_t px = (int64_t)j * W + i;
                double db[3] = {d_beauty[px], d_beauty[np + px], d_beauty[2 * np + px]};
                if (db[0] == 0.0 && db[1] == 0.0 && db[2] == 0.0) continue;
                uint32_t ps = pixel_seed(i, j, st->seed_xor);
                for (int smp = 0; smp < spp; smp++) {
                    SamplerCtx& c = t_smp;
                    c.active = true; c.seed = ps; c.index = (uint32_t)smp; c.dim = 0; c.fixed = nullptr;
                    float ju = random_float();
                    float jv = random_float();
                    float uu = ((float)i + ju) / (float)W;
                    float vv = ((float)(H - j - 1) + jv) / (float)H;
                    Ray ray = s->camera.get_ray(uu, vv, pd_s, pd_t, st->texture_filtering != 2);
                    IorStack iors;
                    for (int k = 0; k < 3; k++) A.aw[k] = db[k] / (double)spp;
                    t_adj = &A;
                    RR.trace_ray(ray, RR.max_depth, 0.0f, true, iors, V3(0.0f), V3(1.0f), kCatCamera);
                    t_adj = nullptr;
                    c.active = false;
                }
            }
        }
        std::lock_guard<std::mutex> g(mu);
        for (int64_t k = 0; k < 5 * plane; k++) d_tex[k] += grad[k];
    };
    if (nthreads <= 1) worker();
    else {
        std::vector<std::thread> th;
        for (int k = 0; k < nthreads; k++) th.emplace_back(worker);
        for (auto& x : th) x.join();
    }
    return 0;
}

// ── Photon map probes ──────────────────────────────────────────────────────
// Builds the caustic photon map of a scene (CausticPhotonTracer.Build +
// PhotonMap ctor).  Returns a handle (null when no photon was deposited);
// *n_photons / *n_emitted / *n_buckets describe it.
void* orc_photon_build(void* sp, int32_t budget, float cell, int32_t nthreads, int64_t* n_photons, int64_t* n_emitted,
                       int64_t* n_buckets) {
    Scene* s = (Scene*)sp;
    PhotonMapO* m = new PhotonMapO();
    int64_t em = 0;
    bool ok = build_photon_map(*s, budget, cell, std::max(1, (int)nthreads), *m, &em);
    if (n_emitted) *n_emitted = em;
    if (!ok) {
        delete m;
        if (n_photons) *n_photons = 0;
        if (n_buckets) *n_buckets = 0;
        return nullptr;
    }
    if (n_photons) *n_photons = (int64_t)m->ph.size();
    if (n_buckets) *n_buckets = (int64_t)m->mask + 1;
    return m;
}

// Copies the map out: pos/inc/pow 3 floats per photon, key (bucket) and
// idx per photon, start (n_buckets + 1), grid = origin xyz, cell.
void orc_photon_get(void* h, float* pos, float* inc, float* pw, int32_t* key, int32_t* idx, int32_t* start,
                    float* grid) {
    PhotonMapO* m = (PhotonMapO*)h;
    for (size_t i = 0; i < m->ph.size(); i++) {
        const PhotonO& p = m->ph[i];
        if (pos) { pos[3 * i] = p.pos.x; pos[3 * i + 1] = p.pos.y; pos[3 * i + 2] = p.pos.z; }
        if (inc) { inc[3 * i] = p.inc.x; inc[3 * i + 1] = p.inc.y; inc[3 * i + 2] = p.inc.z; }
        if (pw) { pw[3 * i] = p.pow.x; pw[3 * i + 1] = p.pow.y; pw[3 * i + 2] = p.pow.z; }
        if (key) key[i] = m->key[i];
        if (idx) idx[i] = m->idx[i];
    }
    if (start) std::copy(m->start.begin(), m->start.end(), start);
    if (grid) { grid[0] = m->origin.x; grid[1] = m->origin.y; grid[2] = m->origin.z; grid[3] = m->cell; }
}

// Builds a photon map from given photons (positions only matter for the
// grid) — the host sort the device grid is compared with.
void* orc_photon_map_from(const float* pos, int64_t n, float cell) {
    PhotonMapO* m = new PhotonMapO();
    m->ph.resize(n);
    for (int64_t i = 0; i < n; i++) {
        m->ph[i].pos = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        m->ph[i].inc = V3(0, 1, 0);
        m->ph[i].pow = V3(0.0f);
    }
    m->build(cell);
    return m;
}

// Sets every photon's power and incoming direction (synthetic maps).
void orc_photon_set_all(void* h, const float* power, const float* inc) {
    PhotonMapO* m = (PhotonMapO*)h;
    for (auto& p : m->ph) {
        p.pow = V3(power[0], power[1], power[2]);
        p.inc = V3(inc[0], inc[1], inc[2]);
    }
}

// k-nearest gather for a batch of points: count, kernel radius and the
// gathered photon indices (k per point, -1 padded).
void orc_photon_gather(void* h, const float* pts, int64_t n, int32_t k, float max_r, int32_t* count, float* radius,
                       int32_t* out_idx) {
    PhotonMapO* m = (PhotonMapO*)h;
    std::vector<int> hi(std::max(1, (int)k));
    std::vector<float> hd(std::max(1, (int)k));
    for (int64_t i = 0; i < n; i++) {
        float r;
        int c = m->gather(V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), k, max_r, hi.data(), hd.data(), r);
        if (count) count[i] = c;
        if (radius) radius[i] = r;
        if (out_idx)
            for (int j = 0; j < k; j++) out_idx[i * k + j] = j < c ? hi[j] : -1;
    }
}

void orc_photon_free(void* h) { delete (PhotonMapO*)h; }

// ── Unit probes for material / light / LUT tests ─────────────────────────
// A shading point is described by (normal, tangent, bitangent, front_face,
// relative_eta); random draws come from the caller's `rnd` list.
static HitRecord probe_rec(const float* frame) {
    HitRecord rec;
    rec.point = V3(0.0f);
    rec.normal = V3(frame[0], frame[1], frame[2]);
    rec.tangent = V3(frame[3], frame[4], frame[5]);
    rec.bitangent = V3(frame[6], frame[7], frame[8]);
    rec.front_face = frame[9] != 0.0f;
    rec.relative_eta = frame[10];
    rec.shading_ng = rec.normal;
    return rec;
}

// out: f[3] = Evaluate(V,L), pdf = Pdf(V,L), fd[3] = EvaluateDirect(L,V,N)
// ThinFilm.Evaluate at n (cos, film ior, thickness nm, f0 rgb) tuples -> rgb.
int orc_thin_film(const float* args, int64_t n, float* out) {
    for (int64_t k = 0; k < n; k++) {
        const float* a = args + 6 * k;
        V3 r = thin_film_fresnel(a[0], a[1], a[2], V3(a[3], a[4], a[5]));
        out[3 * k] = r.x; out[3 * k + 1] = r.y; out[3 * k + 2] = r.z;
    }
    return 0;
}

int orc_bsdf_eval(const rt_material* m, const float* frame, const float* V, const float* L, float* f, float* pdf,
                  float* fd) {
    std::unique_ptr<Material> mat(make_material(*m));
    if (!mat) return RT_ERR_INVALID_ARGUMENT;
    HitRecord rec = probe_rec(frame);
    V3 v(V[0], V[1], V[2]), l(L[0], L[1], L[2]);
    V3 e = mat->evaluate(v, l, rec);
    f[0] = e.x; f[1] = e.y; f[2] = e.z;
    *pdf = mat->pdf(v, l, rec);
    V3 d = mat->evaluate_direct(l, v, rec.normal, rec);
    fd[0] = d.x; fd[1] = d.y; fd[2] = d.z;
    return 0;
}

// out: wo[3], f[3], pdf, flags (bit0 valid, bit1 delta, bit2 has_next_abs,
// bits 4-5 transition, bit6 caustic caster), used = random draws consumed
int orc_bsdf_sample(const rt_material* m, const float* frame, const float* V, const float* rnd, int32_t nrnd, float* wo,
                    float* f, float* pdf, int32_t* flags, int32_t* used) {
    std::unique_ptr<Material> mat(make_material(*m));
    if (!mat) return RT_ERR_INVALID_ARGUMENT;
    HitRecord rec = probe_rec(frame);
    SamplerCtx& c = t_smp;
    c.fixed = rnd; c.nfixed = nrnd; c.fixed_pos = 0;
    BsdfSample s;
    bool ok = mat->sample(V3(V[0], V[1], V[2]), rec, s);
    *used = c.fixed_pos;
    c.fixed = nullptr;
    int fl = 0;
    if (ok) {
        fl |= 1;
        if (s.is_delta) fl |= 2;
        if (s.has_next_absorption) fl |= 4;
        fl |= (s.transition & 3) << 4;
        if (s.caustic_caster) fl |= 64;
        wo[0] = s.wo.x; wo[1] = s.wo.y; wo[2] = s.wo.z;
        f[0] = s.f.x; f[1] = s.f.y; f[2] = s.f.z;
        *pdf = s.pdf;
    }
    *flags = fl;
    return 0;
}

int orc_energy_lut(float* E, float* Eavg) {
    const EnergyLut& L = energy_lut();
    std::memcpy(E, L.E, sizeof(L.E));
    std::memcpy(Eavg, L.Eavg, sizeof(L.Eavg));
    return 0;
}

// ShadowTransmittance / ShadowAbsorption probes
int orc_shadow_tr(const rt_material* m, const float* frame, const float* wi, float* tr, float* sig) {
    std::unique_ptr<Material> mat(make_material(*m));
    if (!mat) return RT_ERR_INVALID_ARGUMENT;
    HitRecord rec = probe_rec(frame);
    V3 t = mat->shadow_transmittance(V3(wi[0], wi[1], wi[2]), rec);
    V3 s = mat->shadow_absorption(rec);
    tr[0] = t.x; tr[1] = t.y; tr[2] = t.z;
    sig[0] = s.x; sig[1] = s.y; sig[2] = s.z;
    return 0;
}

// One NEE sample of light `li` from point p with normal n (sample index si),
// random draws from `rnd`.  out: in_shadow, color[3], dir[3], dist, pdf
// (PdfSolidAngle of the returned direction).
int orc_light_sample(void* sp, int32_t li, const float* p, const float* n, int32_t si, const float* rnd, int32_t nrnd,
                     int32_t* in_shadow, float* color, float* dir, float* dist, float* pdf) {
    Scene* s = (Scene*)sp;
    if (li < 0 || li >= (int)s->light_ptrs.size()) return RT_ERR_INVALID_ARGUMENT;
    SamplerCtx& c = t_smp;
    c.fixed = rnd; c.nfixed = nrnd; c.fixed_pos = 0;
    const Light* L = s->light_ptrs[li];
    V3 P(p[0], p[1], p[2]);
    LightSample ls = L->illuminate(P, V3(n[0], n[1], n[2]), s->world, si);
    c.fixed = nullptr;
    *in_shadow = ls.in_shadow;
    color[0] = ls.color.x; color[1] = ls.color.y; color[2] = ls.color.z;
    dir[0] = ls.dir.x; dir[1] = ls.dir.y; dir[2] = ls.dir.z;
    *dist = ls.dist;
    *pdf = L->pdf_solid_angle(P, ls.dir);
    return 0;
}

float orc_light_pdf(void* sp, int32_t li, const float* p, const float* wi) {
    Scene* s = (Scene*)sp;
    return s->light_ptrs[li]->pdf_solid_angle(V3(p[0], p[1], p[2]), V3(wi[0], wi[1], wi[2]));
}

// Camera basis (Camera.cs:26-51): origin, llc, horizontal, vertical, u, v, lens radius
int orc_camera_basis(const rt_camera* cam, float* out) {
    Camera c(*cam);
    V3 vs[6] = {c.origin, c.llc, c.horiz, c.vert, c.u, c.v};
    for (int k = 0; k < 6; k++) { out[3 * k] = vs[k].x; out[3 * k + 1] = vs[k].y; out[3 * k + 2] = vs[k].z; }
    out[18] = c.lens_radius;
    return 0;
}

}  // extern "C"

extern "C" {

// Batched BSDF probes over n (V, L) pairs with one shading frame.
int orc_bsdf_eval_batch(const rt_material* m, const float* frame, const float* V, const float* L, int64_t n, float* f,
                        float* pdf, float* fd) {
    std::unique_ptr<Material> mat(make_material(*m));
    if (!mat) return RT_ERR_INVALID_ARGUMENT;
    HitRecord rec = probe_rec(frame);
    for (int64_t i = 0; i < n; i++) {
        V3 v(V[3 * i], V[3 * i + 1], V[3 * i + 2]), l(L[3 * i], L[3 * i + 1], L[3 * i + 2]);
        V3 e = mat->evaluate(v, l, rec);
        f[3 * i] = e.x; f[3 * i + 1] = e.y; f[3 * i + 2] = e.z;
        pdf[i] = mat->pdf(v, l, rec);
        V3 d = mat->evaluate_direct(l, v, rec.normal, rec);
        fd[3 * i] = d.x; fd[3 * i + 1] = d.y; fd[3 * i + 2] = d.z;
    }
    return 0;
}

// n Sample() calls with the PCG stream seeded by `seed` (the reference's
// PRNG fallback of Sampler.Sample1D when no Sobol context is open).
int orc_bsdf_sample_batch(const rt_material* m, const float* frame, const float* V, int64_t n, uint64_t seed, float* wo,
                          float* f, float* pdf, int32_t* flags) {
    std::unique_ptr<Material> mat(make_material(*m));
    if (!mat) return RT_ERR_INVALID_ARGUMENT;
    HitRecord rec = probe_rec(frame);
    SamplerCtx& c = t_smp;
    c.active = false;
    c.fixed = nullptr;
    c.rng = seed * 2654435761ull + 0x9E3779B97F4A7C15ull;
    V3 v(V[0], V[1], V[2]);
    for (int64_t i = 0; i < n; i++) {
        BsdfSample s;
        bool ok = mat->sample(v, rec, s);
        int fl = 0;
        if (ok) {
            fl |= 1;
            if (s.is_delta) fl |= 2;
            if (s.has_next_absorption) fl |= 4;
            fl |= (s.transition & 3) << 4;
            if (s.caustic_caster) fl |= 64;
            wo[3 * i] = s.wo.x; wo[3 * i + 1] = s.wo.y; wo[3 * i + 2] = s.wo.z;
            f[3 * i] = s.f.x; f[3 * i + 1] = s.f.y; f[3 * i + 2] = s.f.z;
            pdf[i] = s.pdf;
        }
        flags[i] = fl;
    }
    return 0;
}

// Closest hit with explicit [tmin, tmax] through the world (BVH) or through a
// flat HittableList over every object (the brute-force reference of
// BvhEquivalenceTests.cs:54-72).
int orc_trace_rays_ex(void* sp, const float* rays, int64_t n, float tmin, float tmax, int32_t linear, int32_t* prim,
                      float* t) {
    Scene* s = (Scene*)sp;
    HittableList flat;
    flat.objs = s->objects;
    const Hittable* w = linear ? (const Hittable*)&flat : (const Hittable*)s->world;
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        Ray ray(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]));
        HitRecord rec;
        bool h = w->hit(ray, tmin, tmax, rec);
        prim[i] = h ? rec.prim : -1;
        t[i] = h ? rec.t : 0.0f;
    }
    flat.objs.clear();
    return 0;
}

}  // extern "C"
