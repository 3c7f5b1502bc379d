// orc_core.h — CPU oracle, core math.  TEST INFRASTRUCTURE ONLY: this is the
// checker the HIP path is compared against; nothing in the product links it.
//
// Float semantics follow the reference's System.Numerics.Vector3 / MathF
// code: fp32, no contraction (built with -ffp-contract=off), operations in
// the reference's left-to-right order.  Dot = (x*x' + y*y') + z*z'
// (the SSE4.1 DPPS reduction order with a zero 4th lane), Normalize =
// componentwise true division by Length.  The exact bits of .NET's Vector3
// reductions and MathF transcendentals cannot be checked here (no .NET
// runtime); the primary-hit path uses none of the transcendentals.
#pragma once
#include <cmath>
#include <cstdint>
#include <cfloat>
#include <limits>

namespace orc {

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit V3(float s) : x(s), y(s), z(s) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, V3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
inline V3& operator*=(V3& a, V3 b) { a = a * b; return a; }
inline V3& operator*=(V3& a, float s) { a = a * s; return a; }
inline V3& operator/=(V3& a, float s) { a = a / s; return a; }
inline bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length_sq(V3 a) { return dot(a, a); }
inline float length(V3 a) { return std::sqrt(length_sq(a)); }
inline V3 normalize(V3 a) { return a / length(a); }
// IEEE 754-2019 maximum/minimum (NaN-propagating), as MathF.Max/MathF.Min.
inline float fmax_p(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : (b > a ? b : (std::signbit(a) ? b : a));
}
inline float fmin_p(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    return a < b ? a : (b < a ? b : (std::signbit(a) ? a : b));
}
inline V3 vmin(V3 a, V3 b) { return V3(fmin_p(a.x, b.x), fmin_p(a.y, b.y), fmin_p(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return V3(fmax_p(a.x, b.x), fmax_p(a.y, b.y), fmax_p(a.z, b.z)); }
inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline V3 lerp(V3 a, V3 b, float t) { return a + (b - a) * t; }
inline float lerpf(float a, float b, float t) { return a + (b - a) * t; }

constexpr float kPi = 3.14159265358979323846f;      // MathF.PI
constexpr float kEpsilon = 1e-4f;                    // MathUtils.Epsilon
constexpr float kInfinity = FLT_MAX;                 // MathUtils.Infinity
constexpr float kInv4Pi = 0.0795774715f;             // MathUtils.Inv4Pi

// MathUtils.Luminance (MathUtils.cs:209)
inline float luminance(V3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }
// MathUtils.NearZero (MathUtils.cs:199-203)
inline bool near_zero(V3 v) {
    const float s = 1e-8f;
    return std::fabs(v.x) < s && std::fabs(v.y) < s && std::fabs(v.z) < s;
}
// MathUtils.OffsetOrigin (MathUtils.cs:217)
inline V3 offset_origin(V3 p, V3 n) { return p + n * kEpsilon; }
// MathUtils.Reflect / Refract (MathUtils.cs:157-165)
inline V3 reflect(V3 v, V3 n) { return v - 2.0f * dot(v, n) * n; }
inline V3 refract(V3 uv, V3 n, float eta) {
    float cos_t = std::fmin(dot(-uv, n), 1.0f);
    V3 perp = eta * (uv + cos_t * n);
    V3 par = -std::sqrt(std::fabs(1.0f - length_sq(perp))) * n;
    return perp + par;
}
// MathUtils.Schlick (MathUtils.cs:167-176)
inline float schlick(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 *= r0;
    float x = 1.0f - cosine;
    float x2 = x * x;
    return r0 + (1.0f - r0) * x2 * x2 * x;
}
// MathUtils.FresnelDielectric (MathUtils.cs:189-197)
inline float fresnel_dielectric(float cos_i, float eta) {
    cos_i = clampf(cos_i, 0.0f, 1.0f);
    float sin2_i = std::fmax(0.0f, 1.0f - cos_i * cos_i);
    float sin2_t = eta * eta * sin2_i;
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = std::sqrt(1.0f - sin2_t);
    float r_parl = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    float r_perp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    return 0.5f * (r_parl * r_parl + r_perp * r_perp);
}
// MathUtils.ShadowTerminatorTerm (MathUtils.cs:44-61)
inline float shadow_terminator(V3 ng, V3 ns, V3 w) {
    float ngw = dot(ng, w);
    if (ngw <= 0.0f) return 0.0f;
    float nsw = dot(ns, w);
    if (nsw <= 0.0f) return 0.0f;
    float ngns = dot(ng, ns);
    if (ngns >= 0.9999f) return 1.0f;
    float g = ngw / (nsw * std::fmax(ngns, 1e-4f));
    if (g >= 1.0f) return 1.0f;
    if (g <= 0.0f) return 0.0f;
    float g2 = g * g;
    return (-g2 * g) + g2 + g;
}

// ── Ray (Ray.cs:39-57) ───────────────────────────────────────────────────
struct Ray {
    V3 o, d, inv;
    // RayDifferential of a camera ray (RayDifferential.cs:28-43): both
    // auxiliary rays start at dox
    bool has_diff = false;
    V3 dox, ddx, ddy;
    Ray() {}
    Ray(V3 origin, V3 dir) : o(origin), d(dir), inv(V3(1.0f) / dir) {}
    V3 at(float t) const { return o + t * d; }
};

// ── AABB (AABB.cs:28-78) ─────────────────────────────────────────────────
struct AABB {
    V3 mn, mx;
    AABB() : mn(FLT_MAX), mx(-FLT_MAX) {}
    AABB(V3 a, V3 b) : mn(a), mx(b) {}
    static AABB empty() { return AABB(V3(FLT_MAX), V3(-FLT_MAX)); }
    bool hit(const Ray& r, float tmin, float tmax) const {
        if ((mn.x > mx.x) | (mn.y > mx.y) | (mn.z > mx.z)) return false;
        V3 t0 = (mn - r.o) * r.inv;
        V3 t1 = (mx - r.o) * r.inv;
        V3 ts = vmin(t0, t1);
        V3 tl = vmax(t0, t1);
        float te = fmax_p(tmin, fmax_p(ts.x, fmax_p(ts.y, ts.z)));
        float tx = fmin_p(tmax, fmin_p(tl.x, fmin_p(tl.y, tl.z)));
        return tx >= te;
    }
    float area() const {
        V3 d = mx - mn;
        if (d.x < 0.0f || d.y < 0.0f || d.z < 0.0f) return 0.0f;
        return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
};
inline AABB surround(const AABB& a, const AABB& b) { return AABB(vmin(a.mn, b.mn), vmax(a.mx, b.mx)); }

// ── Owen-scrambled Sobol (OwenSobol.cs:48-210) ──────────────────────────
inline uint32_t reverse_bits(uint32_t v) {
    v = ((v & 0xaaaaaaaau) >> 1) | ((v & 0x55555555u) << 1);
    v = ((v & 0xccccccccu) >> 2) | ((v & 0x33333333u) << 2);
    v = ((v & 0xf0f0f0f0u) >> 4) | ((v & 0x0f0f0f0fu) << 4);
    v = ((v & 0xff00ff00u) >> 8) | ((v & 0x00ff00ffu) << 8);
    return (v >> 16) | (v << 16);
}
inline uint32_t lk_permute(uint32_t v, uint32_t seed) {
    v += seed;
    v ^= v * 0x6c50b47cu;
    v ^= v * 0xb82f1e52u;
    v ^= v * 0xc7afe638u;
    v ^= v * 0x8d22f6e6u;
    return v;
}
inline uint32_t sobol_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
inline uint32_t owen_scramble(uint32_t v, uint32_t seed) {
    return reverse_bits(lk_permute(reverse_bits(v), seed));
}
// Dim-1 direction numbers p(x) = x^2 + x + 1, m1 = 1 (OwenSobol.cs:132-142),
// generated here by the Sobol recurrence v_k = v_{k-1} ^ (v_{k-1} >> 1) from
// v_0 = 1<<31 rather than tabulated.
inline uint32_t sobol_dim1(uint32_t index) {
    uint32_t v = 0, dir = 0x80000000u;
    while (index != 0u) {
        if (index & 1u) v ^= dir;
        index >>= 1;
        dir ^= dir >> 1;
    }
    return v;
}
// PRNG sampler mode: the reference's System.Random streams are not
// reproducible (thread-local, seeded by thread start order), so this mode
// draws a counter-based hash of (index, dim, seed) instead: the 32-bit
// PCG-RXS-M-XS permutation as a hash (Jarzynski and Olano, JCGT 2020),
// chained over the three words; value = top 24 bits * 2^-24.
inline uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline float prng_sample(uint32_t index, uint32_t dim, uint32_t pixel_seed) {
    uint32_t h = pcg_hash(pixel_seed ^ pcg_hash(index ^ pcg_hash(dim + 0x9e3779b9u)));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
inline float sobol_sample(uint32_t index, uint32_t dim, uint32_t pixel_seed) {
    uint32_t v;
    if (dim == 0u) v = reverse_bits(index);
    else if (dim == 1u) v = sobol_dim1(index);
    else {
        uint32_t shuffle = sobol_hash(dim * 0x68bc21ebu + 0xc7afe638u);
        v = reverse_bits(owen_scramble(index, shuffle));
    }
    uint32_t scr = sobol_hash(pixel_seed ^ sobol_hash(dim * 0xa511e9b3u + 0x736f6c34u));
    v = owen_scramble(v, scr);
    return (float)(v >> 8) * (1.0f / 16777216.0f);
}

}  // namespace orc
