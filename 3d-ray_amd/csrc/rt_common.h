// rt_common.h — fp32 vector math, Owen-Sobol sampler and primitive
// intersection shared by the host scene builder and the gfx950 kernels.
//
// Every operation keeps the reference's fp32 operation order
// (System.Numerics.Vector3 / MathF in src/RayTracer/Core/*.cs) and the whole
// library is compiled with -ffp-contract=off, IEEE-correct division/sqrt and
// no denormal flush, so the primary-hit path (camera ray, slab test,
// Möller–Trumbore, quadratic) produces the same bits on host and device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#define RT_HD __host__ __device__ __forceinline__

namespace rt {

struct f3 {
    float x, y, z;
};
RT_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
RT_HD f3 splat(float s) { return f3{s, s, s}; }
RT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RT_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
RT_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RT_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
RT_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
RT_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
RT_HD f3& operator+=(f3& a, f3 b) { a = a + b; return a; }
RT_HD f3& operator*=(f3& a, f3 b) { a = a * b; return a; }
RT_HD f3& operator*=(f3& a, float s) { a = a * s; return a; }
RT_HD bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
// Vector3.Dot with the DPPS reduction order (x*x' + y*y') + z*z'
RT_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RT_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RT_HD float length_sq(f3 a) { return dot(a, a); }
RT_HD float length(f3 a) { return sqrtf(length_sq(a)); }
RT_HD f3 normalize(f3 a) { return a / length(a); }
RT_HD float get(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RT_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
RT_HD f3 lerp3(f3 a, f3 b, float t) { return a + (b - a) * t; }
RT_HD float lerpf(float a, float b, float t) { return a + (b - a) * t; }
RT_HD float fmax3(f3 a) {
    // MathF.Max(x, MathF.Max(y, z)) on finite values
    float m = a.y > a.z ? a.y : a.z;
    return a.x > m ? a.x : m;
}
// Min/Max of two values where a NaN operand is never produced by the
// callers (box corners, finite vertex data).
RT_HD float mn(float a, float b) { return a < b ? a : b; }
RT_HD float mx(float a, float b) { return a > b ? a : b; }
RT_HD f3 vmin(f3 a, f3 b) { return f3{mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)}; }
RT_HD f3 vmax(f3 a, f3 b) { return f3{mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)}; }

constexpr float kPi = 3.14159265358979323846f;  // MathF.PI
constexpr float kEps = 1e-4f;                    // MathUtils.Epsilon
constexpr float kInfinity = FLT_MAX;             // MathUtils.Infinity
constexpr float kInv4Pi = 0.0795774715f;         // MathUtils.Inv4Pi

RT_HD float luminance(f3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }
RT_HD bool near_zero(f3 v) { return fabsf(v.x) < 1e-8f && fabsf(v.y) < 1e-8f && fabsf(v.z) < 1e-8f; }
RT_HD f3 offset_origin(f3 p, f3 n) { return p + n * kEps; }

// ── Owen-scrambled Sobol (OwenSobol.cs:48-210) ──────────────────────────
RT_HD uint32_t bitrev(uint32_t v) {
    v = ((v & 0xaaaaaaaau) >> 1) | ((v & 0x55555555u) << 1);
    v = ((v & 0xccccccccu) >> 2) | ((v & 0x33333333u) << 2);
    v = ((v & 0xf0f0f0f0u) >> 4) | ((v & 0x0f0f0f0fu) << 4);
    v = ((v & 0xff00ff00u) >> 8) | ((v & 0x00ff00ffu) << 8);
    return (v >> 16) | (v << 16);
}
RT_HD uint32_t sobol_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
RT_HD uint32_t owen(uint32_t v, uint32_t seed) {
    v = bitrev(v);
    v += seed;
    v ^= v * 0x6c50b47cu;
    v ^= v * 0xb82f1e52u;
    v ^= v * 0xc7afe638u;
    v ^= v * 0x8d22f6e6u;
    return bitrev(v);
}
// Dimension 1: the Sobol matrix of p(x) = x^2 + x + 1 (m1 = 1) is the
// Pascal triangle mod 2: direction k is built by k steps of d ^= d >> 1.
RT_HD uint32_t sobol_dim1(uint32_t index) {
    uint32_t v = 0u, d = 0x80000000u;
    while (index != 0u) {
        if (index & 1u) v ^= d;
        index >>= 1;
        d ^= d >> 1;
    }
    return v;
}
RT_HD float sobol(uint32_t index, uint32_t dim, uint32_t pixel_seed) {
    uint32_t v;
    if (dim == 0u) v = bitrev(index);
    else if (dim == 1u) v = sobol_dim1(index);
    else v = bitrev(owen(index, sobol_hash(dim * 0x68bc21ebu + 0xc7afe638u)));
    v = owen(v, sobol_hash(pixel_seed ^ sobol_hash(dim * 0xa511e9b3u + 0x736f6c34u)));
    return (float)(v >> 8) * (1.0f / 16777216.0f);
}
// PRNG sampler mode (SamplerKind.Prng, Sampler.cs:9-16).  The reference
// draws from a thread-local System.Random whose stream depends on thread
// scheduling, so no value of it is reproducible; here each draw is a
// counter-based hash of (sample index, dimension, pixel seed): the
// PCG-RXS-M-XS 32-bit permutation used as a hash (Jarzynski and Olano,
// "Hash Functions for GPU Rendering", JCGT 9(3), 2020), chained over the
// three words, then the same 24-bit float conversion as the Sobol path.
RT_HD uint32_t pcg_hash(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
RT_HD float prng_sample(uint32_t index, uint32_t dim, uint32_t pixel_seed) {
    const uint32_t h = pcg_hash(pixel_seed ^ pcg_hash(index ^ pcg_hash(dim + 0x9e3779b9u)));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
RT_HD uint32_t pixel_seed(int i, int j, uint32_t x) {
    return (((uint32_t)i * 73856093u) ^ ((uint32_t)j * 19349663u)) ^ x;
}

}  // namespace rt
