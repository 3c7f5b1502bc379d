// adjoint.h — derivatives of the Disney BSDF with respect to the textured
// parameters p = (base r, g, b, roughness, metallic), for the path-replay
// adjoint (no reference counterpart).
//
// Forward-mode dual numbers: every quantity carries its value and its five
// partial derivatives with respect to p.  The functions below restate the
// parameter derivation (disney_gmat) and the two evaluations the integrator
// weights paths with — Evaluate (the sampled bounce, DisneyBsdf.cs:649-771)
// and EvaluateDirect (NEE, DisneyBsdf.cs:418-535) — for a textured material,
// whose spec_trans and diff_trans are 0 (the transmission lobes vanish).
// The clearcoat lobe does not depend on p and contributes its value only.
// A clamp passes the derivative through inside its closed interval.
#pragma once
#include "shade.h"

namespace rt {

constexpr int kNd = 5;

struct Dv {
    float v;
    float d[kNd];
};
__device__ __forceinline__ Dv dconst(float v) {
    Dv r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < kNd; k++) r.d[k] = 0.0f;
    return r;
}
__device__ __forceinline__ Dv dvar(float v, int k) {
    Dv r = dconst(v);
    r.d[k] = 1.0f;
    return r;
}
__device__ __forceinline__ Dv operator+(Dv a, Dv b) {
    Dv r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < kNd; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
__device__ __forceinline__ Dv operator-(Dv a, Dv b) {
    Dv r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < kNd; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
__device__ __forceinline__ Dv operator*(Dv a, Dv b) {
    Dv r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < kNd; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
__device__ __forceinline__ Dv operator/(Dv a, Dv b) {
    Dv r;
    r.v = a.v / b.v;
    const float ib = 1.0f / b.v;
#pragma unroll
    for (int k = 0; k < kNd; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
__device__ __forceinline__ Dv operator*(Dv a, float s) {
    Dv r;
    r.v = a.v * s;
#pragma unroll
    for (int k = 0; k < kNd; k++) r.d[k] = a.d[k] * s;
    return r;
}
__device__ __forceinline__ Dv operator*(float s, Dv a) { return a * s; }
__device__ __forceinline__ Dv operator/(Dv a, float s) { return a * (1.0f / s); }
__device__ __forceinline__ Dv operator+(Dv a, float s) {
    a.v += s;
    return a;
}
__device__ __forceinline__ Dv operator+(float s, Dv a) { return a + s; }
__device__ __forceinline__ Dv operator-(Dv a, float s) {
    a.v -= s;
    return a;
}
__device__ __forceinline__ Dv operator-(float s, Dv a) { return dconst(s) - a; }
__device__ __forceinline__ Dv operator/(float s, Dv a) { return dconst(s) / a; }
__device__ __forceinline__ Dv dsqrt(Dv a) {
    Dv r;
    r.v = sqrtf(a.v);
    const float h = r.v > 0.0f ? 0.5f / r.v : 0.0f;
#pragma unroll
    for (int k = 0; k < kNd; k++) r.d[k] = a.d[k] * h;
    return r;
}
// fmaxf(a, b) with a constant b: the larger operand and its derivative
__device__ __forceinline__ Dv dmax(Dv a, float b) { return a.v >= b ? a : dconst(b); }
__device__ __forceinline__ Dv dclamp(Dv a, float lo, float hi) {
    if (a.v < lo) return dconst(lo);
    if (a.v > hi) return dconst(hi);
    return a;
}

struct D3 {
    Dv x, y, z;
};
__device__ __forceinline__ D3 d3(Dv x, Dv y, Dv z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 d3c(f3 a) { return D3{dconst(a.x), dconst(a.y), dconst(a.z)}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, D3 b) { return D3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ D3 operator/(D3 a, D3 b) { return D3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, Dv s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ D3 operator*(Dv s, D3 a) { return a * s; }
__device__ __forceinline__ D3 operator/(D3 a, Dv s) { return D3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ D3 operator*(D3 a, float s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ D3 operator*(float s, D3 a) { return a * s; }
__device__ __forceinline__ D3 operator*(D3 a, f3 b) { return D3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ D3 operator+(D3 a, f3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 dsplat(Dv s) { return D3{s, s, s}; }
__device__ __forceinline__ Dv dlum(D3 c) { return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z; }
__device__ __forceinline__ D3 dlerp3(D3 a, D3 b, Dv t) { return a + (b - a) * t; }
__device__ __forceinline__ D3 dlerp3(D3 a, D3 b, float t) { return a + (b - a) * t; }

// Derived parameters that depend on p (the rest are read from the per-hit
// float record, DisneyP).
struct DisneyD {
    D3 base, f0;
    Dv metallic, roughness, ax, ay, alpha_iso, eavg, dw;
};

// EnergyCompensationLut.SampleEavg / SampleE with the alpha derivative
__device__ __forceinline__ Dv eavg_d(const float* Eavg, Dv alpha) {
    alpha = dclamp(alpha, 0.001f, 1.0f);
    Dv fal = alpha * 32.0f - 0.5f;
    int a0 = (int)floorf(fal.v);
    a0 = a0 < 0 ? 0 : (a0 > 31 ? 31 : a0);
    int a1 = a0 + 1 > 31 ? 31 : a0 + 1;
    Dv ta = dclamp(fal - (float)a0, 0.0f, 1.0f);
    return Eavg[a0] + (Eavg[a1] - Eavg[a0]) * ta;
}
__device__ __forceinline__ Dv lut_e_d(const float* E, float mu, Dv alpha) {
    mu = clampf(mu, 0.0f, 1.0f);
    alpha = dclamp(alpha, 0.001f, 1.0f);
    float fmu = mu * 32.0f - 0.5f;
    Dv fal = alpha * 32.0f - 0.5f;
    int m0 = min(max((int)floorf(fmu), 0), 31), a0 = min(max((int)floorf(fal.v), 0), 31);
    int m1 = min(m0 + 1, 31), a1 = min(a0 + 1, 31);
    float tu = clampf(fmu - (float)m0, 0.0f, 1.0f);
    Dv ta = dclamp(fal - (float)a0, 0.0f, 1.0f);
    float e00 = E[m0 * 32 + a0], e01 = E[m0 * 32 + a1], e10 = E[m1 * 32 + a0], e11 = E[m1 * 32 + a1];
    Dv e0 = e00 + (e01 - e00) * ta;
    Dv e1 = e10 + (e11 - e10) * ta;
    return e0 + (e1 - e0) * tu;
}

// disney_gmat's p-dependent part for material knobs m and parameters p.
__device__ __forceinline__ DisneyD disney_d(const rt_material& m, const float p[kNd], const float* Eavg) {
    DisneyD D;
    D3 c = d3(dvar(p[0], 0), dvar(p[1], 1), dvar(p[2], 2));
    D.base = c;
    D.metallic = dclamp(dvar(p[4], 4), 0.0f, 1.0f);
    D.roughness = dclamp(dvar(p[3], 3), 0.0f, 1.0f);
    float specular = clampf(m.specular, 0, 2), spec_tint = clampf(m.specular_tint, 0, 1);
    float ior = mx(m.ior, 1.0001f), aniso = clampf(m.anisotropic, 0, 1);
    Dv alpha = dmax(D.roughness * D.roughness, 0.001f);
    float aspect = sqrtf(1.0f - 0.9f * aniso);
    D.ax = dmax(alpha / aspect, 0.001f);
    D.ay = dmax(alpha * aspect, 0.001f);
    float r = (ior - 1.0f) / (ior + 1.0f);
    float scaled = r * r * 2.0f * specular;
    Dv lum = dlum(c);
    D3 tint = lum.v > 0.0f ? c / lum : d3c(splat(1.0f));
    D3 diel = dlerp3(d3c(splat(scaled)), tint * scaled, spec_tint);
    D.f0 = dlerp3(diel, c, D.metallic);
    D.alpha_iso = dsqrt(D.ax * D.ay);
    D.eavg = eavg_d(Eavg, D.alpha_iso);
    D.dw = 1.0f - D.metallic;   // (1 - metallic)(1 - spec_trans)(1 - diff_trans), both 0
    return D;
}

__device__ __forceinline__ Dv d_ggx_d(f3 h, Dv ax, Dv ay) {
    Dv hx = h.x / ax, hy = h.y / ay;
    Dv d = (hx * hx + hy * hy) + h.z * h.z;
    return 1.0f / dmax(kPi * ax * ay * d * d, 1e-20f);
}
__device__ __forceinline__ Dv g1_ggx_d(f3 w, Dv ax, Dv ay) {
    float c2 = w.z * w.z;
    if (c2 < 1e-14f) return dconst(1.0f / (1.0f + 1e10f));
    Dv a = ax * w.x, b = ay * w.y;
    Dv num = a * a + b * b;
    Dv lam = 0.5f * (dsqrt(1.0f + num / c2) - 1.0f);
    return 1.0f / (1.0f + lam);
}
__device__ __forceinline__ D3 fresnel_d(float c, D3 f0) {
    float w = schlick_weight(c);
    return f0 + (d3c(splat(1.0f)) - f0) * w;
}
__device__ __forceinline__ D3 ms_weight_d(const DisneyD& D, const float* E, float ndv, float ndl) {
    Dv eo = lut_e_d(E, ndv, D.alpha_iso), ei = lut_e_d(E, ndl, D.alpha_iso), ea = D.eavg;
    Dv om = dmax(1.0f - ea, 1e-6f);
    D3 favg = D.f0 + (d3c(splat(1.0f)) - D.f0) * (1.0f / 21.0f);
    D3 den = d3c(splat(1.0f)) - favg * (1.0f - ea);
    den = d3(dmax(den.x, 1e-6f), dmax(den.y, 1e-6f), dmax(den.z, 1e-6f));
    D3 fms = favg * favg * ea / den;
    return fms * ((1.0f - eo) * (1.0f - ei) / om);
}
__device__ __forceinline__ D3 sheen_d(const DisneyP& P, const DisneyD& D, float charlie) {
    Dv lum = dlum(D.base);
    D3 tint = lum.v > 0.0f ? D.base / lum : d3c(splat(1.0f));
    D3 col = dlerp3(d3c(splat(1.0f)), tint, P.sheen_tint);
    return col * (P.sheen * charlie);
}

// Evaluate (DisneyBsdf.cs:649-771) with derivatives.
__device__ D3 disney_eval_d(const DisneyP& P, const DisneyD& D, const float* E, f3 V, f3 L, const Surf& s) {
    f3 N = s.n;
    float ndl = dot(N, L), ndv = dot(N, V);
    D3 zero = d3c(splat(0.0f));
    if (ndv <= 0.0f || ndl <= 0.0f) return zero;
    f3 hr = V + L;
    float hl2 = length_sq(hr);
    if (hl2 < 1e-14f) return zero;
    f3 H = hr / sqrtf(hl2);
    float ndh = fmaxf(dot(N, H), 0.0f), vdh = fmaxf(dot(V, H), 0.0f), ldh = fmaxf(dot(L, H), 0.0f);
    D3 diffuse = zero;
    if (D.dw.v > 0.0f) {
        Dv fd90 = 0.5f + 2.0f * D.roughness * (ldh * ldh);
        float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
        Dv fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
        diffuse = D.base * fd * (D.dw / kPi);
    }
    D3 sheen = zero;
    if (P.sheen > 0.0f && D.dw.v > 0.0f) sheen = sheen_d(P, D, charlie_brdf(ndv, ndl, ndh, P.sheen_rough));
    Frame fr = shading_frame(s, P.rot);
    f3 vl = fr.to_local(V), ll = fr.to_local(L), hlc = fr.to_local(H);
    Dv Dg = d_ggx_d(hlc, D.ax, D.ay);
    Dv G = g1_ggx_d(vl, D.ax, D.ay) * g1_ggx_d(ll, D.ax, D.ay);
    D3 F = fresnel_d(vdh, D.f0);
    D3 spec = F * (Dg * G / fmaxf(4.0f * ndv * ndl, 1e-7f));
    f3 cc = splat(0.0f);
    if (P.clearcoat > 0.0f) {
        float cndh = fmaxf(dot(N, H), 0.0f), cvdh = fmaxf(dot(V, H), 0.0f);
        float a2 = P.cc_alpha * P.cc_alpha;
        float den = cndh * cndh * (a2 - 1.0f) + 1.0f;
        float cD = a2 / (kPi * den * den);
        float cG = smith_g1(ndv, P.cc_alpha) * smith_g1(ndl, P.cc_alpha);
        float cF = P.cc_f0 + (1.0f - P.cc_f0) * schlick_weight(cvdh);
        cc = splat(P.clearcoat * 0.25f * cD * cG * cF / fmaxf(4.0f * ndv * ndl, 1e-7f));
    }
    D3 ms = ms_weight_d(D, E, ndv, ndl) * (1.0f / kPi);
    return (((diffuse + sheen) + spec) + ms) + cc;
}

// EvaluateDirect (DisneyBsdf.cs:418-535) with derivatives.
__device__ D3 disney_eval_direct_d(const DisneyP& P, const DisneyD& D, const float* E, f3 L, f3 V, const Surf& s) {
    f3 N = s.n;
    D3 zero = d3c(splat(0.0f));
    float ndl = fmaxf(dot(N, L), 0.0f);
    if (ndl <= 0.0f) return zero;
    float ndv = fmaxf(dot(N, V), 0.001f);
    f3 H = normalize(L + V);
    float ndh = fmaxf(dot(N, H), 0.0f), vdh = fmaxf(dot(V, H), 0.0f), ldh = fmaxf(dot(L, H), 0.0f);
    D3 diffuse = zero;
    if (D.dw.v > 0.0f) {
        Dv fd90 = 0.5f + 2.0f * D.roughness * (ldh * ldh);
        float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
        Dv fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
        diffuse = D.base * fd * (D.dw * (ndl / kPi));
    }
    D3 spec = zero;
    // the delta test uses the per-hit alpha (DisneyP) so the branch matches
    // the float evaluation
    if (P.alpha > kDeltaAlpha) {
        Frame fr = shading_frame(s, P.rot);
        f3 vl = fr.to_local(V), ll = fr.to_local(L), hl = fr.to_local(H);
        Dv Dg = d_ggx_d(hl, D.ax, D.ay);
        Dv G = g1_ggx_d(vl, D.ax, D.ay) * g1_ggx_d(ll, D.ax, D.ay);
        D3 F = fresnel_d(vdh, D.f0);
        spec = F * (Dg * G / fmaxf(4.0f * ndv, 1e-6f));
    }
    f3 cc = splat(0.0f);
    if (P.clearcoat > 0.0f && P.cc_alpha > kDeltaAlpha) {
        float cndv = fmaxf(dot(N, V), 0.001f), cndl = dot(N, L);
        if (cndl > 0.0f) {
            float cndh = fmaxf(dot(N, H), 0.0f), cvdh = fmaxf(dot(V, H), 0.0f);
            float a2 = P.cc_alpha * P.cc_alpha;
            float den = cndh * cndh * (a2 - 1.0f) + 1.0f;
            float cD = a2 / (kPi * den * den);
            float cG = smith_g1(cndv, P.cc_alpha) * smith_g1(cndl, P.cc_alpha);
            float cF = P.cc_f0 + (1.0f - P.cc_f0) * schlick_weight(cvdh);
            cc = splat(P.clearcoat * 0.25f * cD * cG * cF / fmaxf(4.0f * cndv, 1e-6f));
        }
    }
    D3 sheen = zero;
    if (P.sheen > 0.0f && D.dw.v > 0.0f) sheen = sheen_d(P, D, charlie_brdf(ndv, ndl, ndh, P.sheen_rough) * ndl);
    D3 ms = ms_weight_d(D, E, ndv, ndl) * ((1.0f / kPi) * ndl);
    return (((diffuse + sheen) + spec) + ms) + cc;
}

// The weight a Disney lobe sampler returns (disney_scatter, DisneyBsdf.cs
// 1263-1709) for the sampled direction dir, without the division by the lobe
// probability (held fixed): the estimator weight of delta and fallback
// bounces.  The transmission lobes vanish for textured materials and the
// clearcoat weight does not depend on p (both return zero derivatives).
__device__ D3 lobe_att_d(int lobe, const DisneyP& P, const DisneyD& D, const float* E, f3 V, f3 dir, const Surf& s) {
    f3 N = s.n;
    D3 zero = d3c(splat(0.0f));
    if (lobe == kLDiffuse) {
        float ndv = fmaxf(dot(N, V), 0.001f), ndl = fmaxf(dot(N, dir), 0.001f);
        f3 hr = V + dir;
        float hl2 = length_sq(hr);
        f3 H = hl2 > 1e-7f ? hr / sqrtf(hl2) : N;
        float ldh = fmaxf(dot(dir, H), 0.0f);
        Dv fd90 = 0.5f + 2.0f * D.roughness * (ldh * ldh);
        float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
        Dv fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
        return D.base * fd * (1.0f - P.diff_trans);
    }
    if (lobe == kLSpecular) {
        Frame fr = shading_frame(s, P.rot);
        f3 H = normalize(V + dir);
        float vdh = fmaxf(dot(V, H), 1e-4f);
        return fresnel_d(vdh, D.f0) * g1_ggx_d(fr.to_local(dir), D.ax, D.ay);
    }
    if (lobe == kLSheen) {
        float ndv = fmaxf(dot(N, V), 1e-4f), ndl = fmaxf(dot(N, dir), 1e-4f);
        f3 hr = V + dir;
        float hl2 = length_sq(hr);
        f3 H = hl2 > 1e-7f ? hr / sqrtf(hl2) : N;
        float ndh = fmaxf(dot(N, H), 0.0f);
        return sheen_d(P, D, kPi * charlie_brdf(ndv, ndl, ndh, P.sheen_rough));
    }
    if (lobe == kLMultiscatter) {
        float ndv = fmaxf(dot(N, V), 1e-3f), ndl = fmaxf(dot(N, dir), 1e-3f);
        return ms_weight_d(D, E, ndv, ndl);
    }
    return zero;
}

// Per-channel logarithmic derivative dW/dp / W of a weight, scaled by w:
// out[c * kNd + j] = w_c dW_c/dp_j / W_c (0 where W_c = 0).
__device__ __forceinline__ void log_deriv(const D3& W, f3 w, float out[16]) {
    const Dv* ch[3] = {&W.x, &W.y, &W.z};
    const float wc[3] = {w.x, w.y, w.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float f = ch[c]->v;
        const float sc = f > 0.0f ? wc[c] / f : 0.0f;
#pragma unroll
        for (int j = 0; j < kNd; j++) out[c * kNd + j] = ch[c]->d[j] * sc;
    }
    out[15] = 0.0f;
}

// Gradient texture update: g (per parameter) spread over the bilinear
// footprint of texture coordinate uv with float atomics.
__device__ __forceinline__ void tex_scatter(float* dt, int W, int H, float u, float v, const float g[kNd]) {
    TexFoot f = tex_footprint(W, H, u, v);
    const int64_t plane = (int64_t)W * H;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (f.w[k] == 0.0f) continue;
#pragma unroll
        for (int j = 0; j < kNd; j++)
            if (g[j] != 0.0f) unsafeAtomicAdd(&dt[j * plane + f.idx[k]], f.w[k] * g[j]);
    }
}

}  // namespace rt
