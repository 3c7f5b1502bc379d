// shade.h — device-side materials and lights for the wavefront shade stage.
//
// Semantics follow the reference's C# code (file:line at each function);
// per-material constants are precomputed on the host into GMat
// (rt_internal.h).  Random draws come from the path's Owen-Sobol stream in
// the reference's dimension order (Sampler.Sample1D, Sampler.cs:106-118).
#pragma once
#include "traverse.h"

namespace rt {

// Per-path sampler state: one Sobol dimension per draw.
// Sampler draw stream of one camera sample: (sample index, pixel seed,
// next dimension); prng (uniform over a render) selects the PRNG mode.
struct Rng {
    uint32_t index, seed, dim;
    bool prng;
    __device__ __forceinline__ float next() {
        const uint32_t d = dim++;
        return prng ? prng_sample(index, d, seed) : sobol(index, d, seed);
    }
};

// Surface record reconstructed from the hit entry (HitRecord.cs:9-163).
struct Surf {
    f3 p, n, tan, bitan;
    float t;
    bool front;
    int32_t mat;
    int32_t prim;
    float rel_eta;
};

RT_HD f3 face(f3 d, f3 outward, bool& front) {
    front = dot(d, outward) < 0.0f;
    return front ? outward : -outward;
}

// Frisvad ONB + local→world (SphereLight.cs:284-305)
__device__ __forceinline__ f3 local_to_world(f3 local, f3 w) {
    f3 u, v;
    if (w.z < -0.999f) {
        u = mk(0.0f, -1.0f, 0.0f);
        v = mk(-1.0f, 0.0f, 0.0f);
    } else {
        float a = 1.0f / (1.0f + w.z);
        float b = -w.x * w.y * a;
        u = mk(1.0f - w.x * w.x * a, b, -w.x);
        v = mk(b, 1.0f - w.y * w.y * a, -w.y);
    }
    f3 res = local.x * u + local.y * v + local.z * w;
    float l2 = length_sq(res);
    return l2 > 1e-8f ? res / sqrtf(l2) : w;
}

// Microfacet.BuildTangentFrame (Microfacet.cs)
__device__ __forceinline__ void tangent_frame(f3 n, f3& t, f3& b) {
    if (n.z < -0.999f) {
        t = mk(0.0f, -1.0f, 0.0f);
        b = mk(-1.0f, 0.0f, 0.0f);
    } else {
        float a = 1.0f / (1.0f + n.z);
        float bb = -n.x * n.y * a;
        t = mk(1.0f - n.x * n.x * a, bb, -n.x);
        b = mk(bb, 1.0f - n.y * n.y * a, -n.y);
    }
}

__device__ __forceinline__ f3 random_unit_vector(Rng& r) {
    float u = r.next();
    float v = r.next();
    float z = 1.0f - 2.0f * u;
    float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = 2.0f * kPi * v;
    return mk(rr * cosf(phi), rr * sinf(phi), z);
}

RT_HD f3 reflect(f3 v, f3 n) { return v - 2.0f * dot(v, n) * n; }
RT_HD f3 refract(f3 uv, f3 n, float eta) {
    float ct = fminf(dot(-uv, n), 1.0f);
    f3 perp = eta * (uv + ct * n);
    f3 par = -sqrtf(fabsf(1.0f - length_sq(perp))) * n;
    return perp + par;
}
RT_HD float schlick(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 *= r0;
    float x = 1.0f - cosine;
    float x2 = x * x;
    return r0 + (1.0f - r0) * x2 * x2 * x;
}
RT_HD float fresnel_dielectric(float ci, float eta) {
    ci = clampf(ci, 0.0f, 1.0f);
    float s2i = fmaxf(0.0f, 1.0f - ci * ci);
    float s2t = eta * eta * s2i;
    if (s2t >= 1.0f) return 1.0f;
    float ct = sqrtf(1.0f - s2t);
    float rpl = (ci - eta * ct) / (ci + eta * ct);
    float rpp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5f * (rpl * rpl + rpp * rpp);
}
RT_HD float shadow_terminator(f3 ng, f3 ns, f3 w) {
    float ngw = dot(ng, w);
    if (ngw <= 0.0f) return 0.0f;
    float nsw = dot(ns, w);
    if (nsw <= 0.0f) return 0.0f;
    float ngns = dot(ng, ns);
    if (ngns >= 0.9999f) return 1.0f;
    float g = ngw / (nsw * fmaxf(ngns, 1e-4f));
    if (g >= 1.0f) return 1.0f;
    if (g <= 0.0f) return 0.0f;
    float g2 = g * g;
    return (-g2 * g) + g2 + g;
}

// ── microfacet (Microfacet.cs) ────────────────────────────────────────────
__device__ __forceinline__ float d_ggx(f3 h, float ax, float ay) {
    float hx = h.x / ax, hy = h.y / ay;
    float d = hx * hx + hy * hy + h.z * h.z;
    return 1.0f / fmaxf(kPi * ax * ay * d * d, 1e-20f);
}
__device__ __forceinline__ float g1_ggx(f3 w, float ax, float ay) {
    float c2 = w.z * w.z;
    float lam;
    if (c2 < 1e-14f) lam = 1e10f;
    else {
        float num = (ax * w.x) * (ax * w.x) + (ay * w.y) * (ay * w.y);
        lam = 0.5f * (-1.0f + sqrtf(1.0f + num / c2));
    }
    return 1.0f / (1.0f + lam);
}
__device__ __forceinline__ float smith_g1(float ndx, float alpha) {
    float a2 = alpha * alpha;
    float den = ndx + sqrtf(a2 + (1.0f - a2) * (ndx * ndx));
    return 2.0f * ndx / fmaxf(den, 1e-7f);
}
__device__ __forceinline__ f3 sample_vndf(f3 v, float ax, float ay, float u1, float u2) {
    f3 vh = normalize(mk(ax * v.x, ay * v.y, v.z));
    float l2 = vh.x * vh.x + vh.y * vh.y;
    f3 t1 = l2 > 0.0f ? mk(-vh.y, vh.x, 0.0f) / sqrtf(l2) : mk(1.0f, 0.0f, 0.0f);
    f3 t2 = cross(vh, t1);
    float r = sqrtf(u1);
    float phi = 2.0f * kPi * u2;
    float a = r * cosf(phi);
    float b = r * sinf(phi);
    float s = 0.5f * (1.0f + vh.z);
    b = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - a * a)) + s * b;
    f3 nh = a * t1 + b * t2 + sqrtf(fmaxf(0.0f, 1.0f - a * a - b * b)) * vh;
    return normalize(mk(ax * nh.x, ay * nh.y, fmaxf(0.0f, nh.z)));
}

// EnergyCompensationLut.SampleE (EnergyCompensationLut.cs:140-160)
__device__ __forceinline__ float lut_e(const float* E, float mu, float alpha) {
    mu = clampf(mu, 0.0f, 1.0f);
    alpha = clampf(alpha, 0.001f, 1.0f);
    float fmu = mu * 32.0f - 0.5f, fal = alpha * 32.0f - 0.5f;
    int m0 = min(max((int)floorf(fmu), 0), 31), a0 = min(max((int)floorf(fal), 0), 31);
    int m1 = min(m0 + 1, 31), a1 = min(a0 + 1, 31);
    float tu = clampf(fmu - (float)m0, 0.0f, 1.0f), ta = clampf(fal - (float)a0, 0.0f, 1.0f);
    float e00 = E[m0 * 32 + a0], e01 = E[m0 * 32 + a1], e10 = E[m1 * 32 + a0], e11 = E[m1 * 32 + a1];
    float e0 = e00 + (e01 - e00) * ta;
    float e1 = e10 + (e11 - e10) * ta;
    return e0 + (e1 - e0) * tu;
}

// SheenCharlie.cs
__device__ __forceinline__ float charlie_lambda_log(float c, bool rough) {
    float a, b, cc, d, e;
    if (!rough) { a = 25.3245f; b = 3.32435f; cc = 0.16801f; d = -1.27393f; e = -4.85967f; }
    else { a = 21.5473f; b = 3.82987f; cc = 0.19823f; d = -1.97760f; e = -4.32054f; }
    return a / (1.0f + b * powf(c, cc)) + d * c + e;
}
__device__ __forceinline__ float charlie_lambda(float c, float alpha) {
    c = clampf(c, 1e-4f, 1.0f);
    if (alpha <= 0.25f - 0.05f) return expf(charlie_lambda_log(c, false));
    if (alpha >= 0.25f + 0.05f) return expf(charlie_lambda_log(c, true));
    float t = (alpha - (0.25f - 0.05f)) / (2.0f * 0.05f);
    float lo = expf(charlie_lambda_log(c, false)), hi = expf(charlie_lambda_log(c, true));
    return lo + (hi - lo) * t;
}
__device__ __forceinline__ float charlie_brdf(float ndv, float ndl, float ndh, float alpha) {
    float den = 4.0f * fmaxf(ndv, 1e-6f) * fmaxf(ndl, 1e-6f);
    float a = fmaxf(alpha, 0.04f);
    float inv = 1.0f / a;
    float s2 = fmaxf(1.0f - ndh * ndh, 0.0f);
    float D = (2.0f + inv) * powf(s2, 0.5f * inv) / (2.0f * kPi);
    float G = 1.0f / ((1.0f + charlie_lambda(ndv, alpha)) * (1.0f + charlie_lambda(ndl, alpha)));
    return D * G / den;
}

// ── Disney BSDF (DisneyBsdf.cs) on a GMat record ─────────────────────────
struct DisneyP {
    f3 base, f0, ttint, sigma;
    float metallic, roughness, alpha, ax, ay, ior;
    float pD, pS, pT, pSh, pMs, pDt, pC, spec_trans, spec_trans_raw;
    float sheen, sheen_tint, sheen_rough, diff_trans;
    float clearcoat, cc_alpha, cc_f0, rot, alpha_iso, eavg;
    float tf_thick, tf_ior;   // thin film (thickness 0: plain Schlick Fresnel)
    int32_t flags;
};
__device__ __forceinline__ DisneyP load_disney(const GMat& g) {
    DisneyP p;
    p.flags = g.m0.y;
    p.spec_trans_raw = __int_as_float(g.m0.w);
    p.base = f3of(g.m1); p.metallic = g.m1.w;
    p.f0 = f3of(g.m2); p.roughness = g.m2.w;
    p.alpha = g.m3.x; p.ax = g.m3.y; p.ay = g.m3.z; p.ior = g.m3.w;
    p.pD = g.m4.x; p.pS = g.m4.y; p.pT = g.m4.z; p.pSh = g.m4.w;
    p.pMs = g.m5.x; p.pDt = g.m5.y; p.pC = g.m5.z; p.spec_trans = g.m5.w;
    p.sheen = g.m6.x; p.sheen_tint = g.m6.y; p.sheen_rough = g.m6.z; p.diff_trans = g.m6.w;
    p.clearcoat = g.m7.x; p.cc_alpha = g.m7.y; p.cc_f0 = g.m7.z; p.rot = g.m7.w;
    p.ttint = f3of(g.m8); p.alpha_iso = g.m8.w;
    p.sigma = f3of(g.m9); p.eavg = g.m9.w;
    p.tf_thick = g.m10.x; p.tf_ior = g.m10.y;
    return p;
}

constexpr float kDeltaAlpha = 2.5e-3f;
RT_HD float schlick_weight(float c) {
    float x = clampf(1.0f - c, 0.0f, 1.0f);
    float x2 = x * x;
    return x2 * x2 * x;
}

struct Frame {
    f3 t, b, n;
    RT_HD f3 to_local(f3 w) const { return mk(dot(w, t), dot(w, b), dot(w, n)); }
    RT_HD f3 to_world(f3 w) const { return w.x * t + w.y * b + w.z * n; }
};
// BuildShadingFrame (DisneyBsdf.cs:1825-1866)
__device__ __forceinline__ Frame shading_frame(const Surf& s, float rot) {
    f3 n = s.n, t, b;
    if (length_sq(s.tan) > 1e-10f && length_sq(s.bitan) > 1e-10f) {
        t = s.tan - dot(s.tan, n) * n;
        float tl = length_sq(t);
        if (tl > 1e-10f) {
            t = t / sqrtf(tl);
            b = cross(n, t);
        } else tangent_frame(n, t, b);
    } else tangent_frame(n, t, b);
    if (rot != 0.0f) {
        float ang = 2.0f * kPi * rot;
        float c = cosf(ang), sn = sinf(ang);
        f3 tr = c * t + sn * b;
        f3 br = -sn * t + c * b;
        t = tr;
        b = br;
    }
    return Frame{t, b, n};
}

__device__ __forceinline__ f3 ms_weight(const DisneyP& P, const float* E, float ndv, float ndl) {
    float eo = lut_e(E, ndv, P.alpha_iso), ei = lut_e(E, ndl, P.alpha_iso), ea = P.eavg;
    float om = fmaxf(1.0f - ea, 1e-6f);
    f3 favg = P.f0 + (splat(1.0f) - P.f0) * (1.0f / 21.0f);
    f3 den = splat(1.0f) - favg * (1.0f - ea);
    den = mk(fmaxf(den.x, 1e-6f), fmaxf(den.y, 1e-6f), fmaxf(den.z, 1e-6f));
    f3 fms = favg * favg * ea / den;
    return fms * ((1.0f - eo) * (1.0f - ei) / om);
}
// ThinFilm.Evaluate (ThinFilm.cs:40-136): Belcour-Barla 2017 iridescent
// Fresnel of a film (ior eta, thickness d nm) over a substrate of normal
// reflectance f0, seen from air at cos(theta1) = c1: Airy series of the two
// interfaces truncated after two harmonics, each projected to linear
// Rec.709 through the Gaussian fit of the CIE sensitivities (EvalSensitivity,
// ThinFilm.cs:150-186).
__device__ __forceinline__ float tf_schlick(float f0, float c) {
    float x = clampf(1.0f - c, 0.0f, 1.0f);
    float x5 = x * x;
    x5 *= x5 * x;
    return f0 + (1.0f - f0) * x5;
}
__device__ __forceinline__ f3 tf_sensitivity(float opd, f3 shift) {
    const float phase = 2.0f * kPi * opd * 1e-9f;
    const f3 val = mk(5.4856e-13f, 4.4201e-13f, 5.2481e-13f);
    const f3 pos = mk(1.6810e+06f, 1.7953e+06f, 2.2084e+06f);
    const f3 var = mk(4.3278e+09f, 9.3046e+09f, 6.6121e+09f);
    const f3 s2v = mk(sqrtf(2.0f * kPi * var.x), sqrtf(2.0f * kPi * var.y), sqrtf(2.0f * kPi * var.z));
    const f3 ct = mk(cosf(pos.x * phase + shift.x), cosf(pos.y * phase + shift.y), cosf(pos.z * phase + shift.z));
    const float ph2 = phase * phase;
    const f3 et = mk(expf(-ph2 * var.x), expf(-ph2 * var.y), expf(-ph2 * var.z));
    f3 xyz = val * s2v * ct * et;
    xyz.x += 9.7470e-14f * sqrtf(2.0f * kPi * 4.5282e+09f) * cosf(2.2399e+06f * phase + shift.x) * expf(-ph2 * 4.5282e+09f);
    xyz = xyz / 1.0685e-7f;
    return mk(3.2404542f * xyz.x - 1.5371385f * xyz.y - 0.4985314f * xyz.z,
              -0.9692660f * xyz.x + 1.8760108f * xyz.y + 0.0415560f * xyz.z,
              0.0556434f * xyz.x - 0.2040259f * xyz.y + 1.0572252f * xyz.z);
}
__device__ f3 thin_film(float cos1, float film_ior, float thick_nm, f3 f0) {
    const float outside = 1.0f;
    const float thickness = fmaxf(thick_nm, 0.0f);
    const float tt = clampf((thickness - 0.0f) / (30.0f - 0.0f), 0.0f, 1.0f);
    const float sm = tt * tt * (3.0f - 2.0f * tt);
    const float eta2 = fmaxf(outside + (film_ior - outside) * sm, 1.0001f);
    const float c1 = clampf(cos1, 0.0f, 1.0f);
    const float er = outside / eta2;
    const float s2sq = er * er * (1.0f - c1 * c1);
    if (s2sq >= 1.0f) return splat(1.0f);
    const float cos2 = sqrtf(1.0f - s2sq);
    const float q12 = (eta2 - outside) / (eta2 + outside);
    const float r12 = tf_schlick(q12 * q12, c1);
    const float t121 = 1.0f - r12;
    const float phi12 = eta2 < outside ? kPi : 0.0f;
    const float phi21 = kPi - phi12;
    const f3 bf = mk(clampf(f0.x, 0.0f, 0.9999f), clampf(f0.y, 0.0f, 0.9999f), clampf(f0.z, 0.0f, 0.9999f));
    const f3 sq = mk(sqrtf(bf.x), sqrtf(bf.y), sqrtf(bf.z));
    const f3 bior = (splat(1.0f) + sq) / (splat(1.0f) - sq);
    const f3 q23 = (bior - splat(eta2)) / (bior + splat(eta2));
    const f3 r023 = q23 * q23;
    const f3 r23 = mk(tf_schlick(r023.x, cos2), tf_schlick(r023.y, cos2), tf_schlick(r023.z, cos2));
    const f3 phi23 = mk(bior.x < eta2 ? kPi : 0.0f, bior.y < eta2 ? kPi : 0.0f, bior.z < eta2 ? kPi : 0.0f);
    const float opd = 2.0f * eta2 * thickness * cos2;
    const f3 phi = splat(phi21) + phi23;
    const f3 R = splat(r12) * r23;
    const f3 R123 = mk(clampf(R.x, 1e-5f, 0.9999f), clampf(R.y, 1e-5f, 0.9999f), clampf(R.z, 1e-5f, 0.9999f));
    const f3 r123 = mk(sqrtf(R123.x), sqrtf(R123.y), sqrtf(R123.z));
    const f3 Rs = (t121 * t121) * r23 / (splat(1.0f) - R123);
    f3 I = splat(r12) + Rs;
    f3 Cm = Rs - splat(t121);
    for (int m = 1; m <= 2; m++) {
        Cm *= r123;
        const f3 Sm = 2.0f * tf_sensitivity((float)m * opd, (float)m * phi);
        I += Cm * Sm;
    }
    return mk(fmaxf(I.x, 0.0f), fmaxf(I.y, 0.0f), fmaxf(I.z, 0.0f));
}
// EvalFresnel (DisneyBsdf.cs:1984-1989): Schlick, or the thin film
__device__ __forceinline__ f3 eval_fresnel(float c, f3 f0, const DisneyP& P) {
    if (P.tf_thick > 0.0f) return thin_film(c, P.tf_ior, P.tf_thick, f0);
    float w = schlick_weight(c);
    return f0 + (splat(1.0f) - f0) * w;
}
__device__ __forceinline__ f3 sheen_color(const DisneyP& P) {
    float lum = luminance(P.base);
    f3 tint = lum > 0.0f ? P.base / lum : splat(1.0f);
    return lerp3(splat(1.0f), tint, P.sheen_tint);
}

// EvaluateDirect (DisneyBsdf.cs:418-535)
__device__ f3 disney_eval_direct(const DisneyP& P, const float* E, f3 L, f3 V, const Surf& s) {
    f3 N = s.n;
    float ndl = fmaxf(dot(N, L), 0.0f);
    if (ndl <= 0.0f) return splat(0.0f);
    float ndv = fmaxf(dot(N, V), 0.001f);
    f3 H = normalize(L + V);
    float ndh = fmaxf(dot(N, H), 0.0f), vdh = fmaxf(dot(V, H), 0.0f), ldh = fmaxf(dot(L, H), 0.0f);
    float dw = (1.0f - P.metallic) * (1.0f - P.spec_trans) * (1.0f - P.diff_trans);
    f3 diffuse = splat(0.0f);
    if (dw > 0.0f) {
        float fd90 = 0.5f + 2.0f * P.roughness * ldh * ldh;
        float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
        float fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
        diffuse = P.base * fd * (dw * ndl / kPi);
    }
    f3 spec = splat(0.0f);
    if (P.alpha > kDeltaAlpha) {
        Frame fr = shading_frame(s, P.rot);
        f3 vl = fr.to_local(V), ll = fr.to_local(L), hl = fr.to_local(H);
        float D = d_ggx(hl, P.ax, P.ay);
        float G = g1_ggx(vl, P.ax, P.ay) * g1_ggx(ll, P.ax, P.ay);
        f3 F = eval_fresnel(vdh, P.f0, P);
        spec = D * G * F / fmaxf(4.0f * ndv, 1e-6f);
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
    f3 sheen = splat(0.0f);
    if (P.sheen > 0.0f && dw > 0.0f)
        sheen = P.sheen * charlie_brdf(ndv, ndl, ndh, P.sheen_rough) * sheen_color(P) * ndl;
    f3 ms = ms_weight(P, E, ndv, ndl) * (1.0f / kPi) * ndl;
    return diffuse + sheen + spec + ms + cc;
}

// Evaluate (DisneyBsdf.cs:649-771)
__device__ f3 disney_eval(const DisneyP& P, const float* E, f3 V, f3 L, const Surf& s) {
    f3 N = s.n;
    float ndl = dot(N, L), ndv = dot(N, V);
    if (ndv <= 0.0f) return splat(0.0f);
    if (ndl <= 0.0f) {
        float dall = (1.0f - P.metallic) * (1.0f - P.spec_trans);
        if (dall <= 0.0f || P.diff_trans <= 0.0f) return splat(0.0f);
        return P.base * (dall * P.diff_trans / kPi);
    }
    f3 hr = V + L;
    float hl2 = length_sq(hr);
    if (hl2 < 1e-14f) return splat(0.0f);
    f3 H = hr / sqrtf(hl2);
    float ndh = fmaxf(dot(N, H), 0.0f), vdh = fmaxf(dot(V, H), 0.0f), ldh = fmaxf(dot(L, H), 0.0f);
    float ds = (1.0f - P.metallic) * (1.0f - P.spec_trans) * (1.0f - P.diff_trans);
    f3 diffuse = splat(0.0f);
    if (ds > 0.0f) {
        float fd90 = 0.5f + 2.0f * P.roughness * ldh * ldh;
        float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
        float fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
        diffuse = P.base * fd * (ds / kPi);
    }
    f3 sheen = splat(0.0f);
    if (P.sheen > 0.0f && ds > 0.0f) sheen = P.sheen * charlie_brdf(ndv, ndl, ndh, P.sheen_rough) * sheen_color(P);
    Frame fr = shading_frame(s, P.rot);
    f3 vl = fr.to_local(V), ll = fr.to_local(L), hlc = fr.to_local(H);
    float D = d_ggx(hlc, P.ax, P.ay);
    float G = g1_ggx(vl, P.ax, P.ay) * g1_ggx(ll, P.ax, P.ay);
    f3 F = eval_fresnel(vdh, P.f0, P);
    f3 spec = D * G / fmaxf(4.0f * ndv * ndl, 1e-7f) * F;
    f3 cc = splat(0.0f);
    if (P.clearcoat > 0.0f) {
        float cndv = dot(N, V), cndl = dot(N, L);
        if (cndv > 0.0f && cndl > 0.0f) {
            float cndh = fmaxf(dot(N, H), 0.0f), cvdh = fmaxf(dot(V, H), 0.0f);
            float a2 = P.cc_alpha * P.cc_alpha;
            float den = cndh * cndh * (a2 - 1.0f) + 1.0f;
            float cD = a2 / (kPi * den * den);
            float cG = smith_g1(cndv, P.cc_alpha) * smith_g1(cndl, P.cc_alpha);
            float cF = P.cc_f0 + (1.0f - P.cc_f0) * schlick_weight(cvdh);
            cc = splat(P.clearcoat * 0.25f * cD * cG * cF / fmaxf(4.0f * cndv * cndl, 1e-7f));
        }
    }
    f3 ms = ms_weight(P, E, ndv, ndl) * (1.0f / kPi);
    return diffuse + sheen + spec + ms + cc;
}

// Pdf (DisneyBsdf.cs:783-870)
__device__ float disney_pdf(const DisneyP& P, f3 V, f3 L, const Surf& s) {
    f3 N = s.n;
    float ndl = dot(N, L), ndv = dot(N, V);
    if (ndv <= 0.0f) return 0.0f;
    float cos_pdf = ndl > 0.0f ? ndl / kPi : 0.0f;
    float dt_pdf = ndl < 0.0f ? -ndl / kPi : 0.0f;
    float spec_pdf = 0.0f, cc_pdf = 0.0f;
    f3 hr = V + L;
    float hl2 = length_sq(hr);
    if (hl2 >= 1e-14f) {
        f3 H = hr / sqrtf(hl2);
        float ndh = dot(N, H);
        if (ndh > 0.0f && P.alpha > kDeltaAlpha) {
            float sndv = fmaxf(ndv, 1e-7f);
            Frame fr = shading_frame(s, P.rot);
            f3 vl = fr.to_local(V), hlc = fr.to_local(H);
            spec_pdf = g1_ggx(vl, P.ax, P.ay) * d_ggx(hlc, P.ax, P.ay) / (4.0f * sndv);
        }
        if (P.cc_alpha > kDeltaAlpha) {
            float cndv = dot(N, V), cndh = dot(N, H);
            if (cndv > 0.0f && cndh > 0.0f) {
                float sc = fmaxf(cndv, 1e-7f);
                float a2 = P.cc_alpha * P.cc_alpha;
                float den = cndh * cndh * (a2 - 1.0f) + 1.0f;
                float cD = a2 / (kPi * den * den);
                cc_pdf = smith_g1(cndv, P.cc_alpha) * cD / (4.0f * sc);
            }
        }
    }
    return P.pD * cos_pdf + P.pSh * cos_pdf + P.pMs * cos_pdf + P.pS * spec_pdf + P.pC * cc_pdf + P.pDt * dt_pdf;
}

enum Lobe { kLDiffuse, kLSpecular, kLTransmission, kLSheen, kLMultiscatter, kLDiffTrans, kLClearcoat };

// ScatterInternal + lobe samplers (DisneyBsdf.cs:1176-1709)
__device__ bool disney_scatter(const DisneyP& P, const float* E, const Surf& s, f3 din, Rng& rng, f3& att, f3& dir,
                               int& lobe) {
    f3 N = s.n;
    f3 V = normalize(-din);
    float rnd = rng.next();
    bool res = true;
    if (rnd < P.pD) {
        lobe = kLDiffuse;
        f3 sd = N + random_unit_vector(rng);
        if (near_zero(sd)) sd = N;
        sd = normalize(sd);
        dir = sd;
        float ndv = fmaxf(dot(N, V), 0.001f), ndl = fmaxf(dot(N, sd), 0.001f);
        f3 hr = V + sd;
        float hl2 = length_sq(hr);
        f3 H = hl2 > 1e-7f ? hr / sqrtf(hl2) : N;
        float ldh = fmaxf(dot(sd, H), 0.0f);
        float fd90 = 0.5f + 2.0f * P.roughness * ldh * ldh;
        float fi = schlick_weight(ndv), fo = schlick_weight(ndl);
        float fd = (1.0f + (fd90 - 1.0f) * fi) * (1.0f + (fd90 - 1.0f) * fo);
        att = P.base * fd;
        att *= (1.0f - P.diff_trans);
        att = att / fmaxf(P.pD, 0.1f);
    } else if ((rnd -= P.pD) < P.pS) {
        lobe = kLSpecular;
        Frame fr = shading_frame(s, P.rot);
        f3 vl = fr.to_local(V);
        float u1 = rng.next();
        float u2 = rng.next();
        f3 H = fr.to_world(sample_vndf(vl, P.ax, P.ay, u1, u2));
        f3 L = reflect(-V, H);
        if (dot(L, N) <= 0.0f) {
            att = splat(0.0f);
            dir = N;
            res = false;
        } else {
            dir = L;
            float vdh = fmaxf(dot(V, H), 1e-4f);
            f3 F = eval_fresnel(vdh, P.f0, P);
            att = F * g1_ggx(fr.to_local(L), P.ax, P.ay);
            att = att / fmaxf(P.pS, 0.1f);
        }
    } else if ((rnd -= P.pS) < P.pT) {
        lobe = kLTransmission;
        bool thin = (P.flags & kMatThinWalled) != 0;
        float eta = thin ? (1.0f / P.ior) : (s.rel_eta > 0.0f ? s.rel_eta : (s.front ? (1.0f / P.ior) : P.ior));
        f3 ud = normalize(din);
        Frame fr = shading_frame(s, P.rot);
        bool rough = P.roughness > 0.01f;
        f3 Ht;
        if (rough) {
            f3 vl = fr.to_local(V);
            float u1 = rng.next();
            float u2 = rng.next();
            Ht = fr.to_world(sample_vndf(vl, P.ax, P.ay, u1, u2));
        } else Ht = N;
        if (dot(Ht, ud) > 0.0f) Ht = -Ht;
        float ct = fminf(dot(-ud, Ht), 1.0f);
        float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        bool cannot = !thin && eta * st > 1.0f;
        float frs = cannot ? 1.0f : fresnel_dielectric(ct, eta);
        f3 d;
        bool refl;
        if (frs > rng.next()) { d = reflect(ud, Ht); refl = true; }
        else if (thin) { d = ud; refl = false; }
        else { d = refract(ud, Ht, eta); refl = false; }
        if (near_zero(d)) { d = reflect(ud, N); refl = true; }
        dir = normalize(d);
        att = refl ? splat(1.0f) : P.ttint;
        if (rough) {
            f3 ll = fr.to_local(dir);
            ll.z = fabsf(ll.z);
            att *= g1_ggx(ll, P.ax, P.ay);
        }
        att = att / fmaxf(P.pT, 0.1f);
    } else if ((rnd -= P.pT) < P.pSh) {
        lobe = kLSheen;
        f3 sd = N + random_unit_vector(rng);
        if (near_zero(sd)) sd = N;
        sd = normalize(sd);
        dir = sd;
        float ndv = fmaxf(dot(N, V), 1e-4f), ndl = fmaxf(dot(N, sd), 1e-4f);
        f3 hr = V + sd;
        float hl2 = length_sq(hr);
        f3 H = hl2 > 1e-7f ? hr / sqrtf(hl2) : N;
        float ndh = fmaxf(dot(N, H), 0.0f);
        att = P.sheen * kPi * charlie_brdf(ndv, ndl, ndh, P.sheen_rough) * sheen_color(P);
        att = att / fmaxf(P.pSh, 0.1f);
    } else if ((rnd -= P.pSh) < P.pMs) {
        lobe = kLMultiscatter;
        f3 sd = N + random_unit_vector(rng);
        if (near_zero(sd)) sd = N;
        sd = normalize(sd);
        dir = sd;
        float ndv = fmaxf(dot(N, V), 1e-3f), ndl = fmaxf(dot(N, sd), 1e-3f);
        att = ms_weight(P, E, ndv, ndl) / fmaxf(P.pMs, 0.05f);
    } else if ((rnd -= P.pMs) < P.pDt) {
        lobe = kLDiffTrans;
        f3 sd = -N + random_unit_vector(rng);
        if (near_zero(sd)) sd = -N;
        dir = normalize(sd);
        att = P.base * P.diff_trans;
        att = att / fmaxf(P.pDt, 0.05f);
    } else {
        lobe = kLClearcoat;
        Frame fr;
        fr.n = N;
        tangent_frame(N, fr.t, fr.b);
        if (dot(N, V) <= 0.0f) {
            att = splat(0.0f);
            dir = N;
            res = false;
        } else {
            f3 vl = fr.to_local(V);
            float u1 = rng.next();
            float u2 = rng.next();
            f3 H = fr.to_world(sample_vndf(vl, P.cc_alpha, P.cc_alpha, u1, u2));
            f3 L = reflect(-V, H);
            if (dot(L, N) <= 0.0f) {
                att = splat(0.0f);
                dir = N;
                res = false;
            } else {
                dir = L;
                float cndl = fmaxf(dot(N, L), 1e-4f), cvdh = fmaxf(dot(V, H), 1e-4f);
                float fres = P.cc_f0 + (1.0f - P.cc_f0) * schlick_weight(cvdh);
                att = splat(fres * smith_g1(cndl, P.cc_alpha));
                att = att / fmaxf(P.pC, 0.1f);
            }
        }
    }
    if (isnan(att.x) || isinf(att.x) || isnan(att.y) || isinf(att.y) || isnan(att.z) || isinf(att.z)) att = P.base;
    return res;
}

struct BsdfS {
    f3 wo, f;
    float pdf;
    bool delta;
    bool caster;      // BsdfSample.CausticCaster (followed by the photon walk)
    bool has_next_abs;
    f3 next_abs;
    int transition;   // 0 none, 1 enter, 2 exit
    int lobe;         // Disney lobe that produced the sample (Lobe)
};

// Sample (DisneyBsdf.cs:893-997)
__device__ bool disney_sample(const DisneyP& P, const float* E, f3 V, const Surf& s, Rng& rng, BsdfS& o) {
    f3 att, wo;
    int lobe;
    if (!disney_scatter(P, E, s, -V, rng, att, wo, lobe)) return false;
    float ndw = dot(s.n, wo);
    o.wo = wo;
    o.lobe = lobe;
    o.has_next_abs = false;
    o.next_abs = splat(0.0f);
    o.transition = 0;
    o.caster = true;
    if (lobe == kLDiffTrans) {
        f3 fb = disney_eval(P, E, V, wo, s);
        float pb = disney_pdf(P, V, wo, s);
        if (pb <= 0.0f) { o.f = att; o.pdf = 1.0f; o.delta = true; return true; }
        o.f = fb; o.pdf = pb; o.delta = false;
        return true;
    }
    if (lobe == kLTransmission) {
        o.f = att;
        o.pdf = 1.0f;
        o.delta = true;
        if (ndw < 0.0f && !(P.flags & kMatThinWalled)) {
            bool has = P.sigma.x > 0.0f || P.sigma.y > 0.0f || P.sigma.z > 0.0f;
            if (s.front && has) { o.has_next_abs = true; o.next_abs = P.sigma; }
            else if (!s.front) { o.has_next_abs = true; o.next_abs = splat(0.0f); }
            o.transition = s.front ? 1 : 2;
        }
        return true;
    }
    if (ndw <= 0.0f) return false;
    if ((lobe == kLSpecular && P.alpha <= kDeltaAlpha) || (lobe == kLClearcoat && P.cc_alpha <= kDeltaAlpha)) {
        o.f = att;
        o.pdf = 1.0f;
        o.delta = true;
        // only a mirror-like (metallic >= 0.5) delta reflection casts caustics
        o.caster = P.metallic >= 0.5f;
        return true;
    }
    o.f = disney_eval(P, E, V, wo, s);
    o.pdf = disney_pdf(P, V, wo, s);
    o.delta = false;
    return true;
}

// Dielectric.Scatter (Dielectric.cs:75-105): Schlick-weighted reflect /
// refract with the relative IOR (IorStack when set, else the material's).
__device__ __forceinline__ void dielectric_scatter(const GMat& g, const Surf& s, f3 d, Rng& rng, f3& att, f3& wo,
                                                   f3& origin) {
    float ior = g.m3.w;
    float ri = s.rel_eta > 0.0f ? s.rel_eta : (s.front ? (1.0f / ior) : ior);
    f3 ud = normalize(d);
    float ct = fminf(dot(-ud, s.n), 1.0f);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    bool cannot = ri * st > 1.0f;
    f3 dir;
    if (cannot || schlick(ct, ri) > rng.next()) dir = reflect(ud, s.n);
    else dir = refract(ud, s.n, ri);
    f3 on = dot(dir, s.n) >= 0.0f ? s.n : -s.n;
    origin = offset_origin(s.p, on);
    wo = dir;
    att = f3of(g.m1);
}

// ShadowTransmittance / ShadowAbsorption of any material kind
// (DisneyBsdf.cs:564-621, Dielectric.cs:46-61, IMaterial defaults)
__device__ __forceinline__ f3 mat_shadow_tr(const GMat& g, f3 wi, f3 n) {
    int kind = g.m0.x;
    if (kind == RT_MAT_DIELECTRIC) {
        float ct = fminf(fabsf(dot(wi, n)), 1.0f);
        float fr = fresnel_dielectric(ct, 1.0f / g.m3.w);
        return (1.0f - fr) * f3of(g.m1);
    }
    if (kind == RT_MAT_DISNEY) {
        float st = __int_as_float(g.m0.w);
        if (st <= 0.0f) return splat(0.0f);
        float ct = fminf(fabsf(dot(wi, n)), 1.0f);
        float fr = fresnel_dielectric(ct, 1.0f / g.m3.w);
        return st * (1.0f - fr) * f3of(g.m8);
    }
    return splat(0.0f);
}
__device__ __forceinline__ f3 mat_shadow_abs(const GMat& g) {
    if (g.m0.x == RT_MAT_DISNEY) return f3of(g.m9);
    return splat(0.0f);
}

}  // namespace rt
