// material.h — Disney parameter derivation and the texture set, shared by
// the host scene builder (per-material constants) and the shading kernels
// (per-hit constants of textured materials).
#pragma once
#include "../../include/rt3d.h"
#include "rt_internal.h"

namespace rt {

// EnergyCompensationLut.SampleEavg (EnergyCompensationLut.cs:162-171)
RT_HD float eavg_lookup(const float* Eavg, float alpha) {
    alpha = clampf(alpha, 0.001f, 1.0f);
    float fal = alpha * 32.0f - 0.5f;
    int a0 = (int)floorf(fal);
    a0 = a0 < 0 ? 0 : (a0 > 31 ? 31 : a0);
    int a1 = a0 + 1 > 31 ? 31 : a0 + 1;
    float ta = clampf(fal - (float)a0, 0.0f, 1.0f);
    return Eavg[a0] + (Eavg[a1] - Eavg[a0]) * ta;
}

// Disney ShadingParams, ComputeF0, ComputeLobeWeights and ResolveTransmission
// (DisneyBsdf.cs:213-293, 1016-1049, 1097-1139, 1943-1959) for material m
// with base colour c, metallic and roughness given separately (the material's
// own knobs, or a texture lookup).  m0 = (RT_MAT_DISNEY, flags, -1, raw
// spec_trans bits).
RT_HD GMat disney_gmat(const rt_material& m, f3 c, float metallic_in, float roughness_in, const float* Eavg) {
    GMat g{};
    int32_t flags = 0;
    float metallic = clampf(metallic_in, 0, 1), roughness = clampf(roughness_in, 0, 1);
    float specular = clampf(m.specular, 0, 2), spec_tint = clampf(m.specular_tint, 0, 1);
    float sheen = clampf(m.sheen, 0, 1), sheen_tint = clampf(m.sheen_tint, 0, 1);
    float sheen_rough = clampf(m.sheen_roughness, 0.04f, 1);
    float clearcoat = clampf(m.clearcoat, 0, 1), cc_gloss = clampf(m.clearcoat_gloss, 0, 1);
    float spec_trans = clampf(m.spec_trans, 0, 1), ior = mx(m.ior, 1.0001f);
    float aniso = clampf(m.anisotropic, 0, 1);
    float rot = m.anisotropic_rotation - floorf(m.anisotropic_rotation);
    float diff_trans = clampf(m.diff_trans, 0, 1);
    float alpha = mx(roughness * roughness, 0.001f);
    float aspect = sqrtf(1.0f - 0.9f * aniso);
    float ax = mx(alpha / aspect, 0.001f), ay = mx(alpha * aspect, 0.001f);
    float cr = (m.flags & RT_MATF_COAT_ROUGHNESS) ? m.coat_roughness : -1.0f;
    float cc_alpha = cr >= 0.0f ? mx(cr * cr, 0.001f) : lerpf(0.1f, 0.001f, cc_gloss);
    float ec = mx(m.coat_ior, 1.0001f);
    float r0 = (ec - 1.0f) / (ec + 1.0f);
    float cc_f0 = r0 * r0;
    // ComputeF0
    float r = (ior - 1.0f) / (ior + 1.0f);
    float scaled = r * r * 2.0f * specular;
    float lum = luminance(c);
    f3 tint = lum > 0.0f ? c / lum : splat(1.0f);
    f3 diel = lerp3(splat(scaled), scaled * tint, spec_tint);
    f3 f0 = lerp3(diel, c, metallic);
    // ComputeLobeWeights
    float dall = (1.0f - metallic) * (1.0f - spec_trans);
    float dw = dall * (1.0f - diff_trans), dtw = dall * diff_trans;
    float sf0;
    if (metallic > 0.5f) sf0 = luminance(c);
    else sf0 = r * r * 2.0f * specular;
    float sfloor = 0.1f * (1.0f - spec_trans * 0.9f);
    float sw = mx(sfloor, lerpf(sf0, 1.0f, metallic));
    float tw = (1.0f - metallic) * spec_trans;
    float cw = clearcoat * 0.04f;
    float shw = sheen * 0.25f * dall;
    float alpha_iso = sqrtf(ax * ay);
    float eavg = eavg_lookup(Eavg, alpha_iso);
    float msw = sw * (1.0f - eavg);
    float sum = dw + sw + tw + cw + shw + msw + dtw;
    float tot = sum < 1e-6f ? 1.0f : sum;
    // ResolveTransmission
    f3 ttint, sigma = splat(0.0f);
    bool has_tc = (m.flags & RT_MATF_TRANSMISSION_COLOR) != 0;
    bool thin = (m.flags & RT_MATF_THIN_WALLED) != 0;
    if (!has_tc) ttint = mk(sqrtf(mx(c.x, 0.0f)), sqrtf(mx(c.y, 0.0f)), sqrtf(mx(c.z, 0.0f)));
    else {
        f3 tc = mk(m.transmission_color[0], m.transmission_color[1], m.transmission_color[2]);
        float td = mx(m.transmission_depth, 0.0f);
        if (td <= 0.0f) ttint = tc;
        else {
            f3 cc = mk(mx(tc.x, 1e-6f), mx(tc.y, 1e-6f), mx(tc.z, 1e-6f));
            float inv = 1.0f / td;
            sigma = mk(-logf(cc.x) * inv, -logf(cc.y) * inv, -logf(cc.z) * inv);
            ttint = splat(1.0f);
        }
    }
    if (thin) flags |= kMatThinWalled;
    if (has_tc) flags |= kMatHasTransColor;
    // IsSpecularTransmissive / TryGetDielectricIor read the raw knobs
    if (!thin && m.spec_trans >= 0.5f && m.roughness <= 0.05f) flags |= kMatSpecTransmissive;
    if (!thin && m.spec_trans > 0.0f) flags |= kMatDielectricIor;
    if (m.flags & RT_MATF_TEXTURED) flags |= kMatTextured;
    if (m.flags & RT_MATF_IMAGE) flags |= kMatImage;
    g.m1 = make_float4(c.x, c.y, c.z, metallic);
    g.m2 = make_float4(f0.x, f0.y, f0.z, roughness);
    g.m3 = make_float4(alpha, ax, ay, ior);
    g.m4 = make_float4(dw / tot, sw / tot, tw / tot, shw / tot);
    g.m5 = make_float4(msw / tot, dtw / tot, cw / tot, spec_trans);
    g.m6 = make_float4(sheen, sheen_tint, sheen_rough, diff_trans);
    g.m7 = make_float4(clearcoat, cc_alpha, cc_f0, rot);
    g.m8 = make_float4(ttint.x, ttint.y, ttint.z, alpha_iso);
    g.m9 = make_float4(sigma.x, sigma.y, sigma.z, eavg);
    g.m10 = make_float4(mx(m.thin_film_thickness, 0.0f), mx(m.thin_film_ior, 1.0001f), 0.0f, 0.0f);
    // raw spec_trans for ShadowTransmittance (DisneyBsdf.cs:570-584)
    g.m0 = make_int4(RT_MAT_DISNEY, flags, -1, __builtin_bit_cast(int32_t, m.spec_trans));
    return g;
}

// ── texture set (no reference counterpart) ─────────────────────────────────
// Planar fp32 [kTexChannels][height][width]: base colour r, g, b, roughness,
// metallic.  Bilinear filtering with texel centres at (i + 0.5) / width;
// u wraps (the seam of spherical and cylindrical maps), v clamps.
constexpr int kTexChannels = 5;

struct TexFoot {
    int idx[4];     // texel index within a plane: (i0,j0) (i1,j0) (i0,j1) (i1,j1)
    float w[4];     // bilinear weights, summing to 1
};

RT_HD TexFoot tex_footprint(int W, int H, float u, float v) {
    float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    // wrap in u with the float floor (no int overflow for large u)
    float wx = fx0 - (float)W * floorf(fx0 / (float)W);
    int i0 = (int)wx;
    i0 = i0 < 0 ? 0 : (i0 >= W ? W - 1 : i0);
    int i1 = i0 + 1 >= W ? 0 : i0 + 1;
    int j0 = fy0 < 0.0f ? 0 : (fy0 > (float)(H - 1) ? H - 1 : (int)fy0);
    int j1 = fy0 + 1.0f < 0.0f ? 0 : (fy0 + 1.0f > (float)(H - 1) ? H - 1 : (int)fy0 + 1);
    TexFoot t;
    t.idx[0] = j0 * W + i0;
    t.idx[1] = j0 * W + i1;
    t.idx[2] = j1 * W + i0;
    t.idx[3] = j1 * W + i1;
    t.w[0] = (1.0f - fx) * (1.0f - fy);
    t.w[1] = fx * (1.0f - fy);
    t.w[2] = (1.0f - fx) * fy;
    t.w[3] = fx * fy;
    return t;
}

// Filtered value of channel ch.
RT_HD float tex_value(const float* tex, int W, int H, const TexFoot& t, int ch) {
    const float* pl = tex + (int64_t)ch * W * H;
    return ((t.w[0] * pl[t.idx[0]] + t.w[1] * pl[t.idx[1]]) + t.w[2] * pl[t.idx[2]]) + t.w[3] * pl[t.idx[3]];
}

}  // namespace rt
