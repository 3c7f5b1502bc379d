"""Disney / Lambert / Dielectric restatement in the oracle, pinned by the
reference's own BSDF invariants (DisneyBsdfTests.cs, MisMaterialsTests.cs)
and closed forms."""
import numpy as np
import pytest

from raytracer import abi

UP = np.array([0, 1, 0], np.float32)


def disney(metallic=0.0, roughness=0.5, specular=0.5, sheen=0.0, clearcoat=0.0, clearcoat_gloss=1.0,
           base=(0.7, 0.5, 0.3), spec_trans=0.0, ior=1.5, **kw):
    m = abi.material_defaults(abi.RT_MAT_DISNEY)
    m.color[:] = base
    m.metallic, m.roughness, m.specular, m.specular_tint = metallic, roughness, specular, 0.0
    m.sheen, m.sheen_tint, m.clearcoat, m.clearcoat_gloss = sheen, 0.5, clearcoat, clearcoat_gloss
    m.spec_trans, m.ior = spec_trans, ior
    for k, v in kw.items():
        if k == "transmission_color":
            m.flags |= abi.RT_MATF_TRANSMISSION_COLOR
            m.transmission_color[:] = v
        elif k == "thin_walled":
            m.flags |= abi.RT_MATF_THIN_WALLED if v else 0
        else:
            setattr(m, k, v)
    return m


OPAQUE = [(0, 1.0, 0.5, 0, 0, 1), (0, 0.25, 0.5, 0, 0, 1), (0, 0.05, 0.5, 0, 0, 1), (1, 0.15, 0, 0, 0, 1),
          (1, 0.6, 0, 0, 0, 1), (0, 0.4, 0.5, 0.6, 0, 1), (0, 0.3, 0.5, 0, 1, 0.9), (0, 0.8, 0.5, 0.3, 0.5, 0.5)]
BROAD = [(0, 1.0, 0.5, 0, 0, 1), (0, 0.8, 0.5, 0, 0, 1), (1, 0.6, 0, 0, 0, 1), (0, 0.5, 0.5, 0.6, 0, 1)]


def sphere_dirs(n, seed):
    rng = np.random.default_rng(seed)
    z = 2 * rng.random(n) - 1
    r = np.sqrt(np.maximum(0, 1 - z * z))
    phi = 2 * np.pi * rng.random(n)
    return np.stack([r * np.cos(phi), z, r * np.sin(phi)], 1).astype(np.float32)


def hemi_dirs(n, seed):
    d = sphere_dirs(n, seed)
    d[d[:, 1] < 0] *= -1
    return d


@pytest.mark.parametrize("p", OPAQUE)
def test_evaluate_is_reciprocal(oracle, p):
    m = disney(*p)
    v, l = hemi_dirs(256, 1), hemi_dirs(256, 2)
    fvl, _, _ = oracle.bsdf_eval_batch(m, UP, v, l)
    flv, _, _ = oracle.bsdf_eval_batch(m, UP, l, v)
    assert np.all(np.abs(fvl - flv) <= 1e-6 + 1e-4 * np.maximum(np.abs(fvl), np.abs(flv)))


@pytest.mark.parametrize("p", BROAD)
def test_pdf_integrates_to_one(oracle, p):
    m = disney(*p)
    view = np.array([0.3, 0.9, 0.2], np.float32)
    view /= np.linalg.norm(view)
    l = sphere_dirs(32768, 3)
    _, pdf, _ = oracle.bsdf_eval_batch(m, UP, np.repeat(view[None], len(l), 0), l)
    integral = 4 * np.pi * pdf.astype(np.float64).mean()
    assert 0.92 <= integral <= 1.08


@pytest.mark.parametrize("p", OPAQUE)
def test_sample_f_and_pdf_match_evaluate_and_pdf(oracle, p):
    m = disney(*p)
    view = np.array([0.2, 1.0, 0.3], np.float32)
    view /= np.linalg.norm(view)
    wo, f, pdf, flags = oracle.bsdf_sample_batch(m, UP, view, 256, seed=5)
    ok = (flags & 1) != 0
    nd = ok & ((flags & 2) == 0)
    assert nd.sum() > 100
    fe, pe, _ = oracle.bsdf_eval_batch(m, UP, np.repeat(view[None], nd.sum(), 0), wo[nd])
    assert np.all(np.abs(f[nd] - fe) <= 1e-6 + 1e-3 * np.abs(fe))
    assert np.all((pdf[nd] >= pe * 0.999 - 1e-6) & (pdf[nd] <= pe * 1.001 + 1e-6))


def test_diffuse_furnace(oracle):
    m = disney(0, 1.0, 0.0, base=(0.9, 0.9, 0.9))
    l = sphere_dirs(32768, 1234)
    up = l[:, 1] > 0
    f, _, _ = oracle.bsdf_eval_batch(m, UP, np.repeat(UP[None], len(l), 0), l)
    refl = 4 * np.pi * (f[up] * l[up, 1:2]).astype(np.float64).sum(0) / len(l)
    assert up.sum() > len(l) / 3
    assert np.all((refl >= 0.85) & (refl <= 0.95))


def test_multiscatter_rough_white_metal(oracle):
    m = disney(1, 0.9, 0.5, base=(1, 1, 1))
    l = sphere_dirs(131072, 2026)
    up = l[:, 1] > 0
    f, _, _ = oracle.bsdf_eval_batch(m, UP, np.repeat(UP[None], len(l), 0), l)
    refl = 4 * np.pi * (f[up, 0] * l[up, 1]).astype(np.float64).sum() / len(l)
    assert 0.90 < refl <= 1.05


def test_multiscatter_gold_keeps_hue(oracle):
    m = disney(1, 0.9, 0.5, base=(1.0, 0.71, 0.29))
    l = sphere_dirs(131072, 77)
    up = l[:, 1] > 0
    f, _, _ = oracle.bsdf_eval_batch(m, UP, np.repeat(UP[None], len(l), 0), l)
    refl = 4 * np.pi * (f[up] * l[up, 1:2]).astype(np.float64).sum(0) / len(l)
    assert refl[0] > 0.85 and refl[0] / refl[2] > 2.0


def test_energy_lut_properties(oracle):
    E, Eavg = oracle.energy_lut()
    assert np.all((E >= 0) & (E <= 1)) and np.all((Eavg >= 0) & (Eavg <= 1))
    # smooth limit: single-scatter GGX loses (almost) no energy at small alpha
    # away from grazing view angles
    assert np.all(E[4:, 0] > 0.98)
    # energy loss grows with roughness (E_avg decreasing overall)
    assert Eavg[-1] < Eavg[0] - 0.2
    assert np.all(np.diff(Eavg) <= 1e-2)


def test_lambert_closed_forms(oracle):
    m = abi.material_defaults(abi.RT_MAT_LAMBERTIAN)
    m.color[:] = (0.5, 0.25, 0.125)
    l = hemi_dirs(64, 9)
    v = hemi_dirs(64, 10)
    f, pdf, fd = oracle.bsdf_eval_batch(m, UP, v, l)
    assert np.allclose(f, np.array([0.5, 0.25, 0.125]) / np.pi, rtol=1e-6)
    assert np.allclose(pdf, l[:, 1] / np.pi, rtol=1e-6)
    assert np.allclose(fd, np.array([0.5, 0.25, 0.125]) * (l[:, 1:2] / np.pi), rtol=1e-5)
    wo, fs, ps, flags = oracle.bsdf_sample_batch(m, UP, v[0], 4096, seed=3)
    ok = (flags & 1) != 0
    # cosine-weighted: E[cos] = 2/3
    assert abs(wo[ok, 1].mean() - 2.0 / 3.0) < 0.02
    assert np.allclose(ps[ok], wo[ok, 1] / np.pi, rtol=1e-5)


def test_beer_lambert_sigma(oracle):
    # σa = −ln(C)/D (DisneyBsdf.cs:1037-1048)
    m = disney(0, 0.0, 0.5, base=(1, 1, 1), spec_trans=1.0, transmission_color=(0.5, 0.25, 0.8),
               transmission_depth=2.0)
    tr, sig = oracle.shadow_tr(m, UP, np.array([0, -1, 0], np.float32))
    assert np.allclose(sig, -np.log([0.5, 0.25, 0.8]) / 2.0, rtol=1e-5)
    # normal incidence straight-through transmittance (1 - F0) with F0 = ((n-1)/(n+1))^2
    f0 = ((1.5 - 1) / (1.5 + 1)) ** 2
    assert np.allclose(tr, 1 - f0, rtol=1e-5)


def test_near_delta_glass_direct_and_pdf_negligible(oracle):
    m = disney(0, 0.0, 0.5, base=(1, 1, 1), spec_trans=1.0)
    v = hemi_dirs(256, 4)
    l = hemi_dirs(256, 5)
    _, pdf, fd = oracle.bsdf_eval_batch(m, UP, v, l)
    assert np.all(pdf < 1e-4 + 1e-3)   # only the multiscatter sliver remains
    assert np.all(fd < 1e-3)


def test_transmission_samples_carry_medium_switch(oracle):
    m = disney(0, 0.0, 0.5, base=(1, 1, 1), spec_trans=1.0, transmission_color=(0.5, 0.5, 0.5),
               transmission_depth=1.0)
    view = np.array([0.2, 0.9, 0.1], np.float32)
    view /= np.linalg.norm(view)
    wo, f, pdf, flags = oracle.bsdf_sample_batch(m, UP, view, 512, seed=8)
    refr = ((flags & 1) != 0) & (wo[:, 1] < 0)
    assert refr.sum() > 0
    assert np.all((flags[refr] & 2) != 0)          # delta
    assert np.all((flags[refr] & 4) != 0)          # next-segment absorption
    assert np.all(((flags[refr] >> 4) & 3) == 1)   # enter


def test_dielectric_shadow_transmittance(oracle):
    m = abi.material_defaults(abi.RT_MAT_DIELECTRIC)
    m.color[:] = (1, 1, 1)
    m.ior = 1.5
    tr, _ = oracle.shadow_tr(m, UP, np.array([0, -1, 0], np.float32))
    assert np.allclose(tr, 1 - 0.04, rtol=1e-5)
