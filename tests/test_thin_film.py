"""Thin-film Disney Fresnel (ThinFilm.Evaluate, ThinFilm.cs:40-136, used by
DisneyBsdf.EvalFresnel, DisneyBsdf.cs:1984-1989) in the oracle.

Pinned by
  * an independent float64 restatement of the published model (Belcour &
    Barla 2017, "A practical extension to microfacet theory for the modeling
    of varying iridescence"; the KHR_materials_iridescence formulation with
    its Gaussian fit of the CIE sensitivities) on random inputs;
  * the incoherent limit: once the optical path difference is long, the
    Gaussian envelopes exp(-phase^2 var) kill both interference harmonics
    and the reflectance is the incoherent two-interface sum
    R12 + T12^2 R23 / (1 - R12 R23) (R123 clamped as the model clamps it);
  * the Disney BSDF invariants (reciprocity, Sample <-> Evaluate/Pdf) with a
    film on the specular lobe.
"""
import numpy as np
import pytest

from test_materials import UP, disney, hemi_dirs


def _schlick(f0, c):
    x = np.clip(1.0 - c, 0.0, 1.0)
    return f0 + (1.0 - f0) * x ** 5


def _sens(opd, shift):
    """Belcour-Barla spectral sensitivity, XYZ Gaussian fit, to linear Rec.709."""
    phase = 2.0 * np.pi * opd * 1e-9
    val = np.array([5.4856e-13, 4.4201e-13, 5.2481e-13])
    pos = np.array([1.6810e6, 1.7953e6, 2.2084e6])
    var = np.array([4.3278e9, 9.3046e9, 6.6121e9])
    xyz = val * np.sqrt(2 * np.pi * var) * np.cos(pos * phase + shift) * np.exp(-phase ** 2 * var)
    xyz[0] += 9.7470e-14 * np.sqrt(2 * np.pi * 4.5282e9) * np.cos(2.2399e6 * phase + shift[0]) * np.exp(
        -phase ** 2 * 4.5282e9)
    xyz /= 1.0685e-7
    M = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]])
    return M @ xyz


def model(cos1, film_ior, thick, f0):
    """Iridescent Fresnel in float64 (air above the film)."""
    d = max(thick, 0.0)
    t = np.clip(d / 30.0, 0.0, 1.0)
    eta2 = max(1.0 + (film_ior - 1.0) * t * t * (3 - 2 * t), 1.0001)
    c1 = np.clip(cos1, 0.0, 1.0)
    s2 = (1.0 / eta2) ** 2 * (1 - c1 * c1)
    if s2 >= 1:
        return np.ones(3)
    c2 = np.sqrt(1 - s2)
    r12 = _schlick(((eta2 - 1) / (eta2 + 1)) ** 2, c1)
    t121 = 1 - r12
    f0c = np.clip(np.asarray(f0, float), 0.0, 0.9999)
    bior = (1 + np.sqrt(f0c)) / (1 - np.sqrt(f0c))
    r23 = _schlick(((bior - eta2) / (bior + eta2)) ** 2, c2)
    phi = np.pi + np.where(bior < eta2, np.pi, 0.0)
    opd = 2 * eta2 * d * c2
    R123 = np.clip(r12 * r23, 1e-5, 0.9999)
    Rs = t121 ** 2 * r23 / (1 - R123)
    I = r12 + Rs
    Cm = Rs - t121
    for m in (1, 2):
        Cm = Cm * np.sqrt(R123)
        I = I + Cm * 2 * _sens(m * opd, m * phi)
    return np.maximum(I, 0.0)


def test_matches_published_model(oracle):
    rng = np.random.default_rng(11)
    n = 4000
    args = np.stack([rng.uniform(0, 1, n), rng.uniform(1.0, 2.4, n), rng.uniform(0, 1200, n),
                     rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n)], 1).astype(np.float32)
    args[:50, 2] = 0.0                 # no film
    args[50:100, 2] = rng.uniform(0, 30, 50).astype(np.float32)   # inside the ior blend
    got = oracle.thin_film(args)
    want = np.array([model(*a[:3].astype(float), a[3:].astype(float)) for a in args])
    err = np.abs(got - want)
    print("max |oracle - model|", err.max())
    assert np.all(got >= 0.0)
    assert np.all(err <= 2e-4 + 2e-4 * np.abs(want)), err.max()


def test_incoherent_limit_at_large_thickness(oracle):
    rng = np.random.default_rng(5)
    for _ in range(200):
        c, ior, f0 = rng.uniform(0.05, 1.0), rng.uniform(1.2, 2.2), rng.uniform(0.02, 0.9, 3)
        got = oracle.thin_film(np.array([[c, ior, 20000.0, *f0]], np.float32))[0]
        c2 = np.sqrt(1 - (1 - c * c) / ior ** 2)
        r12 = _schlick(((ior - 1) / (ior + 1)) ** 2, c)
        bior = (1 + np.sqrt(f0)) / (1 - np.sqrt(f0))
        r23 = _schlick(((bior - ior) / (bior + ior)) ** 2, c2)
        R = np.clip(r12 * r23, 1e-5, 0.9999)
        incoherent = r12 + (1 - r12) ** 2 * r23 / (1 - R)
        assert np.allclose(got, incoherent, rtol=1e-4, atol=1e-6), (got, incoherent)


FILMS = [dict(thin_film_thickness=380.0, thin_film_ior=1.33), dict(thin_film_thickness=650.0, thin_film_ior=1.8)]


@pytest.mark.parametrize("film", FILMS)
@pytest.mark.parametrize("metallic,roughness", [(1.0, 0.25), (0.0, 0.4)])
def test_thin_film_disney_invariants(oracle, film, metallic, roughness):
    m = disney(metallic, roughness, 0.5, **film)
    plain = disney(metallic, roughness, 0.5)
    v, l = hemi_dirs(256, 21), hemi_dirs(256, 22)
    fvl, _, _ = oracle.bsdf_eval_batch(m, UP, v, l)
    flv, _, _ = oracle.bsdf_eval_batch(m, UP, l, v)
    assert np.all(np.abs(fvl - flv) <= 1e-6 + 1e-4 * np.maximum(np.abs(fvl), np.abs(flv)))
    fp, _, _ = oracle.bsdf_eval_batch(plain, UP, v, l)
    assert np.abs(fvl - fp).max() > 1e-3          # the film changes the specular colour
    view = np.array([0.2, 1.0, 0.3], np.float32)
    view /= np.linalg.norm(view)
    wo, f, pdf, flags = oracle.bsdf_sample_batch(m, UP, view, 256, seed=9)
    nd = ((flags & 1) != 0) & ((flags & 2) == 0)
    fe, pe, _ = oracle.bsdf_eval_batch(m, UP, np.repeat(view[None], nd.sum(), 0), wo[nd])
    assert np.all(np.abs(f[nd] - fe) <= 1e-6 + 1e-3 * np.abs(fe))
    assert np.all((pdf[nd] >= pe * 0.999 - 1e-6) & (pdf[nd] <= pe * 1.001 + 1e-6))
