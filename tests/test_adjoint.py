"""CPU checks of the texture-gradient oracle (FP64 adjoint of the render
recursion) and of the texture lookup it shares with the renderer.

  * the bilinear footprint (texel centres at (i + 0.5) / width, u wraps, v
    clamps) equals an independent numpy restatement;
  * a constant texture equal to a material's own knobs renders the same image
    as the untextured material;
  * the adjoint is linear in dLoss/dBeauty;
  * against central finite differences of the oracle's forward render with
    common random numbers.  At metallic <= 0.5 the lobe probabilities, pdfs,
    MIS weights and sampled directions do not depend on the base colour, and
    with Russian roulette off the estimator is a smooth function of it, so
    the FD of the loss equals the detached-sampling gradient up to O(h^2).
"""
import numpy as np
import pytest

from raytracer import abi, scenes

W = H = 32
SPP, DEPTH = 4, 3


def _settings(w=W, h=H, spp=SPP, depth=DEPTH):
    st = abi.settings_defaults(w, h, spp, depth)
    st.max_sample_radiance = 3.0e38      # clamps inactive (the adjoint differentiates the unclamped estimator)
    st.indirect_clamp_factor = 1.0
    st.rr_min_bounces = 100              # no Russian roulette
    return st


def _scene(texture):
    return scenes.textured_icosphere_scene(W, H, subdiv=3, texture=texture)


def _np_bilinear(tex, u, v):
    C, h, w = tex.shape
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = x - x0, y - y0
    i0 = np.mod(x0.astype(np.int64), w)
    i1 = np.mod(i0 + 1, w)
    j0 = np.clip(y0, 0, h - 1).astype(np.int64)
    j1 = np.clip(y0 + 1, 0, h - 1).astype(np.int64)
    t = tex.astype(np.float64)
    return ((1 - fx) * (1 - fy) * t[:, j0, i0] + fx * (1 - fy) * t[:, j0, i1] + (1 - fx) * fy * t[:, j1, i0]
            + fx * fy * t[:, j1, i1])


def test_texture_lookup_matches_numpy(oracle):
    tex = np.random.default_rng(1).random((5, 8, 12)).astype(np.float32)
    d = scenes.textured_icosphere_scene(8, 8, subdiv=1, texture=tex)
    o = oracle.OracleScene(d)
    rng = np.random.default_rng(2)
    uv = np.concatenate([rng.uniform(-1.5, 2.5, (400, 2)), [[0.0, 0.0], [1.0, 1.0], [0.999, 0.001], [-0.01, 1.2]]])
    uv = uv.astype(np.float32)
    got = o.texture_sample(uv)
    want = _np_bilinear(tex, uv[:, 0].astype(np.float64), uv[:, 1].astype(np.float64))
    assert np.allclose(got, want.T, rtol=0, atol=2e-6)


def test_constant_texture_equals_untextured_material(oracle):
    # texture (0.5, 0.5, 0.5, 0.5, 0.5) on the sphere vs the same Disney knobs
    # given as a constant material: identical paths, same radiance up to the
    # rounding of the bilinear weights
    tex = np.full((5, 16, 16), 0.5, np.float32)
    d = _scene(tex)
    mats = d.mats.copy()
    k = np.nonzero(mats["flags"] & abi.RT_MATF_TEXTURED)[0]
    assert len(k) == 1
    mats["flags"][k] &= ~abi.RT_MATF_TEXTURED
    mats["color"][k] = 0.5
    mats["roughness"][k] = 0.5
    mats["metallic"][k] = 0.5
    plain = scenes.SceneDesc(d.prims, mats, d.lights, d.camera)
    a = oracle.OracleScene(d).render(_settings())["beauty"]
    b = oracle.OracleScene(plain).render(_settings())["beauty"]
    assert np.abs(a).sum() > 0
    assert np.allclose(a, b, rtol=1e-4, atol=1e-6)


def test_oracle_adjoint_is_linear(oracle):
    d = _scene(np.full((5, 16, 16), 0.5, np.float32))
    o = oracle.OracleScene(d)
    rng = np.random.default_rng(3)
    d1 = rng.normal(size=(3, H, W)).astype(np.float32)
    d2 = rng.normal(size=(3, H, W)).astype(np.float32)
    g1, g2 = o.adjoint(_settings(), d1), o.adjoint(_settings(), d2)
    g12 = o.adjoint(_settings(), d1 + d2)
    assert np.abs(g1).max() > 0
    assert np.allclose(g12, g1 + g2, rtol=1e-6, atol=1e-9 * np.abs(g12).max())
    assert np.allclose(o.adjoint(_settings(), 2 * d1), 2 * g1, rtol=1e-9, atol=0)


def _loss(img, target):
    return float(((img.astype(np.float64) - target) ** 2).sum())


def test_oracle_adjoint_matches_crn_fd_basecolor(oracle):
    gt = scenes.material_texture(16, seed=7, cell=2)
    target = oracle.OracleScene(_scene(gt)).render(_settings())["beauty"].astype(np.float64)
    theta = np.full((5, 16, 16), 0.5, np.float32)
    img = oracle.OracleScene(_scene(theta)).render(_settings())["beauty"]
    grad = oracle.OracleScene(_scene(theta)).adjoint(_settings(), (2.0 * (img - target)).astype(np.float32))
    # the 8 base-colour texels with the largest gradients
    flat = np.argsort(-np.abs(grad[:3]).ravel())[:8]
    h = 1e-3
    worst = 0.0
    for f in flat:
        c, j, i = np.unravel_index(f, grad[:3].shape)
        tp, tm = theta.copy(), theta.copy()
        tp[c, j, i] += h
        tm[c, j, i] -= h
        lp = _loss(oracle.OracleScene(_scene(tp)).render(_settings())["beauty"], target)
        lm = _loss(oracle.OracleScene(_scene(tm)).render(_settings())["beauty"], target)
        fd = (lp - lm) / (2 * h)
        rel = abs(fd - grad[c, j, i]) / max(abs(fd), 1e-12)
        print(f"texel ({c},{j},{i}): adjoint {grad[c, j, i]:.6e} fd {fd:.6e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    assert worst <= 1e-3
