"""Textured rendering and the path-replay adjoint on the GPU.

  * textured render vs the oracle with the same sampler streams;
  * rt_scene_set_textures: replacing and restoring a texture set reproduces
    the films bit for bit;
  * rt_adjoint vs the oracle's FP64 adjoint (central differences of the
    BSDF at each textured hit, summed along the render recursion) on the same
    paths, all five channels;
  * rt_adjoint vs central finite differences of the GPU render with common
    random numbers for base-colour texels (at metallic <= 0.5 with Russian
    roulette off the estimator is smooth in the base colour and its sampling
    does not depend on it, so the two agree up to O(h^2) and fp32 noise);
  * rt_adjoint vs the same differences for roughness and metallic texels,
    averaged over independent seeds (their sampling depends on them);
  * the torch.autograd wrapper and a few optimisation steps.
"""
import numpy as np
import pytest

from raytracer import abi, scenes
from test_gpu_render import _same_stream

pytestmark = pytest.mark.gpu


def _settings(native, sc, spp, depth):
    r = native.Renderer(sc, spp, depth, max_sample_radiance=3.0e38, indirect_clamp_factor=1.0, rr_min_bounces=100)
    return r


def _loss(img, target):
    return float(((img.astype(np.float64) - target) ** 2).sum())


def test_textured_render_matches_oracle(native, oracle, gpu_ctx):
    d = scenes.textured_icosphere_scene(64, 48, subdiv=4, texture=scenes.material_texture(64, cell=4))
    _same_stream(native, oracle, gpu_ctx, d, 64, 48, 8, 6)


def test_set_textures_reproduces_films(native, gpu_ctx):
    a_tex = scenes.material_texture(64, seed=1, cell=4)
    b_tex = scenes.material_texture(64, seed=2, cell=4)
    sc = native.Scene(gpu_ctx, scenes.textured_icosphere_scene(48, 48, subdiv=3, texture=a_tex))
    r = native.Renderer(sc, 4, 4)
    a = r.render(48, 48).buffers.beauty.copy()
    sc.set_textures(b_tex)
    b = r.render(48, 48).buffers.beauty.copy()
    fresh = native.Renderer(native.Scene(gpu_ctx, scenes.textured_icosphere_scene(48, 48, subdiv=3, texture=b_tex)),
                            4, 4).render(48, 48).buffers.beauty
    sc.set_textures(a_tex)
    a2 = r.render(48, 48).buffers.beauty
    assert not np.array_equal(a, b)
    assert np.array_equal(b, fresh)
    assert np.array_equal(a, a2)


def test_adjoint_matches_oracle_adjoint(native, oracle, gpu_ctx):
    W = H = 48
    spp, depth = 8, 3
    rng = np.random.default_rng(11)
    theta = rng.uniform(0.2, 0.8, (5, 24, 24)).astype(np.float32)
    gt = scenes.material_texture(24, seed=12, cell=3)
    d_gt = scenes.textured_icosphere_scene(W, H, subdiv=3, texture=gt)
    d = scenes.textured_icosphere_scene(W, H, subdiv=3, texture=theta)
    r = _settings(native, native.Scene(gpu_ctx, d), spp, depth)
    target = _settings(native, native.Scene(gpu_ctx, d_gt), spp, depth).render(W, H).buffers.beauty.astype(np.float64)
    img = r.render(W, H).buffers.beauty
    db = (2.0 * (img - target)).astype(np.float32)
    g, st = r.adjoint(W, H, db)
    o = oracle.OracleScene(d).adjoint(r.settings, db)
    err = np.linalg.norm(g - o) / np.linalg.norm(o)
    per = [np.linalg.norm(g[c] - o[c]) / max(np.linalg.norm(o[c]), 1e-30) for c in range(5)]
    print(f"relative L2 error {err:.3e}; per channel {['%.2e' % x for x in per]}; "
          f"adjoint device {st.device_ms:.1f} ms", flush=True)
    assert np.abs(o).max() > 0
    assert err <= 1e-2
    assert max(per) <= 2e-2


def test_adjoint_matches_gpu_crn_fd_basecolor(native, gpu_ctx):
    W = H = 64
    spp, depth = 8, 4
    gt = scenes.material_texture(32, seed=21, cell=2)
    theta = np.full((5, 32, 32), 0.5, np.float32)
    d_gt = scenes.textured_icosphere_scene(W, H, subdiv=4, texture=gt)
    sc = native.Scene(gpu_ctx, scenes.textured_icosphere_scene(W, H, subdiv=4, texture=theta))
    r = _settings(native, sc, spp, depth)
    target = _settings(native, native.Scene(gpu_ctx, d_gt), spp, depth).render(W, H).buffers.beauty.astype(np.float64)
    img = r.render(W, H).buffers.beauty
    g, _ = r.adjoint(W, H, (2.0 * (img - target)).astype(np.float32))
    base = np.abs(g[:3])
    cand = np.flatnonzero(base.ravel() >= 0.1 * base.max())
    pick = np.random.default_rng(5).choice(cand, size=min(16, len(cand)), replace=False)
    h = 1e-3
    worst = 0.0
    for f in pick:
        c, j, i = np.unravel_index(f, base.shape)
        tp, tm = theta.copy(), theta.copy()
        tp[c, j, i] += h
        tm[c, j, i] -= h
        sc.set_textures(tp)
        lp = _loss(r.render(W, H).buffers.beauty, target)
        sc.set_textures(tm)
        lm = _loss(r.render(W, H).buffers.beauty, target)
        fd = (lp - lm) / float(tp[c, j, i] - tm[c, j, i])
        rel = abs(fd - g[c, j, i]) / max(abs(fd), 1e-12)
        print(f"texel ({c},{j},{i}): adjoint {g[c, j, i]:.6e} fd {fd:.6e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    sc.set_textures(theta)
    assert worst <= 1e-2


def test_adjoint_matches_gpu_crn_fd_roughness_metallic(native, gpu_ctx):
    """Roughness and metallic texels against render-level central differences
    (h = 1e-3, common random numbers), averaged over 8 independent seeds.

    Both parameters move the sampling: the lobe-selection probabilities and
    the VNDF directions depend on them, so one realisation's difference
    quotient also counts the samples that switch lobe between theta - h and
    theta + h (a term whose mean is zero and whose variance grows as 1/h),
    while the path-replay adjoint holds the sampling fixed; the two agree in
    expectation, not per realisation.  The loss is linear (a fixed dLoss /
    dBeauty image from an independent 4096-spp render), so no correlation
    between the image and its weights biases either side.  Metallic sits at
    0.25: at 0.5 the reference switches the specular lobe's selection weight
    (DisneyBsdf.cs:1112, metallic > 0.5), a jump in every realisation.
    The texels: those whose gradient is at least a tenth of the channel's
    largest, up to 16 (8 roughness and 10 metallic texels here; measured
    worst relative difference 5.6e-3 and 7.2e-3)."""
    W = H = 64
    depth, spp, seeds, h = 4, 4096, 8, 1e-3
    gt = scenes.material_texture(32, seed=21, cell=2)
    theta = np.full((5, 32, 32), 0.5, np.float32)
    theta[4] = 0.25
    sc = native.Scene(gpu_ctx, scenes.textured_icosphere_scene(W, H, subdiv=4, texture=theta))
    sc_gt = native.Scene(gpu_ctx, scenes.textured_icosphere_scene(W, H, subdiv=4, texture=gt))

    def rend(scene, n, seed):
        return native.Renderer(scene, n, depth, max_sample_radiance=3.0e38, indirect_clamp_factor=1.0,
                               rr_min_bounces=100, seed_xor=seed)

    target = rend(sc_gt, 1024, 0x999).render(W, H).buffers.beauty.astype(np.float64)
    D = (2.0 * (rend(sc, 4096, 0x777).render(W, H).buffers.beauty.astype(np.float64) - target)).astype(np.float32)
    D64 = D.astype(np.float64)
    pilot, _ = rend(sc, 4096, 0x778).adjoint(W, H, D)
    rng = np.random.default_rng(5)
    picks = {}
    for c in (3, 4):
        a = np.abs(pilot[c]).ravel()
        cand = np.flatnonzero(a >= 0.1 * a.max())
        picks[c] = rng.choice(cand, size=min(16, len(cand)), replace=False)
    G = {c: [] for c in picks}
    F = {c: [] for c in picks}
    for k in range(seeds):
        r = rend(sc, spp, 0x1000 + k)
        g, _ = r.adjoint(W, H, D)
        for c, texels in picks.items():
            G[c].append(g[c].ravel()[texels])
            fd = []
            for f in texels:
                j, i = np.unravel_index(f, (32, 32))
                tp, tm = theta.copy(), theta.copy()
                tp[c, j, i] += h
                tm[c, j, i] -= h
                sc.set_textures(tp)
                lp = float((r.render(W, H).buffers.beauty.astype(np.float64) * D64).sum())
                sc.set_textures(tm)
                lm = float((r.render(W, H).buffers.beauty.astype(np.float64) * D64).sum())
                fd.append((lp - lm) / float(tp[c, j, i] - tm[c, j, i]))
            F[c].append(fd)
        sc.set_textures(theta)
    for c, name in ((3, "roughness"), (4, "metallic")):
        g, f = np.mean(G[c], 0), np.mean(F[c], 0)
        rel = np.abs(g - f) / np.abs(f)
        print(f"{name}: {len(g)} texels, adjoint vs FD relative difference median {np.median(rel):.2e}, "
              f"max {rel.max():.2e}", flush=True)
        assert rel.max() <= 1e-2


def test_adjoint_rejects_unsupported(native, gpu_ctx):
    import ctypes as C
    tex = np.full((5, 8, 8), 0.5, np.float32)
    d = scenes.textured_icosphere_scene(16, 16, subdiv=1, texture=tex)
    db = np.zeros((3, 16, 16), np.float32)
    with pytest.raises(native.RtError) as e:
        native.Renderer(native.Scene(gpu_ctx, d), 1, 2).adjoint(16, 16, db)   # default clamps are active
    assert e.value.code == abi.RT_ERR_UNSUPPORTED
    fog = scenes.SceneDesc(d.prims, d.mats, d.lights, d.camera, medium=((0.01,) * 3, (0.04,) * 3),
                           texcoords=d.texcoords, texture=d.texture)
    with pytest.raises(native.RtError) as e:
        _settings(native, native.Scene(gpu_ctx, fog), 1, 2).adjoint(16, 16, db)
    assert e.value.code == abi.RT_ERR_UNSUPPORTED
    plain = native.Scene(gpu_ctx, scenes.cornell(16, 16))
    r = _settings(native, plain, 1, 2)
    r.settings.width = r.settings.height = 16
    out = np.zeros(5, np.float32)
    rc = native.load().rt_adjoint(gpu_ctx.h, plain.h, C.byref(r.settings), native._f(db), native._f(out), None)
    assert rc == abi.RT_ERR_INVALID_ARGUMENT


def test_torch_autograd_wrapper(native, gpu_ctx):
    import torch
    from raytracer.diff import render_textures
    W = H = 32
    gt = scenes.material_texture(16, seed=31, cell=2)
    sc = native.Scene(gpu_ctx, scenes.textured_icosphere_scene(W, H, subdiv=3, texture=gt))
    r = _settings(native, sc, 4, 3)
    target = torch.from_numpy(r.render(W, H).buffers.beauty.copy())
    theta = torch.full((5, 16, 16), 0.5, dtype=torch.float32, requires_grad=True)
    img = render_textures(theta, r, W, H)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    want, _ = r.adjoint(W, H, (2.0 * (img.detach() - target)).numpy())
    assert np.allclose(theta.grad.numpy(), want, rtol=1e-4, atol=1e-6 * np.abs(want).max())
    # a few steps of gradient descent lower the loss
    opt = torch.optim.Adam([theta], lr=0.05)
    losses = []
    for _ in range(6):
        opt.zero_grad()
        l = ((render_textures(theta, r, W, H) - target) ** 2).sum()
        l.backward()
        opt.step()
        with torch.no_grad():
            theta.clamp_(0.0, 1.0)
        losses.append(float(l.detach()))
    print("losses", ["%.4f" % x for x in losses], flush=True)
    assert losses[-1] < 0.8 * losses[0]
