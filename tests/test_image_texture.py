"""ImageTexture (ImageTexture.cs:42-289), FootprintMath (FootprintMath.cs:38-119)
and the camera differentials (Camera.cs:127-153, Renderer.cs:616-624) in the
oracle, pinned by what the definitions fix:

  * level 0 is the gamma-2.2 decode of the bytes (within one ulp of x^2.2
    in float64); level sizes halve with floor down to 1x1, 1 +
    floor(log2(max(w, h))) levels; a power-of-two image's level m is the
    mean of its 2^m x 2^m blocks;
  * bilinear lookups return the texel at its centre, the average halfway
    between texels, wrap across the u seam, and v = 0 is the bottom row;
  * a footprint of 2^k texels selects level k, anisotropic footprints
    average taps along the major axis, any footprint of a constant image
    returns the constant;
  * a camera facing a textured quad whose pixel covers exactly 4 texels
    shades with level 2; a 1-texel checkerboard seen from afar averages to
    mid-grey with filtering on and aliases with it off.
Plus the YAML front-end's image textures (no GPU needed)."""
import math
import os

import numpy as np
import pytest

from raytracer import abi, scene_loader, scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rand_image(h, w, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, (h, w, 4), dtype=np.uint8)


def bilinear(level, u, v):
    """Texel-centre bilinear with repeat addressing, float64."""
    h, w = level.shape[:2]
    px, py = u * w - 0.5, v * h - 0.5
    x0, y0 = math.floor(px), math.floor(py)
    fx, fy = px - x0, py - y0
    t = lambda x, y: level[y % h, x % w].astype(np.float64)
    top = t(x0, y0) * (1 - fx) + t(x0 + 1, y0) * fx
    bot = t(x0, y0 + 1) * (1 - fx) + t(x0 + 1, y0 + 1) * fx
    return top * (1 - fy) + bot * fy


def test_decode_table_and_mip_chain(oracle):
    img = rand_image(20, 37, 1)
    lv = oracle.image_levels(img)
    assert [l.shape[:2] for l in lv] == [(20, 37), (10, 18), (5, 9), (2, 4), (1, 2), (1, 1)]
    # MathF.Pow(i * (1/255f), 2.2f): the exponent is the float 2.2f
    want = np.power(img[..., :3].astype(np.float32) * np.float32(1 / 255), float(np.float32(2.2)), dtype=np.float64)
    # the C library's powf is within one ulp of the exact power: 2^-23 relative
    assert np.all(np.abs(lv[0] - want) <= want * 2.0 ** -23 + 1e-45)
    sq = rand_image(16, 16, 2)
    lv = oracle.image_levels(sq)
    assert len(lv) == 5
    base = lv[0].astype(np.float64)
    for m in range(1, 5):
        b = 2 ** m
        means = base.reshape(16 // b, b, 16 // b, b, 3).mean(axis=(1, 3))
        assert np.abs(lv[m] - means).max() < 2e-6, m


def test_bilinear_closed_forms(oracle):
    img = rand_image(8, 12, 3)
    lv0 = oracle.image_levels(img)[0].astype(np.float64)
    h, w = 8, 12
    # centre of texel (i, j): Value flips v, so row j sits at v = 1 - (j + 0.5) / h
    uv, want = [], []
    for (i, j) in [(0, 0), (5, 3), (11, 7), (7, 0)]:
        uv.append(((i + 0.5) / w, 1 - (j + 0.5) / h))
        want.append(lv0[j, i])
    uv.append((6.0 / w, 1 - 2.5 / h))                      # halfway between columns 5 and 6 of row 2
    want.append(0.5 * (lv0[2, 5] + lv0[2, 6]))
    uv.append((0.25 / w, 1 - 4.5 / h))                     # the u seam: a quarter towards column 0 from column 11
    want.append(0.25 * lv0[4, 11] + 0.75 * lv0[4, 0])
    got = oracle.image_sample(img, np.array(uv, np.float32))
    assert np.abs(got - np.array(want)).max() < 2e-6
    # uv_scale tiles: u * 3 wraps to the same texel
    g2 = oracle.image_sample(img, np.array([[(5.5 / w) / 3 + 1 / 3, 1 - 3.5 / h]], np.float32), scale=(3.0, 1.0))
    assert np.abs(g2[0] - lv0[3, 5]).max() < 2e-6


def test_footprint_selects_level_and_filters(oracle):
    img = rand_image(64, 64, 4)
    lv = oracle.image_levels(img)
    rng = np.random.default_rng(5)
    uv = rng.random((40, 2)).astype(np.float32) * 0.8 + 0.1
    for k in (1, 2, 3):
        a = 2.0 ** k / 64
        got = oracle.image_sample(img, uv, np.tile([a, 0, 0, a], (len(uv), 1)))
        want = np.array([bilinear(lv[k], u, 1 - v) for u, v in uv])
        assert np.abs(got - want).max() < 1e-5, k
        # a plausible slip (a factor 2 in the footprint) lands on another level
        assert np.abs(oracle.image_sample(img, uv, np.tile([2 * a, 0, 0, 2 * a], (len(uv), 1))) - want).max() > 1e-2
    # anisotropy 4: four taps along the major (v) axis at the minor axis' level 0
    a = 1.0 / 64
    got = oracle.image_sample(img, uv, np.tile([a, 0, 0, 4 * a], (len(uv), 1)))
    want = []
    for u, v in uv:
        taps = [bilinear(lv[0], u, 1 - (v + t * 4 * a)) for t in (-0.375, -0.125, 0.125, 0.375)]
        want.append(np.mean(taps, axis=0))
    assert np.abs(got - np.array(want)).max() < 1e-5
    flat = np.full((32, 32, 4), 77, np.uint8)
    c = np.float32(77 * np.float32(1 / 255)) ** 2.2
    fps = rng.normal(size=(40, 4)).astype(np.float32) * 0.05
    assert np.abs(oracle.image_sample(flat, uv, fps) - c).max() < 1e-6


def textured_quad(W, H, img, fov=90.0, dist=1.0, lens=0.0):
    b = scenes.SceneBuilder()
    m = b.lambertian((0.8, 0.8, 0.8), image=b.image(img))
    b.quad((-1.0, -1.0, 0.0), (2.0, 0.0, 0.0), (0.0, 2.0, 0.0), m)
    b.directional_light((0.0, 0.0, -1.0), intensity=1.0)
    cam = scenes.make_camera((0.0, 0.0, dist), (0.0, 0.0, 0.0), fov, W / H, aperture=lens,
                             focus_dist=max(dist, 1.0))
    return b.build(cam)


def test_camera_footprint_of_four_texels_uses_level_two(oracle):
    # 90 degree fov at distance 1: a 64-pixel frame spans the 2-unit quad
    # (u, v in [0, 1]), so one pixel is 1/64 of the texture = 4 texels of a
    # 256-texel image -> level 2
    W = H = 64
    img = rand_image(256, 256, 6)
    lv = oracle.image_levels(img)
    st = abi.settings_defaults(W, H, 1, 1)
    st.capture_aovs = 1
    osc = oracle.OracleScene(textured_quad(W, H, img))
    prim, _, u, v, _, _ = osc.trace_primary(st, 0, 1)
    alb = osc.render(st, aovs=True)["albedo_a"].reshape(3, -1).T
    hit = prim == 0
    assert hit.mean() > 0.9
    want = np.array([bilinear(lv[2], uu, 1 - vv) for uu, vv in zip(u[hit], v[hit])])
    e1 = np.abs(alb[hit] - want).max(1)
    # where rounding leaves major/minor a hair above 1, ceil() takes two taps
    # a quarter of the footprint either side along the major axis
    d = 0.25 / 64
    tu = np.array([0.5 * (bilinear(lv[2], uu - d, 1 - vv) + bilinear(lv[2], uu + d, 1 - vv))
                   for uu, vv in zip(u[hit], v[hit])])
    tv = np.array([0.5 * (bilinear(lv[2], uu, 1 - vv - d) + bilinear(lv[2], uu, 1 - vv + d))
                   for uu, vv in zip(u[hit], v[hit])])
    e = np.minimum(e1, np.minimum(np.abs(alb[hit] - tu).max(1), np.abs(alb[hit] - tv).max(1)))
    print(f"one tap on level 2: {(e1 < 2e-4).mean():.3f} of the pixels; worst {e.max():.2e}", flush=True)
    assert e.max() < 2e-4 and (e1 < 2e-4).mean() > 0.9
    other = np.array([bilinear(lv[1], uu, 1 - vv) for uu, vv in zip(u[hit], v[hit])])
    assert np.abs(alb[hit] - other).max() > 1e-2


def test_distant_checkerboard_filters_to_grey(oracle):
    W = H = 32
    ck = np.zeros((256, 256, 4), np.uint8)
    ck[(np.arange(256)[:, None] + np.arange(256)[None, :]) % 2 == 0] = 255
    desc = textured_quad(W, H, ck, fov=20.0, dist=4.0)
    st = abi.settings_defaults(W, H, 1, 1)
    st.capture_aovs = 1
    on = oracle.OracleScene(desc).render(st, aovs=True)["albedo_a"].reshape(3, -1)[0]
    st.texture_filtering = 2
    off = oracle.OracleScene(desc).render(st, aovs=True)["albedo_a"].reshape(3, -1)[0]
    print(f"filtered: mean {on.mean():.3f} max |x - 0.5| {np.abs(on - 0.5).max():.3f}; "
          f"point-sampled (bilinear level 0): spread {off.std():.3f}", flush=True)
    assert np.abs(on - 0.5).max() < 0.05
    assert off.std() > 0.1


def test_yaml_image_textures(tmp_path):
    from PIL import Image
    Image.fromarray(rand_image(8, 16, 7)).save(tmp_path / "t.png")
    doc = {"lights": [], "world": {"sky": {"color": [0, 0, 0]}},
           "materials": [{"id": "a", "type": "lambertian", "texture": {"type": "image", "path": "t.png",
                                                                       "uv_scale": [4, 2]}},
                         {"id": "b", "type": "disney", "roughness": 0.3,
                          "texture": {"type": "image", "path": os.path.join(HERE, "textures", "grid-uv.png")}},
                         {"id": "c", "type": "lambertian", "texture": {"type": "image", "path": "nope.png"}}],
           "entities": [{"type": "sphere", "material": m} for m in ("a", "b", "c")]}
    w = []
    d = scene_loader.load_data(doc, 16, 16, base_dir=str(tmp_path), warnings=w)
    ma, mb, mc = (d.mats[i] for i in d.prims["material"])
    assert ma["flags"] & abi.RT_MATF_IMAGE and ma["image"] == 0 and mb["flags"] & abi.RT_MATF_IMAGE
    assert mb["kind"] == abi.RT_MAT_DISNEY and mb["image"] == 1
    assert not mc["flags"] & abi.RT_MATF_IMAGE and np.allclose(mc["color"], [1, 0, 1]) and "not found" in w[0]
    assert len(d.images) == 2 and d.images[0][1] == (4.0, 2.0) and d.images[1][0].shape == (1024, 1024, 4)
    with pytest.raises(scene_loader.SceneLoadError, match="checker"):
        scene_loader.load_data({"lights": [], "materials": [{"id": "x", "type": "lambertian",
                                                             "texture": {"type": "checker"}}]}, 8, 8)
    with pytest.raises(scene_loader.SceneLoadError, match="dielectric"):
        scene_loader.load_data({"lights": [], "materials": [{"id": "x", "type": "dielectric",
                                                             "texture": {"type": "image", "path": "t.png"}}]},
                               8, 8, base_dir=str(tmp_path))
