"""Image textures on the GPU (rt_image, kMatImage, k_fp) against the oracle:
primary hits bit-exact, first-hit albedo AOVs within 2e-4 (the GPU's
log2f and the C library's may differ in the last bit of the LOD) and the
radiance on the same sample stream, for a textured quad at a known mip
level, a thin-lens camera, filtering on and off, a mixed scene (tiled plane,
Disney sphere, smooth OBJ mesh with UVs), and a smooth triangle that keeps
the DpDu an earlier accepted quad hit left on the reference's record."""
import os

import numpy as np
import pytest

from raytracer import abi, native, obj_loader, scene_loader, scenes
from test_image_texture import HERE, rand_image, textured_quad

pytestmark = pytest.mark.gpu


def _compare(name, desc, gpu_ctx, oracle, W, H, spp, depth, filt=0, alb_tol=2e-4, alb_frac=1.0, sampler="sobol"):
    st = abi.settings_defaults(W, H, spp, depth)
    st.capture_aovs = 1
    st.texture_filtering = filt
    prng = sampler == "prng"
    sc = native.Scene(gpu_ctx, desc)
    try:
        if not prng:   # the primary-hit hook traces Sobol samples
            p, t, _, _ = sc.trace_primary(st, 0, 2)
        r = native.Renderer(sc, spp, depth, sampler=sampler)
        r.settings.texture_filtering = filt
        buf = r.render(W, H, capture_aovs=True).buffers
    finally:
        sc.close()
    osc = oracle.OracleScene(desc)
    if not prng:
        op, ot, _, _, _, _ = osc.trace_primary(st, 0, 2)
        assert np.array_equal(p, op) and np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    else:
        st.sampler = abi.RT_SAMPLER_PRNG
    o = osc.render(st, aovs=True)
    diff = np.abs(buf.albedo_a - o["albedo_a"])
    within = float((diff < alb_tol).mean())
    rel = abs(float(buf.beauty.mean()) - float(o["beauty"].mean())) / max(float(o["beauty"].mean()), 1e-6)
    print(f"{name}: max |albedo gpu - oracle| {diff.max():.2e}, {within:.4f} within {alb_tol}; mean radiance "
          f"gpu {buf.beauty.mean():.6f} oracle {o['beauty'].mean():.6f} rel {rel:.2e}", flush=True)
    assert buf.beauty.mean() > 0 and within >= alb_frac and rel < 1e-3
    if alb_frac < 1.0:
        assert diff.max() < 0.05
    return buf, o


@pytest.mark.parametrize("filt", [0, 2])
@pytest.mark.parametrize("spp", [1, 4])
def test_textured_quad(gpu_ctx, oracle, filt, spp):
    img = rand_image(256, 256, 6)
    _compare(f"quad spp={spp} filtering={filt}", textured_quad(64, 64, img), gpu_ctx, oracle, 64, 64, spp, 2, filt)


def test_textured_quad_prng_sampler(gpu_ctx, oracle):
    # the footprint's ray differentials start from the stratified jitter of
    # the camera ray (Renderer.cs:677-697)
    img = rand_image(256, 256, 6)
    _compare("quad prng spp=6", textured_quad(64, 64, img), gpu_ctx, oracle, 64, 64, 6, 2, 0, sampler="prng")


@pytest.mark.parametrize("filt", [0, 2])
def test_thin_lens_camera(gpu_ctx, oracle, filt):
    img = rand_image(128, 96, 8)
    desc = textured_quad(48, 40, img, fov=60.0, dist=2.5, lens=0.15)
    # the lens sample goes through the device's cosf/sinf and the C
    # library's; a last-bit difference in the footprint can move
    # major/minor across 1.0, where ImageTexture's ceil() switches between one
    # and two taps, so a few pixels may differ by a quarter-texel shift
    _compare(f"thin lens filtering={filt}", desc, gpu_ctx, oracle, 48, 40, 4, 2, filt, alb_frac=0.99)


def _mixed_scene(W, H):
    from PIL import Image
    b = scenes.SceneBuilder()
    grid = np.asarray(Image.open(os.path.join(HERE, "textures", "grid-uv.png")).convert("RGBA"))
    check = np.asarray(Image.open(os.path.join(HERE, "textures", "checkerboard.png")).convert("RGBA"))
    floor = b.lambertian((0.5, 0.5, 0.5), image=b.image(check, (4.0, 4.0)))
    b.plane((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), floor)
    ball = b.disney((0.8, 0.8, 0.8), roughness=0.4, metallic=0.3, image=b.image(grid))
    b.sphere((1.3, 1.0, 0.0), 1.0, ball)
    mesh = obj_loader.load(os.path.join(HERE, "models", "test-sphere.obj"))
    tris = (mesh.tris.reshape(-1, 3) + np.float32([-1.3, 1.0, 0.5])).reshape(-1, 9)
    b.mesh_triangles(tris, b.lambertian((0.8, 0.8, 0.8), image=b.image(grid, (2.0, 1.0))), mesh.normals, mesh.uvs)
    b.triangle(np.float32([-3, 0.01, -2]), np.float32([3, 0.01, -2]), np.float32([0, 2.5, -2.5]),
               b.lambertian((0.6, 0.6, 0.6), image=b.image(check)))
    b.area_light((-1.0, 4.0, -1.0), (2.0, 0.0, 0.0), (0.0, 0.0, 2.0), intensity=8.0)
    return b.build(scenes.make_camera((0.0, 2.0, 5.0), (0.0, 0.8, 0.0), 50.0, W / H))


def test_mixed_textured_scene(gpu_ctx, oracle):
    _compare("mixed scene", _mixed_scene(80, 60), gpu_ctx, oracle, 80, 60, 4, 4)


def test_smooth_triangle_keeps_an_earlier_hits_dpdu(gpu_ctx, oracle):
    # world list in object order: the quad (behind, accepted first: it sets
    # DpDu = U), then the smooth triangle in front, which leaves DpDu alone,
    # so its footprint uses the quad's U / V instead of its tangent frame
    W = H = 48
    img = rand_image(256, 256, 9)

    def scene(quad_first):
        b = scenes.SceneBuilder()
        grey = b.lambertian((0.5, 0.5, 0.5))
        tex = b.lambertian((0.8, 0.8, 0.8), image=b.image(img))
        tri = (np.float32([-2, -2, 0]), np.float32([2, -2, 0]), np.float32([-2, 2, 0.5]))
        nrm = [(0.1, 0.0, 1.0), (0.0, 0.1, 1.0), (-0.1, 0.0, 1.0)]
        quad = ((-3.0, -3.0, -1.0), (6.0, 0.0, 0.0), (0.0, 6.0, 0.0), grey)
        if quad_first:
            b.quad(*quad)
        b.triangle(*tri, tex, normals=nrm)
        if not quad_first:
            b.quad(*quad)
        b.directional_light((0.0, 0.0, -1.0), intensity=1.0)
        return b.build(scenes.make_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 60.0, 1.0))

    stale, _ = _compare("smooth triangle after the quad", scene(True), gpu_ctx, oracle, W, H, 1, 2)
    own, _ = _compare("smooth triangle before the quad", scene(False), gpu_ctx, oracle, W, H, 1, 2)
    # same geometry: only the DpDu the footprint starts from differs (the
    # quad's 6-unit U against the triangle's unit tangent)
    d = np.abs(stale.albedo_a - own.albedo_a)
    print(f"stale-DpDu pixels differ by up to {d.max():.3f}", flush=True)
    assert d.max() > 1e-2


def test_image_material_refusals(gpu_ctx):
    img = rand_image(8, 8, 1)
    b = scenes.SceneBuilder()
    m = b.material(abi.RT_MAT_DIELECTRIC, (1, 1, 1), ior=1.5, image=b.image(img))
    b.sphere((0, 0, 0), 1.0, m)
    with pytest.raises(native.RtError) as e:
        native.Scene(gpu_ctx, b.build(scenes.make_camera((0, 0, 3), (0, 0, 0), 40.0, 1.0)))
    assert e.value.code == abi.RT_ERR_UNSUPPORTED
    b = scenes.SceneBuilder()
    m = b.lambertian((1, 1, 1), image=0)
    b.sphere((0, 0, 0), 1.0, m)
    with pytest.raises(native.RtError) as e:
        native.Scene(gpu_ctx, b.build(scenes.make_camera((0, 0, 3), (0, 0, 0), 40.0, 1.0)))
    assert e.value.code == abi.RT_ERR_INVALID_ARGUMENT
