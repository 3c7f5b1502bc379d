"""Smooth triangles (per-vertex normals, SmoothTriangle.cs:68-197) on the
GPU against the oracle: primary hits bit-exact, first-hit shading normals
within 2e-6 of the oracle's, the UV-gradient tangent frame through an
anisotropic Disney lobe, radiance of a smooth mesh on the same sample
stream, and the scene-creation refusals."""
import numpy as np
import pytest

from raytracer import abi, native, scenes
from test_smooth_triangle import ANISO, TILTED, V0, V1, V2, film_normal, plate, smooth_light_scene

pytestmark = pytest.mark.gpu


def _render(gpu_ctx, desc, st):
    sc = native.Scene(gpu_ctx, desc)
    try:
        p, t, _, _ = sc.trace_primary(st, 0, 2)
        r = native.Renderer(sc, st.samples_per_pixel, st.max_depth).render(
            st.width, st.height, capture_aovs=bool(st.capture_aovs))
    finally:
        sc.close()
    return p, t, r.buffers


def _compare(name, desc, gpu_ctx, oracle, spp, depth, W, H, normal_tol):
    st = abi.settings_defaults(W, H, spp, depth)
    st.capture_aovs = 1
    p, t, buf = _render(gpu_ctx, desc, st)
    osc = oracle.OracleScene(desc)
    op, ot, _, _, _, _ = osc.trace_primary(st, 0, 2)
    assert np.array_equal(p, op) and np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    o = osc.render(st, aovs=True)
    hit = np.abs(o["normal_a"].reshape(3, -1)).sum(0) > 0
    dn = np.abs(buf.normal_a.reshape(3, -1).T - film_normal(o))[hit].max()
    rel = abs(float(buf.beauty.mean()) - float(o["beauty"].mean())) / max(float(o["beauty"].mean()), 1e-6)
    print(f"{name}: {hit.sum()} pixels with a normal, max |dn| {dn:.2e}; mean radiance gpu "
          f"{buf.beauty.mean():.6f} oracle {o['beauty'].mean():.6f} rel {rel:.2e}", flush=True)
    assert buf.beauty.mean() > 0
    assert dn < normal_tol
    return rel


@pytest.mark.parametrize("back", [False, True])
def test_tilted_plate_normals_and_radiance(gpu_ctx, oracle, back):
    # one sample: the A-half normal is the single first hit's shading normal
    _compare(f"plate back={back} 1 spp", plate(40, 30, TILTED, back=back), gpu_ctx, oracle, 1, 4, 40, 30, 2e-6)
    rel = _compare(f"plate back={back} 8 spp", plate(40, 30, TILTED, back=back), gpu_ctx, oracle, 8, 6, 40, 30, 1e-3)
    assert rel < 1e-3


@pytest.mark.parametrize("uvs", ["xy", "yx", "skew"])
def test_uv_tangent_frame_through_anisotropic_lobe(gpu_ctx, oracle, uvs):
    xy = [(x, y) for x, y, _ in (V0, V1, V2)]
    table = {"xy": xy, "yx": [(y, x) for x, y in xy], "skew": [(0.0, 0.0), (1.0, 0.3), (-0.2, 0.9)]}
    desc = plate(48, 36, TILTED, uvs=table[uvs], mat=ANISO)
    rel = _compare(f"anisotropic plate uv={uvs}", desc, gpu_ctx, oracle, 8, 4, 48, 36, 1e-3)
    assert rel < 1e-3


def test_smooth_icosphere_mesh(gpu_ctx, oracle):
    W, H = 64, 48
    v, f = scenes.icosphere(3)
    b = scenes.SceneBuilder()
    m = b.disney((0.8, 0.5, 0.3), metallic=0.6, roughness=0.3, clearcoat=0.5)
    b.mesh(v, f, m, normals=v)
    g = b.lambertian((0.5, 0.5, 0.5))
    b.plane((0.0, -1.0, 0.0), (0.0, 1.0, 0.0), g)
    b.sphere_light((2.0, 3.0, 2.0), 0.5, intensity=40.0)
    b.directional_light((-0.5, -1.0, -0.3), intensity=1.0)
    desc = b.build(scenes.make_camera((0.0, 0.6, 3.5), (0.0, 0.0, 0.0), 45.0, W / H))
    rel = _compare("smooth icosphere", desc, gpu_ctx, oracle, 8, 8, W, H, 1e-3)
    assert rel < 1e-3


def test_scene_creation_refusals(gpu_ctx):
    with pytest.raises(native.RtError) as e:
        native.Scene(gpu_ctx, plate(8, 8, [(0, 0, 1), (0, 0, 0), (0, 0, 1)]))
    assert e.value.code == abi.RT_ERR_INVALID_ARGUMENT
    with pytest.raises(native.RtError) as e:
        native.Scene(gpu_ctx, smooth_light_scene())
    assert e.value.code == abi.RT_ERR_UNSUPPORTED
    d = plate(8, 8)
    b = scenes.SceneBuilder()
    m = b.lambertian((0.5, 0.5, 0.5))
    b.sphere((0.0, 0.0, 0.0), 1.0, m)
    s = b.build(d.camera)
    bad = scenes.SceneDesc(s.prims, s.mats, s.lights, s.camera, normals=np.ones((1, 9), np.float32))
    with pytest.raises(native.RtError) as e:
        native.Scene(gpu_ctx, bad)
    assert e.value.code == abi.RT_ERR_INVALID_ARGUMENT
