"""Object hierarchies on the GPU (rt_node: Group / Transform / Instance /
camera-invisible wrappers) against the oracle's restatement of Transform.Hit,
Instance.Hit and Group (Transform.cs:74-240, Instance.cs:54-84,
Group.cs:20-80):

  * closest hits of camera rays and of random world rays are bit-identical
    (primitive, t, u, v): the ray is mapped into each wrapper's frame with the
    same multiply-adds, the shared template BVHs are traversed in the
    reference's order;
  * same-stream renders agree per pixel (point, normal and tangent frame
    mapped back to world space, material overrides);
  * occlusion queries agree with the oracle's closest hits;
  * instances share their template: 64 instances of a mesh cost one mesh of
    device memory, not 64;
  * the refusals (caustics, adjoint, geometry lights under a transform,
    nesting deeper than 4).
"""
import numpy as np
import pytest

import hierarchy_cases as hc
from raytracer import abi, scenes
from test_gpu_primary import _uv_equal
from test_gpu_render import _same_stream

pytestmark = pytest.mark.gpu

POSES = [((1.5, 0.6, 0.9), (30.0, 45.0, 10.0), (0.3, -0.2, 0.5)),
         ((2.0, 0.5, 1.0), (-70.0, 10.0, 120.0), (0.0, 0.0, 0.0))]


def _rays_equal(native, oracle, ctx, desc, rays):
    sc = native.Scene(ctx, desc)
    p, t, u, v = sc.trace_rays(rays)
    op, ot, ou, ov = oracle.OracleScene(desc).trace_rays(rays)
    print(f"rays={p.size} hits={(p >= 0).sum()} prim-mismatch={(p != op).sum()}", flush=True)
    assert (p >= 0).sum() > 0.1 * p.size
    assert np.array_equal(p, op)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    _uv_equal(desc, p, u, v, ou, ov)


@pytest.mark.parametrize("pose", POSES)
def test_transformed_sphere_rays_bit_exact(native, oracle, gpu_ctx, pose):
    M = hc.trs(*pose).astype(np.float32)
    _rays_equal(native, oracle, gpu_ctx, hc.ellipsoid_scene(M), hc.random_rays(50000, 5, 3.0, 1.5))


def test_instanced_template_rays_bit_exact(native, oracle, gpu_ctx):
    _rays_equal(native, oracle, gpu_ctx, hc.instanced_scene(), hc.random_rays(100000, 9, 3.0, 3.0))


@pytest.mark.parametrize("n_children,wrap", [(12, False), (3, False), (9, True)])
def test_group_rays_bit_exact(native, oracle, gpu_ctx, n_children, wrap):
    hier, _ = hc.group_scene(n_children, wrap_group=wrap)
    _rays_equal(native, oracle, gpu_ctx, hier, hc.random_rays(50000, 13))


def test_instanced_camera_hits_bit_exact(native, oracle, gpu_ctx):
    # includes a camera-invisible instance: primary rays pass through it
    W, H, spp = 160, 120, 2
    desc = hc.instanced_scene(W, H)
    st = abi.settings_defaults(W, H, spp, 8)
    p, t, u, v = native.Scene(gpu_ctx, desc).trace_primary(st, 0, spp)
    op, ot, ou, ov, _, _ = oracle.OracleScene(desc).trace_primary(st, 0, spp)
    print(f"rays={p.size} hits={(p >= 0).sum()} prim-mismatch={(p != op).sum()}", flush=True)
    assert np.array_equal(p, op)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    _uv_equal(desc, p, u, v, ou, ov)


def test_instanced_render_same_streams(native, oracle, gpu_ctx):
    _same_stream(native, oracle, gpu_ctx, hc.instanced_scene(96, 72), 96, 72, 8, 8)


def test_transformed_sphere_render_same_streams(native, oracle, gpu_ctx):
    M = hc.trs(*POSES[0]).astype(np.float32)
    _same_stream(native, oracle, gpu_ctx, hc.ellipsoid_scene(M, 80, 60), 80, 60, 8, 8)


def test_occlusion_matches_oracle_hits(native, oracle, gpu_ctx):
    desc = hc.instanced_scene()
    rays = hc.random_rays(50000, 17, 3.0, 3.0)
    rng = np.random.default_rng(3)
    tmax = rng.uniform(1.0, 12.0, len(rays)).astype(np.float32)
    segs = np.concatenate([rays, np.full((len(rays), 1), 1e-4, np.float32), tmax[:, None]], 1)   # MathUtils.Epsilon
    occ = native.Scene(gpu_ctx, desc).occluded(segs)
    op, ot, _, _ = oracle.OracleScene(desc).trace_rays(rays)
    want = (op >= 0) & (ot <= tmax)
    print(f"occluded {occ.mean():.3f}, mismatches {(occ.astype(bool) != want).sum()}", flush=True)
    assert np.array_equal(occ.astype(bool), want)


def test_instances_share_the_template(native, gpu_ctx):
    v, f = scenes.icosphere(5)                      # 20,480 triangles
    n_inst = 64
    rng = np.random.default_rng(1)
    poses = [hc.trs((s, s, s), (0.0, float(a), 0.0), (float(x), 0.0, float(z)))
             for s, a, x, z in zip(rng.uniform(0.2, 0.5, n_inst), rng.uniform(0, 360, n_inst),
                                    rng.uniform(-5, 5, n_inst), rng.uniform(-5, 5, n_inst))]
    cam = scenes.make_camera((0.0, 4.0, 12.0), (0.0, 0.0, 0.0), 45.0, 1.0)
    b = scenes.SceneBuilder()
    m = b.lambertian((0.6, 0.6, 0.6))
    mesh = b.node_prims(b.mesh(v, f, m))
    for M in poses:
        b.add_root(b.node_transform(b.node_instance(mesh), M))
    shared = native.Scene(gpu_ctx, b.build(cam, sky=(0.5, 0.7, 1.0)))
    b = scenes.SceneBuilder()
    m = b.lambertian((0.6, 0.6, 0.6))
    for M in poses:
        vv = (np.c_[v, np.ones(len(v))] @ M)[:, :3]
        b.mesh(vv, f, m)
    baked = native.Scene(gpu_ctx, b.build(cam))
    sb, bb = shared.info().device_bytes, baked.info().device_bytes
    print(f"device bytes: shared template {sb / 1e6:.1f} MB, 64 baked copies {bb / 1e6:.1f} MB", flush=True)
    assert sb * 20 < bb
    r = native.Renderer(shared, 4, 4).render(64, 64)
    assert r.buffers.beauty.mean() > 0


def test_hierarchy_refusals(native, gpu_ctx):
    desc = hc.instanced_scene(32, 24)
    sc = native.Scene(gpu_ctx, desc)
    with pytest.raises(native.RtError) as e:
        native.Renderer(sc, 2, 4, enable_caustics=True, caustic_photons=1000).render(32, 24)
    assert e.value.code == abi.RT_ERR_UNSUPPORTED
    b = scenes.SceneBuilder()
    m = b.lambertian((0.5, 0.5, 0.5))
    n = b.node_prims(b.sphere((0.0, 0.0, 0.0), 1.0, m))
    for _ in range(5):
        n = b.node_transform(n, hc.translate(0.1, 0.0, 0.0))
    b.add_root(n)
    with pytest.raises(native.RtError) as e:
        native.Scene(gpu_ctx, b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40.0, 1.0)))
    assert e.value.code == abi.RT_ERR_UNSUPPORTED
    b = scenes.SceneBuilder()
    em = b.emissive((1.0, 1.0, 1.0), 5.0)
    q = b.quad((-0.5, 2.0, -0.5), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), em)
    b.add_root(b.node_transform(b.node_prims(q), hc.translate(0.0, 1.0, 0.0)))
    with pytest.raises(native.RtError) as e:
        native.Scene(gpu_ctx, b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40.0, 1.0)))
    assert e.value.code == abi.RT_ERR_UNSUPPORTED


def test_batch_with_a_hierarchy_scene(native, gpu_ctx):
    hier = native.Scene(gpu_ctx, hc.instanced_scene(48, 36))
    flat = native.Scene(gpu_ctx, scenes.cornell(40, 40))
    items = [(native.Renderer(hier, 4, 6), 48, 36, False), (native.Renderer(flat, 3, 6), 40, 40, False)]
    single = [r.render(w, h) for r, w, h, _ in items]
    batch, _ = native.render_batch(items)
    for one, b in zip(single, batch):
        for k in ("beauty", "beauty_a", "beauty_b"):
            assert np.array_equal(getattr(one.buffers, k), getattr(b.buffers, k)), k
