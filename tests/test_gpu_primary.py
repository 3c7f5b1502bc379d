"""Bit-exact primary hits: HIP traversal vs the oracle's restatement of the
reference traversal, on the same camera rays (same Sobol indices)."""
import numpy as np
import pytest

from helpers import axis_aligned_rays, random_rays, random_spheres, random_triangles
from raytracer import abi, scenes

pytestmark = pytest.mark.gpu


def _primary_equal(native, oracle, ctx, desc, W, H, spp, begin=0):
    st = abi.settings_defaults(W, H, spp, 8)
    sc = native.Scene(ctx, desc)
    p, t, u, v = sc.trace_primary(st, begin, spp)
    op, ot, ou, ov, _, _ = oracle.OracleScene(desc).trace_primary(st, begin, spp)
    print(f"rays={p.size} hits={(p >= 0).sum()} prim-mismatch={(p != op).sum()}", flush=True)
    assert np.array_equal(p, op)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    _uv_equal(desc, p, u, v, ou, ov)
    return p


def _uv_equal(desc, p, u, v, ou, ov):
    # (u, v) come from the intersection arithmetic (bit-exact) except on
    # spheres, where GetSphereUV uses acos/atan2 (last-ulp libm differences)
    sph = (p >= 0) & (desc.prims["kind"][np.maximum(p, 0)] == abi.RT_PRIM_SPHERE)
    for a, b in ((u, ou), (v, ov)):
        assert np.array_equal(a[~sph].view(np.uint32), b[~sph].view(np.uint32))
        assert np.all(np.abs(a[sph] - b[sph]) <= 4e-7)


def test_cornell_primary_hits_bit_exact(native, oracle, gpu_ctx):
    p = _primary_equal(native, oracle, gpu_ctx, scenes.cornell(256, 256), 256, 256, 4)
    assert len(np.unique(p)) >= 12


def test_cornell_primary_hits_later_samples(native, oracle, gpu_ctx):
    _primary_equal(native, oracle, gpu_ctx, scenes.cornell(128, 128), 128, 128, 3, begin=13)


def test_area_light_proxy_culling_bit_exact(native, oracle, gpu_ctx):
    _primary_equal(native, oracle, gpu_ctx, scenes.cornell(96, 96, light="area"), 96, 96, 2)


def test_mixed_scene_primary_hits_bit_exact(native, oracle, gpu_ctx):
    # triangles, nested meshes, culled quad, camera-invisible sphere, planes
    _primary_equal(native, oracle, gpu_ctx, random_triangles(4000, 21, planes=2), 160, 120, 2)


def test_displaced_icosphere_primary_hits_bit_exact(native, oracle, gpu_ctx):
    d = scenes.displaced_icosphere_scene(160, 90, subdiv=6)
    _primary_equal(native, oracle, gpu_ctx, d, 160, 90, 2)


def test_headline_scene_primary_hits_bit_exact(native, oracle, gpu_ctx):
    # the 1,310,720-triangle scene at a reduced resolution (the CPU oracle
    # traces every ray)
    d = scenes.displaced_icosphere_scene(192, 108, subdiv=8)
    _primary_equal(native, oracle, gpu_ctx, d, 192, 108, 1)


@pytest.mark.parametrize("count,seed", [(1, 1), (4, 4), (5, 5), (200, 7), (2000, 8)])
def test_random_rays_bit_exact_spheres(native, oracle, gpu_ctx, count, seed):
    d = random_spheres(count, seed)
    rays = np.concatenate([random_rays(4096, seed + 1000), axis_aligned_rays()])
    sc = native.Scene(gpu_ctx, d)
    p, t, u, v = sc.trace_rays(rays)
    op, ot, _, _ = oracle.OracleScene(d).trace_rays(rays)
    assert np.array_equal(p, op)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))


def test_random_rays_bit_exact_triangles(native, oracle, gpu_ctx):
    d = random_triangles(3000, 5)
    rays = np.concatenate([random_rays(20000, 6, spread=8.0), axis_aligned_rays()])
    sc = native.Scene(gpu_ctx, d)
    p, t, u, v = sc.trace_rays(rays)
    op, ot, ou, ov = oracle.OracleScene(d).trace_rays(rays)
    assert (p >= 0).sum() > 1000
    assert np.array_equal(p, op)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    _uv_equal(d, p, u, v, ou, ov)


def test_scene_info_matches_oracle(native, oracle, gpu_ctx):
    d = scenes.displaced_icosphere_scene(64, 36, subdiv=5)
    a = native.Scene(gpu_ctx, d).info()
    b = oracle.OracleScene(d).info()
    assert (a.num_nodes, a.num_leaves, a.max_depth, a.rr_min_bounces) == (b.num_nodes, b.num_leaves, b.max_depth,
                                                                        b.rr_min_bounces)
    assert a.rr_min_survival == b.rr_min_survival and a.mean_irradiance == b.mean_irradiance
