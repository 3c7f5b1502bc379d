"""Binned-SAH BVH and closest-hit traversal (BvhNode.cs:91-363, HittableList.cs).

CPU checks: the product's host-built, flattened tree is node-for-node the
oracle's restatement of the reference tree; the oracle's BVH traversal
agrees with a brute-force linear list (BvhEquivalenceTests.cs:54-143).
"""
import numpy as np
import pytest

from helpers import axis_aligned_rays, random_rays, random_spheres, random_triangles
from raytracer import scenes


@pytest.mark.parametrize("count,seed", [(5, 5), (20, 6), (200, 7), (2000, 8)])
def test_product_tree_equals_reference_tree_spheres(native, oracle, count, seed):
    d = random_spheres(count, seed)
    a = native.bvh_dump(d)
    b = oracle.OracleScene(d).bvh_dump()
    assert len(a) > 0 and np.array_equal(a, b)


def test_product_tree_equals_reference_tree_mixed(native, oracle):
    d = random_triangles(3000, 11)
    assert np.array_equal(native.bvh_dump(d), oracle.OracleScene(d).bvh_dump())


def test_product_tree_equals_reference_tree_displaced_icosphere(native, oracle):
    d = scenes.displaced_icosphere_scene(64, 36, subdiv=5)
    assert np.array_equal(native.bvh_dump(d), oracle.OracleScene(d).bvh_dump())


def test_clustered_centroids_build(native, oracle):
    # BvhEquivalenceTests.Bvh_Matches_LinearList_ClusteredPrimitives
    b = scenes.SceneBuilder()
    m = b.lambertian((0.5, 0.5, 0.5))
    for i in range(16):
        b.sphere((0.0, 0.0, 0.0), 0.5 + i * 0.01, m)
    d = b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40, 1))
    assert np.array_equal(native.bvh_dump(d), oracle.OracleScene(d).bvh_dump())
    s = oracle.OracleScene(d)
    rays = random_rays(64, 123)
    pb, tb = oracle.trace_rays_ex(s, rays, 0.001, 1e30, False)
    pl, tl = oracle.trace_rays_ex(s, rays, 0.001, 1e30, True)
    assert np.array_equal(pb >= 0, pl >= 0)
    assert np.all(np.abs(tb - tl)[pb >= 0] <= 1e-4)


@pytest.mark.parametrize("count,seed", [(1, 1), (2, 2), (3, 3), (4, 4), (5, 5), (20, 6), (200, 7), (2000, 8)])
def test_oracle_bvh_matches_linear_list(oracle, count, seed):
    # BvhEquivalenceTests.Bvh_Matches_LinearList
    s = oracle.OracleScene(random_spheres(count, seed))
    rays = random_rays(256, seed + 1000)
    pb, tb = oracle.trace_rays_ex(s, rays, 0.001, 1e30, False)
    pl, tl = oracle.trace_rays_ex(s, rays, 0.001, 1e30, True)
    assert np.array_equal(pb >= 0, pl >= 0)
    assert np.all(np.abs(tb - tl)[pb >= 0] <= 1e-4)


def test_oracle_bvh_axis_aligned_and_inside_rays(oracle):
    s = oracle.OracleScene(random_spheres(50, 42))
    for rays in (axis_aligned_rays(), np.concatenate([np.zeros((64, 3)), random_rays(64, 321)[:, 3:]], 1)):
        pb, tb = oracle.trace_rays_ex(s, rays, 0.001, 1e30, False)
        pl, tl = oracle.trace_rays_ex(s, rays, 0.001, 1e30, True)
        assert np.array_equal(pb >= 0, pl >= 0)
        assert np.all(np.abs(tb - tl)[pb >= 0] <= 1e-4)


def test_oracle_triangles_match_linear_list(oracle):
    s = oracle.OracleScene(random_triangles(2000, 3, with_mesh=True))
    rays = random_rays(2000, 4)
    pb, tb = oracle.trace_rays_ex(s, rays, 1e-4, 3.4e38, False)
    pl, tl = oracle.trace_rays_ex(s, rays, 1e-4, 3.4e38, True)
    assert np.array_equal(pb >= 0, pl >= 0)
    assert np.all(np.abs(tb - tl)[pb >= 0] <= 1e-4)


def test_moller_trumbore_closed_form(oracle):
    # a ray through a known barycentric point of one triangle hits it at the
    # analytic distance with the analytic (u, v); edges u+v=1 and t=tmin/tmax
    b = scenes.SceneBuilder()
    m = b.lambertian((0.5, 0.5, 0.5))
    b.triangle(np.array([0, 0, 0], np.float32), np.array([2, 0, 0], np.float32), np.array([0, 2, 0], np.float32), m)
    s = oracle.OracleScene(b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40, 1)))
    rays = np.array([[0.5, 0.5, 3.0, 0, 0, -1], [1.0, 1.0, 3.0, 0, 0, -1], [1.5, 1.0, 3.0, 0, 0, -1],
                     [0.5, 0.5, -3.0, 0, 0, 1], [0.5, 0.5, 3.0, 1, 0, 0]], np.float32)
    prim, t, u, v = s.trace_rays(rays)
    assert prim.tolist() == [0, 0, -1, 0, -1]
    assert t[0] == 3.0 and u[0] == 0.25 and v[0] == 0.25
    assert t[1] == 3.0 and u[1] + v[1] == 1.0
    assert t[3] == 3.0
    p2, t2 = oracle.trace_rays_ex(s, rays[:1], 1e-4, 3.0, True)   # t == tMax is accepted
    assert p2[0] == 0
    p3, _ = oracle.trace_rays_ex(s, rays[:1], 1e-4, 2.9999998, True)
    assert p3[0] == -1


def test_equal_t_later_primitive_wins(oracle):
    # two coincident triangles: the later one in list order wins (t <= closest)
    b = scenes.SceneBuilder()
    m = b.lambertian((0.5, 0.5, 0.5))
    tri = [np.array(x, np.float32) for x in ([0, 0, 0], [2, 0, 0], [0, 2, 0])]
    b.triangle(*tri, m)
    b.triangle(*tri, m)
    s = oracle.OracleScene(b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40, 1)))
    prim, t, u, v = s.trace_rays(np.array([[0.5, 0.5, 3.0, 0, 0, -1]], np.float32))
    assert prim[0] == 1
