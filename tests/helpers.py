"""Scene and ray generators shared by the tests (seeded numpy streams)."""
import numpy as np

from raytracer import scenes


def random_spheres(count, seed, mesh=False):
    """BvhEquivalenceTests.RandomSpheres: centres U[-10,10]^3, r U[0.05,0.55]."""
    rng = np.random.default_rng(seed)
    b = scenes.SceneBuilder()
    m = b.lambertian((0.5, 0.5, 0.5))
    for _ in range(count):
        c = rng.uniform(-10, 10, 3)
        r = 0.05 + rng.uniform(0, 0.5)
        b.sphere(c, r, m)
    cam = scenes.make_camera((0.0, 0.0, 30.0), (0.0, 0.0, 0.0), 40.0, 1.0)
    return b.build(cam)


def random_triangles(count, seed, with_mesh=True, planes=1):
    """Random triangle soup + (optionally) a nested mesh object, a culled
    quad, a camera-invisible sphere and infinite planes."""
    rng = np.random.default_rng(seed)
    b = scenes.SceneBuilder()
    ma = b.lambertian((0.6, 0.6, 0.6))
    mb = b.disney((0.4, 0.5, 0.6), roughness=0.3)
    cen = rng.uniform(-5, 5, (count, 1, 3))
    tris = cen + rng.normal(0, 0.6, (count, 3, 3))
    b.triangles(tris.astype(np.float32), np.where(rng.random(count) < 0.5, ma, mb))
    if with_mesh:
        v = rng.uniform(-2, 2, (40, 3)) + np.array([2.0, -1.0, 1.0])
        f = rng.integers(0, 40, (60, 3))
        b.mesh(v, f, ma)
        b.mesh(rng.uniform(-1, 1, (5, 3)), np.array([[0, 1, 2], [1, 3, 4], [0, 2, 4]]), mb)
    b.quad((-3, 2, -3), (2, 0, 0), (0, 0, 2), ma, flags=1)
    b.sphere((0.5, 0.5, 0.5), 0.7, mb, flags=2)
    for k in range(planes):
        b.plane((0, -6 - k, 0), (0.01 * k, 1, 0), ma)
    cam = scenes.make_camera((0.0, 2.0, 16.0), (0.0, 0.0, 0.0), 45.0, 1.0)
    return b.build(cam)


def random_rays(n, seed, spread=20.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, (n, 3))
    d = rng.uniform(-1, 1, (n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def axis_aligned_rays():
    """BvhEquivalenceTests.Bvh_Matches_LinearList_AxisAlignedRays"""
    rays = []
    for k in np.arange(-10.0, 10.01, 1.5):
        rays += [[k, 0, -20, 0, 0, 1], [0, k, -20, 0, 0, 1], [-20, k, 0, 1, 0, 0], [k, -20, 0, 0, 1, 0],
                 [0, 0, 20, 0, 0, -1]]
    return np.array(rays, np.float32)
