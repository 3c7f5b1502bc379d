"""Scenes with an object hierarchy (rt_node: Group / Transform / Instance /
camera-invisible wrappers) shared by the oracle tests and the GPU parity
tests.

Matrices are System.Numerics layout (row vectors, translation in row 4):
a point maps as [x, y, z, 1] @ M, and ComputeTransformMatrix's
scale * rotX * rotY * rotZ * translate order (SceneLoader.cs:3788-3822) is
the product of the factors below in that order.
"""
import math

import numpy as np

from raytracer import abi, scenes


def scale(sx, sy=None, sz=None):
    sy = sx if sy is None else sy
    sz = sx if sz is None else sz
    return np.diag([sx, sy, sz, 1.0])


def rot_x(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[1, 0, 0, 0], [0, c, s, 0], [0, -s, c, 0], [0, 0, 0, 1]], np.float64)


def rot_y(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[c, 0, -s, 0], [0, 1, 0, 0], [s, 0, c, 0], [0, 0, 0, 1]], np.float64)


def rot_z(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[c, s, 0, 0], [-s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)


def translate(x, y, z):
    m = np.eye(4)
    m[3, :3] = (x, y, z)
    return m


def trs(s=(1.0, 1.0, 1.0), r=(0.0, 0.0, 0.0), t=(0.0, 0.0, 0.0)):
    return scale(*s) @ rot_x(r[0]) @ rot_y(r[1]) @ rot_z(r[2]) @ translate(*t)


def ellipsoid_scene(M, W=64, H=48):
    """One unit sphere at the origin seen through Transform(M), a Lambertian
    ground and a sphere light."""
    b = scenes.SceneBuilder()
    m_obj = b.lambertian((0.7, 0.5, 0.3))
    m_gnd = b.lambertian((0.5, 0.5, 0.5))
    s = b.sphere((0.0, 0.0, 0.0), 1.0, m_obj)
    b.add_root(b.node_transform(b.node_prims(s), M))
    b.quad((-6.0, -1.5, -6.0), (0.0, 0.0, 12.0), (12.0, 0.0, 0.0), m_gnd)
    b.sphere_light((2.0, 4.0, 3.0), 0.3, intensity=30.0, shadow_samples=1)
    cam = scenes.make_camera((0.0, 1.0, 6.0), (0.0, 0.0, 0.0), 40.0, W / H)
    return b.build(cam)


def instanced_scene(W=96, H=72, identity=False, flat=False, camera_invisible=True, materials=None):
    """A template Group (an icosphere mesh, a sphere and a box of 12
    triangles, each under its own Transform) instanced several times, each
    instance under its own Transform, one with a material override and one
    camera-invisible; a ground quad and a sphere light at the top level.

    identity: every matrix is the identity.  flat: the same objects without
    any node (the flat object list of the same primitives, no transforms
    applied), for structure comparisons (only meaningful with identity)."""
    b = scenes.SceneBuilder()
    m_a = b.disney((0.8, 0.3, 0.2), roughness=0.4)
    m_b = b.lambertian((0.2, 0.6, 0.8))
    m_c = b.disney((0.9, 0.9, 0.9), metallic=1.0, roughness=0.2)
    m_over = b.lambertian((0.9, 0.8, 0.1))
    m_gnd = b.lambertian((0.5, 0.5, 0.5))
    v, f = scenes.icosphere(2)
    mesh = b.mesh(v * 0.5, f, m_a)
    sph = b.sphere((0.0, 0.0, 0.0), 0.4, m_b)
    box = b.mesh_triangles(scenes._BOX_TRIS.reshape(-1, 9) * 0.6, m_c)
    if flat:
        b.quad((-8.0, -1.0, -8.0), (0.0, 0.0, 16.0), (16.0, 0.0, 0.0), m_gnd)
        b.sphere_light((1.5, 5.0, 2.0), 0.3, intensity=40.0, shadow_samples=1)
        return b.build(scenes.make_camera((0.0, 2.5, 7.0), (0.0, 0.0, 0.0), 45.0, W / H))
    I = np.eye(4)
    n_mesh = b.node_prims(mesh)
    n_sph = b.node_prims(sph)
    n_box = b.node_prims(box)
    parts = [b.node_transform(n_mesh, I if identity else trs(t=(-0.7, 0.2, 0.0))),
             b.node_transform(n_sph, I if identity else trs(s=(1.0, 1.6, 1.0), t=(0.7, 0.3, 0.1))),
             b.node_transform(n_box, I if identity else trs(r=(20.0, 35.0, 10.0), t=(0.0, -0.2, 0.6)))]
    template = b.node_group(parts)
    poses = [((1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
             ((0.7, 0.7, 0.7), (0.0, 60.0, 0.0), (-2.4, -0.3, -0.8)),
             ((1.2, 0.8, 1.0), (15.0, -40.0, 5.0), (2.3, 0.1, -1.2)),
             ((0.5, 0.5, 0.5), (0.0, 0.0, 90.0), (0.3, 1.6, -2.5))]
    if identity:
        poses = poses[:1]
    for k, (s_, r_, t_) in enumerate(poses):
        over = m_over if k == 2 else -1
        inst = b.node_instance(template, material=over)
        b.add_root(b.node_transform(inst, I if identity else trs(s_, r_, t_),
                                    camera_invisible=camera_invisible and k == 3))
    b.quad((-8.0, -1.0, -8.0), (0.0, 0.0, 16.0), (16.0, 0.0, 0.0), m_gnd)
    b.sphere_light((1.5, 5.0, 2.0), 0.3, intensity=40.0, shadow_samples=1)
    return b.build(scenes.make_camera((0.0, 2.5, 7.0), (0.0, 0.0, 0.0), 45.0, W / H))


def group_scene(n_children, wrap_group=False):
    """Spheres in one Group node (a BVH over the finite children when more
    than 4, Group.cs:40-60) plus an infinite plane child, against the flat
    list of the same primitives (returned second)."""
    rng = np.random.default_rng(n_children)
    centres = rng.uniform(-3.0, 3.0, (n_children, 3))
    radii = rng.uniform(0.2, 0.6, n_children)

    def build(hier):
        b = scenes.SceneBuilder()
        m = b.lambertian((0.6, 0.6, 0.6))
        ids = [b.sphere(tuple(c), float(r), m) for c, r in zip(centres, radii)]
        pl = b.plane((0.0, -3.5, 0.0), (0.0, 1.0, 0.0), m)
        if hier:
            g = b.node_group([b.node_prims(i) for i in ids] + [b.node_prims(pl)])
            if wrap_group:
                g = b.node_instance(g)
            b.add_root(g)
        return b.build(scenes.make_camera((0.0, 0.0, 9.0), (0.0, 0.0, 0.0), 50.0, 1.0))

    return build(True), build(False)


def random_rays(n, seed, origin_box=4.0, target_box=2.5):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-origin_box, origin_box, (n, 3))
    o[:, 2] = rng.uniform(5.0, 9.0, n)
    t = rng.uniform(-target_box, target_box, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)
