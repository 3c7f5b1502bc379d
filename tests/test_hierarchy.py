"""The oracle's object hierarchy (Group.cs:20-80, Transform.cs:74-240,
Instance.cs:54-84) pinned against closed forms and the flat object list:

  * identity wrappers hit exactly what the flat list hits (Transform's ray
    mapping with the identity is exact; Group / Instance add no arithmetic);
  * a sphere seen through a scaled, rotated, translated Transform hits where
    the float64 ray-ellipsoid intersection puts it, with the ellipsoid's
    gradient as normal (the transposed inverse: using the matrix itself, or
    the inverse untransposed, fails at non-uniform scale);
  * a Group of more than 4 finite children (a BVH) and of fewer (a list)
    hit what the flat list of the same primitives hits;
  * Instance material overrides reach the hit record;
  * the limits and refusals the device shares (wrapper nesting, geometry
    lights on transformed primitives).
"""
import numpy as np
import pytest

from raytracer import abi, scenes
import hierarchy_cases as hc


def _same_hits(a, b):
    pa, ta, ua, va = a
    pb, tb, ub, vb = b
    same = pa == pb
    return same.mean(), same


def test_identity_wrappers_hit_what_the_flat_list_hits(oracle):
    hier = oracle.OracleScene(hc.instanced_scene(identity=True))
    flat = oracle.OracleScene(hc.instanced_scene(identity=True, flat=True))
    rays = hc.random_rays(20000, 1)
    a, b = hier.trace_rays(rays), flat.trace_rays(rays)
    frac, same = _same_hits(a, b)
    print(f"same primitive: {frac:.5f}; hits {np.mean(a[0] >= 0):.3f}")
    assert np.mean(a[0] >= 0) > 0.2
    assert frac >= 0.999
    for k in (1, 2, 3):
        assert np.array_equal(a[k][same], b[k][same])


def _ellipsoid_truth(M32, rays):
    """float64 closest intersection of world rays with the unit sphere seen
    through M (row vectors), its world normal facing the ray."""
    Mi = np.linalg.inv(M32.astype(np.float64))
    o, d = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64)
    lo = o @ Mi[:3, :3] + Mi[3, :3]
    ld = d @ Mi[:3, :3]
    a = (ld * ld).sum(1)
    bh = (lo * ld).sum(1)
    c = (lo * lo).sum(1) - 1.0
    disc = bh * bh - a * c
    hit = disc > 0
    sq = np.sqrt(np.maximum(disc, 0.0))
    t0, t1 = (-bh - sq) / a, (-bh + sq) / a
    t = np.where(t0 > 1e-3, t0, t1)
    hit &= t > 1e-3
    q = lo + t[:, None] * ld
    n = q @ Mi[:3, :3].T
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(((n * d).sum(1) < 0)[:, None], n, -n)
    return hit, t, o + t[:, None] * d, n


@pytest.mark.parametrize("pose", [((1.5, 0.6, 0.9), (30.0, 45.0, 10.0), (0.3, -0.2, 0.5)),
                                  ((0.4, 0.4, 0.4), (0.0, 0.0, 0.0), (1.0, 0.5, -0.5)),
                                  ((2.0, 0.5, 1.0), (-70.0, 10.0, 120.0), (0.0, 0.0, 0.0))])
def test_transformed_sphere_is_the_ellipsoid(oracle, pose):
    M = hc.trs(*pose).astype(np.float32)
    desc = hc.ellipsoid_scene(M)
    sc = oracle.OracleScene(desc)
    rays = hc.random_rays(20000, 7, origin_box=3.0, target_box=1.5)
    r = sc.trace_records(rays)
    hit, t, p, n = _ellipsoid_truth(M, rays)
    on = r["prim"] == 0
    both = on & hit
    # the ground and the light's proxy sphere may be hit first
    occluded = hit & (r["prim"] > 0) & (r["t"] < t)
    bad = (hit & ~on & ~occluded) | (on & ~hit)
    print(f"hits {on.mean():.3f}, disagreements {bad.mean():.5f}")
    assert on.sum() > 500
    assert bad.mean() <= 1e-3
    # fp32 sphere test (quadratic cancellation at grazing incidence) against
    # float64: relative to the hit distance
    rel = np.abs(r["t"][both] - t[both]) / t[both]
    perr = np.linalg.norm(r["point"][both] - p[both], axis=1) / t[both]
    ang = np.linalg.norm(r["normal"][both] - n[both], axis=1)   # the angle, for small angles
    print(f"t rel median {np.median(rel):.1e} max {rel.max():.1e}; point/t median {np.median(perr):.1e} "
          f"max {perr.max():.1e}; normal angle median {np.median(ang):.1e} max {ang.max():.1e}")
    assert np.median(rel) <= 1e-6 and rel.max() <= 1e-4
    assert np.median(perr) <= 1e-6 and perr.max() <= 1e-4
    assert np.median(ang) <= 5e-5 and ang.max() <= 5e-3
    assert np.all(r["material"][both] == 0)


@pytest.mark.parametrize("n_children,wrap", [(12, False), (3, False), (9, True)])
def test_group_hits_what_the_flat_list_hits(oracle, n_children, wrap):
    hier, flat = hc.group_scene(n_children, wrap_group=wrap)
    rays = hc.random_rays(20000, 3)
    a = oracle.OracleScene(hier).trace_rays(rays)
    b = oracle.OracleScene(flat).trace_rays(rays)
    frac, same = _same_hits(a, b)
    print(f"{n_children} children: same primitive {frac:.5f}")
    assert frac >= 0.999
    assert np.array_equal(a[1][same], b[1][same])


def test_instance_override_reaches_the_record(oracle):
    desc = hc.instanced_scene()
    sc = oracle.OracleScene(desc)
    rays = hc.random_rays(40000, 11, origin_box=3.0, target_box=3.0)
    r = sc.trace_records(rays)
    mats = desc.mats
    m_over = 3
    assert mats["color"][m_over].tolist() == pytest.approx([0.9, 0.8, 0.1])
    template_prims = np.arange(0, 320 + 1 + 12)
    over = r["material"] == m_over
    print(f"override hits {over.sum()}, template hits {np.isin(r['prim'], template_prims).sum()}")
    assert over.sum() > 100
    assert np.all(np.isin(r["prim"][over], template_prims))
    # the same template primitives keep their own materials in the other instances
    own = np.isin(r["prim"], template_prims) & ~over
    assert own.sum() > 100
    assert set(np.unique(r["material"][own]).tolist()) <= {0, 1, 2}


def test_wrapper_nesting_limit_and_refusals(oracle):
    b = scenes.SceneBuilder()
    m = b.lambertian((0.5, 0.5, 0.5))
    n = b.node_prims(b.sphere((0.0, 0.0, 0.0), 1.0, m))
    for _ in range(5):
        n = b.node_transform(n, hc.translate(0.1, 0.0, 0.0))
    b.add_root(n)
    with pytest.raises(RuntimeError, match="nesting"):
        oracle.OracleScene(b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40.0, 1.0)))
    # an emissive primitive under a transform cannot be a geometry light
    b = scenes.SceneBuilder()
    e = b.emissive((1.0, 1.0, 1.0), 5.0)
    q = b.quad((-0.5, 2.0, -0.5), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), e)
    b.add_root(b.node_transform(b.node_prims(q), hc.translate(0.0, 1.0, 0.0)))
    d = b.build(scenes.make_camera((0, 0, 5), (0, 0, 0), 40.0, 1.0))
    assert int(d.lights["kind"][0]) == abi.RT_LIGHT_GEOMETRY
    with pytest.raises(RuntimeError, match="transformed"):
        oracle.OracleScene(d)


def test_identity_render_matches_flat(oracle):
    W, H = 40, 30
    st = abi.settings_defaults(W, H, 4, 6)
    a = oracle.OracleScene(hc.instanced_scene(W, H, identity=True)).render(st)["beauty"].astype(np.float64)
    b = oracle.OracleScene(hc.instanced_scene(W, H, identity=True, flat=True)).render(st)["beauty"].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.maximum(a, b), 1e-3)
    print(f"within 1e-3: {(rel <= 1e-3).mean():.4f}; mean rel {abs(a.mean() - b.mean()) / b.mean():.2e}")
    assert b.mean() > 0
    assert (rel <= 1e-3).mean() >= 0.98
    assert abs(a.mean() - b.mean()) <= 1e-3 * b.mean()


def _world_leaf_entries(dump):
    """Entries in the leaves of a pre-order world-BVH dump (rt_debug_bvh_dump)."""
    i, n = 0, 0
    while i < len(dump):
        if dump[i] == 1:
            c = int(dump[i + 1])
            n += c
            i += 8 + c
        else:
            i += 8
    return n


def test_host_build_world_is_the_roots(native):
    # the world list of a hierarchy scene is its roots and nothing else: 3
    # roots stay a list (no world BVH), 6 roots become a BVH over 6 objects
    M = hc.trs(*((1.5, 0.6, 0.9), (30.0, 45.0, 10.0), (0.3, -0.2, 0.5))).astype(np.float32)
    assert len(native.bvh_dump(hc.ellipsoid_scene(M))) == 0
    d = hc.instanced_scene()
    assert len(d.hierarchy[2]) == 6
    assert _world_leaf_entries(native.bvh_dump(d)) == 6
    hier, flat = hc.group_scene(12)
    assert len(native.bvh_dump(hier)) == 0            # one root: the group
    assert _world_leaf_entries(native.bvh_dump(flat)) == 12
