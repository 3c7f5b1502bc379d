"""Any-hit (shadow-ray) queries on the 4-wide BVH against the oracle.

rt_occluded runs the traversal the shadow kernel uses for opaque scenes;
its answer must equal the existence of a closest hit on [tmin, tmax] from
the oracle's restatement of the reference traversal, ray for ray.
"""
import numpy as np
import pytest

from raytracer import scenes
from helpers import random_triangles

pytestmark = pytest.mark.gpu


def _segments(desc, n, seed, axis_frac=0.1):
    rng = np.random.default_rng(seed)
    v = desc.prims["v"]
    tri = desc.prims["kind"] == 0
    pts = v[tri][:, :3] if tri.any() else v[:, :3]
    lo, hi = pts.min(0) - 0.5, pts.max(0) + 0.5
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    k = int(axis_frac * n)
    d[:k] = 0.0
    d[np.arange(k), rng.integers(0, 3, k)] = rng.choice([-1.0, 1.0], k)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


@pytest.mark.parametrize("name", ["mixed", "icosphere", "cornell_area"])
def test_occluded_matches_oracle(native, oracle, gpu_ctx, name):
    desc = {"mixed": lambda: random_triangles(4000, 21, planes=2),
            "icosphere": lambda: scenes.displaced_icosphere_scene(64, 36, subdiv=6),
            "cornell_area": lambda: scenes.cornell(64, 64, light="area")}[name]()
    sc = native.Scene(gpu_ctx, desc)
    osc = oracle.OracleScene(desc)
    rays = _segments(desc, 20000, 7)
    for tmax in (0.05, 0.5, 2.0, 3.0e38):
        segs = np.concatenate([rays, np.full((len(rays), 1), 1e-4, np.float32),
                               np.full((len(rays), 1), tmax, np.float32)], 1)
        got = sc.occluded(segs)
        prim, _ = oracle.trace_rays_ex(osc, rays, 1e-4, tmax, False)
        want = prim >= 0
        print(f"{name} tmax={tmax}: occluded {want.mean():.3f}, mismatches {(got != want).sum()}", flush=True)
        assert np.array_equal(got, want)
