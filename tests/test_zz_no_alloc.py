"""rt_occluded_noalloc, the allocation-free variant of rt_occluded (no
reference counterpart; for callers that cannot allocate device memory):
its staging lives in the library's static per-GPU array, so its answers
must equal rt_occluded's ray for ray, on two calls in a row (the second
smaller, so stale contents of the first would show), and it refuses more
segments than that array holds with a clear error."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_triangles
from raytracer import abi, scenes

pytestmark = pytest.mark.gpu

MAX_SEGMENTS = 1 << 20


def _segs(desc, n, seed, tmax):
    rng = np.random.default_rng(seed)
    v = desc.prims["v"]
    tri = desc.prims["kind"] == 0
    pts = v[tri][:, :3]
    lo, hi = pts.min(0) - 0.5, pts.max(0) + 0.5
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.stack([np.full(n, 1e-4), rng.uniform(0.05, tmax, n)], 1)
    return np.ascontiguousarray(np.concatenate([o, d, t], 1), np.float32)


def _noalloc(native, ctx, sc, segs):
    f = native.load().rt_occluded_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_int32)]
    out = np.full(len(segs), -7, np.int32)
    rc = f(ctx.h, sc.h, segs.ctypes.data_as(C.POINTER(C.c_float)), len(segs), out.ctypes.data_as(C.POINTER(C.c_int32)))
    return rc, out


@pytest.mark.parametrize("name", ["mixed", "icosphere"])
def test_noalloc_matches_rt_occluded(native, name):
    ctx = native.Context(0)
    desc = {"mixed": lambda: random_triangles(4000, 23, planes=2),
            "icosphere": lambda: scenes.displaced_icosphere_scene(64, 36, subdiv=6)}[name]()
    sc = native.Scene(ctx, desc)
    for n, seed in ((600_000, 5), (70_000, 6)):
        segs = _segs(desc, n, seed, 8.0)
        print(f"{name}: rt_occluded_noalloc on {n} segments", flush=True)
        want = sc.occluded(segs).astype(np.int32)
        rc, got = _noalloc(native, ctx, sc, segs)
        assert rc == abi.RT_OK, native.load().rt_last_error(ctx.h)
        print(f"  occluded {want.mean():.3f}, mismatches {(got != want).sum()}", flush=True)
        assert np.array_equal(got, want)
    sc.close()
    ctx.close()


def test_noalloc_refuses_more_than_its_staging(native):
    ctx = native.Context(0)
    desc = random_triangles(200, 3, with_mesh=False, planes=0)
    sc = native.Scene(ctx, desc)
    segs = np.zeros((MAX_SEGMENTS + 1, 8), np.float32)
    segs[:, 4] = 1.0
    segs[:, 7] = 1.0
    rc, out = _noalloc(native, ctx, sc, segs)
    assert rc == abi.RT_ERR_INVALID_ARGUMENT
    assert b"at most 1048576" in native.load().rt_last_error(ctx.h)
    assert np.all(out == -7)
    rc, out = _noalloc(native, ctx, sc, segs[:MAX_SEGMENTS])
    assert rc == abi.RT_OK
    sc.close()
    ctx.close()
