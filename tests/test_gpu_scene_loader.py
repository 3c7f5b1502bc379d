"""Scenes loaded from YAML (raytracer/scene_loader.py) through the GPU
renderer: primary hits bit-exact against the oracle on the reference's own
caustic-lens test scene, the plain-method Cornell config, the OBJ mesh
scene (the reference's model files, smooth and flat) and the image-texture
scene (the reference's texture files, footprint-filtered camera hits), and
the radiance against the oracle on the same sample stream."""
import os

import numpy as np
import pytest

from raytracer import abi, native, scene_loader

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


@pytest.mark.parametrize("name", ["caustic-lens.yaml", "cornell-c1.yaml", "mesh-trio.yaml", "textured.yaml",
                                  "templates.yaml"])
def test_yaml_scene_primary_hits_and_radiance(gpu_ctx, oracle, name):
    W, H, spp = 96, 72, 8
    desc = scene_loader.load(os.path.join(HERE, name), W, H)
    st = abi.settings_defaults(W, H, spp, 8)
    sc = native.Scene(gpu_ctx, desc)
    try:
        p, t, _, _ = sc.trace_primary(st, 0, 2)
        osc = oracle.OracleScene(desc)
        op, ot, _, _, _, _ = osc.trace_primary(st, 0, 2)
        print(f"{name}: rays {p.size}, hits {(p >= 0).sum()}, prim mismatch {(p != op).sum()}", flush=True)
        assert np.array_equal(p, op) and np.array_equal(t.view(np.uint32), ot.view(np.uint32))
        img = native.Renderer(sc, spp, 8).render(W, H).buffers.beauty
    finally:
        sc.close()
    ref = osc.render(st)["beauty"]
    # same Sobol stream on both sides: per-image relative difference of the means
    rel = abs(float(img.mean()) - float(ref.mean())) / max(float(ref.mean()), 1e-6)
    print(f"{name}: mean radiance gpu {img.mean():.6f} oracle {ref.mean():.6f} rel {rel:.2e}", flush=True)
    assert img.mean() > 0 and rel < 1e-3
