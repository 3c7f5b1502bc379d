"""Thin-film Disney (ThinFilm.Evaluate through DisneyBsdf.EvalFresnel,
DisneyBsdf.cs:1984-1989) on the HIP path against the oracle: same sampler
streams, so pixels agree to fp32 re-association and the device's
transcendentals; the film visibly changes the image."""
import numpy as np
import pytest

from raytracer import abi, scenes

pytestmark = pytest.mark.gpu


def _scene(W, H, film=True):
    b = scenes.SceneBuilder()
    b.plane((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), b.lambertian((0.5, 0.5, 0.5)))
    kw = dict(thin_film_thickness=420.0, thin_film_ior=1.35) if film else {}
    b.sphere((-0.7, 0.5, 0.0), 0.5, b.disney((0.9, 0.9, 0.9), metallic=1.0, roughness=0.2, **kw))
    kw2 = dict(thin_film_thickness=780.0, thin_film_ior=1.9) if film else {}
    b.sphere((0.6, 0.45, 0.2), 0.45, b.disney((0.2, 0.3, 0.6), metallic=0.0, roughness=0.35, **kw2))
    b.sphere_light((0.0, 2.5, 1.5), 0.3, (1.0, 1.0, 1.0), 30.0, shadow_samples=2)
    b.directional_light((-0.3, -1.0, -0.4), (1.0, 0.95, 0.9), 2.0)
    cam = scenes.make_camera((0.0, 1.2, 3.2), (0.0, 0.4, 0.0), 45.0, W / H)
    return b.build(cam)


def test_thin_film_same_streams(native, oracle, gpu_ctx):
    W = H = 64
    st = abi.settings_defaults(W, H, 16, 5)
    lum = lambda x: 0.2126 * x[0] + 0.7152 * x[1] + 0.0722 * x[2]
    out = {}
    for film in (True, False):
        desc = _scene(W, H, film)
        r = native.Renderer(native.Scene(gpu_ctx, desc), 16, 5)
        g = r.render(W, H).buffers.beauty.astype(np.float64)
        out[film] = g
        if film:
            c = oracle.OracleScene(desc).render(st)["beauty"].astype(np.float64)
            lg, lc = lum(g), lum(c)
            rel = np.abs(lg - lc) / np.maximum(np.maximum(lg, lc), 1e-3)
            frac = (rel <= 1e-3).mean()
            mrel = abs(lg.mean() - lc.mean()) / lc.mean()
            print(f"thin film same streams: {frac:.4f} of pixels within 1e-3, mean rel {mrel:.2e}", flush=True)
            assert frac >= 0.99 and mrel <= 2e-3
    # the film recolours the spheres' highlights
    d = np.abs(out[True] - out[False]).max(0)
    assert (d > 0.01).mean() > 0.01
