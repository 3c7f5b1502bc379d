"""The HIP integrator end to end against the answers the mathematics fixes
(the cases of integrator_cases.py, run on the oracle in
test_integrator_closed_form.py): direct lighting through rt_render for every
MIS heuristic and light-picking strategy, and the balance / power heuristics
converging to the same image on a multi-light Cornell box."""
import pytest

import integrator_cases as IC

pytestmark = pytest.mark.gpu

ESTIMATORS = [(m, p) for m in ("balance", "power") for p in ("all", "power", "uniform")]


def _gpu_fns(native, ctx):
    scenes_cache = {}

    def scene(desc):
        sc = scenes_cache.get(id(desc))
        if sc is None:
            scenes_cache.clear()
            sc = scenes_cache[id(desc)] = (desc, native.Scene(ctx, desc))
        return sc[1]

    def render(desc, st):
        r = native.Renderer(scene(desc), st.samples_per_pixel, st.max_depth)
        r.settings = type(st).from_buffer_copy(st)
        return r.render(st.width, st.height).buffers.beauty

    def primary(desc, st):
        return scene(desc).trace_primary(st, 0, st.samples_per_pixel)
    return render, primary


@pytest.mark.parametrize("emitters", [("sphere",), ("quad",), ("area",)])
@pytest.mark.parametrize("mis", ["balance", "power"])
def test_direct_lighting_single_emitter(native, gpu_ctx, emitters, mis):
    render, primary = _gpu_fns(native, gpu_ctx)
    IC.check_direct_lighting(render, primary, emitters, mis, "all", spp=64)


@pytest.mark.parametrize("mis,pick", ESTIMATORS)
def test_direct_lighting_three_emitters(native, gpu_ctx, mis, pick):
    render, primary = _gpu_fns(native, gpu_ctx)
    IC.check_direct_lighting(render, primary, ("sphere", "quad", "area"), mis, pick, spp=64)


def test_estimator_invariance_cornell(native, gpu_ctx):
    render, _ = _gpu_fns(native, gpu_ctx)
    IC.check_estimator_invariance(IC.estimator_means(render, W=64, H=64, spp=64, runs=8))


def test_reference_caustic_tests_on_the_device(native, gpu_ctx):
    """CausticRenderTests (CausticRenderTests.cs:46-71, 258-320) through the
    HIP path: the device photon map focuses beneath the lens, and a
    caustics-on render keeps its energy within [0.85, 1.25] of the
    caustics-off one with a visible bright tail and few spikes."""
    import numpy as np
    from raytracer import abi
    import test_reference_render_tests as RT

    desc = RT.lens_scene()
    sc = native.Scene(gpu_ctx, desc)
    st = abi.settings_defaults(160, 160, 24, 8)
    st.caustics, st.caustic_photons, st.caustic_gather_radius = 1, 1_000_000, 0.0
    m = sc.photon_map(st)
    pos = m["pos"]
    assert m["n"] > 1000
    assert (np.abs(pos[:, 1]) < 0.05).sum() > 0.9 * m["n"]
    assert RT._query(pos, (0, 0, 0), 0.6) > 3 * RT._query(pos, (2.4, 0, 0), 0.6)

    def render(caustics, photons):
        r = native.Renderer(sc, 24, 8, enable_caustics=caustics, caustic_photons=photons)
        return r.render(160, 160).buffers.display.astype(np.float64)

    off, on = render(False, 0), render(True, 1_500_000)
    lum = lambda x: 0.2126 * x[0] + 0.7152 * x[1] + 0.0722 * x[2]
    lon, loff = lum(on), lum(off)
    ratio = lon.sum() / loff.sum()
    bon, boff = np.sort(lon[80:].ravel())[::-1], np.sort(loff[80:].ravel())[::-1]
    top = bon.size // 50
    print(f"device caustics: energy ratio {ratio:.4f}, spikes {(lon > 0.99).sum()}, "
          f"tail {bon[:top].mean():.4f} vs {boff[:top].mean():.4f}", flush=True)
    assert np.all(np.isfinite(on))
    assert 0.85 <= ratio <= 1.25
    assert (lon > 0.99).sum() <= 250
    assert bon[:top].mean() > 1.05 * boff[:top].mean()


@pytest.mark.parametrize("mis,pick", [("balance", "all"), ("power", "all"), ("balance", "uniform")])
def test_sky_furnace(native, gpu_ctx, mis, pick):
    render, primary = _gpu_fns(native, gpu_ctx)
    IC.check_sky_furnace(render, primary, mis, pick, spp=32)


def test_environment_light_same_streams(native, oracle, gpu_ctx):
    """A flat sky with its EnvironmentLight over occluding geometry (Cornell
    walls without a ceiling, both spheres, a sphere light): the device follows
    the oracle's paths (same sampler streams), so pixels agree to fp32
    re-association."""
    import numpy as np
    from raytracer import abi, scenes
    W = H = 64
    base = scenes.cornell(W, H)
    keep = [i for i in range(len(base.prims)) if not (2 <= i <= 3)]   # drop the ceiling triangles
    b = scenes.SceneBuilder()
    b.add_materials(base.mats)
    for i in keep:
        p = base.prims[i]
        if p["kind"] == abi.RT_PRIM_TRIANGLE:
            v = p["v"][:9].reshape(3, 3)
            b.triangle(v[0], v[1], v[2], int(p["material"]))
        elif p["kind"] == abi.RT_PRIM_SPHERE:
            b.sphere(tuple(p["v"][:3]), float(p["v"][3]), int(p["material"]))
    b.sphere_light((0.3, 1.6, 0.3), 0.1, (1.0, 0.9, 0.8), 25.0, shadow_samples=2)
    desc = b.build(base.camera, sky=(0.5, 0.7, 1.0))
    assert desc.lights["kind"][-1] == abi.RT_LIGHT_ENVIRONMENT
    st = abi.settings_defaults(W, H, 16, 6)
    render, primary = _gpu_fns(native, gpu_ctx)
    g = render(desc, st).astype(np.float64)
    c = oracle.OracleScene(desc).render(st)["beauty"].astype(np.float64)
    lum = lambda x: 0.2126 * x[0] + 0.7152 * x[1] + 0.0722 * x[2]
    lg, lc = lum(g), lum(c)
    rel = np.abs(lg - lc) / np.maximum(np.maximum(lg, lc), 1e-3)
    frac = (rel <= 1e-3).mean()
    mrel = abs(lg.mean() - lc.mean()) / lc.mean()
    print(f"environment light same streams: {frac:.4f} of pixels within 1e-3, mean rel {mrel:.2e}", flush=True)
    assert frac >= 0.97 and mrel <= 2e-3
