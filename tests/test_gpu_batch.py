"""Batched rendering (rt_render_batch): many (scene, settings) jobs in one
path pool.  Each job's film must be bit-identical to rendering that job on
its own — same per-sample values, same per-pixel summation order — which is
what makes "batched vs plain at equal RMSE" hold by construction."""
import numpy as np
import pytest

from raytracer import scenes

pytestmark = pytest.mark.gpu

PLANES = ("beauty", "beauty_a", "beauty_b", "display")


def _jobs(native, ctx):
    c1 = native.Scene(ctx, scenes.cornell(48, 40))
    c2 = native.Scene(ctx, scenes.cornell(32, 32, light="area"))
    d = scenes.displaced_icosphere_scene(40, 24, subdiv=4)
    ico, ico2 = native.Scene(ctx, d), native.Scene(ctx, d)
    return [
        (native.Renderer(c1, 6, 8), 48, 40, True),
        (native.Renderer(ico, 4, 4, light_sampling_strategy="power"), 40, 24, False),
        (native.Renderer(c2, 3, 12), 32, 32, False),
        (native.Renderer(c1, 2, 0), 48, 40, False),          # depth 0: black
        (native.Renderer(ico2, 5, 6, seed_xor=7), 40, 24, True),
    ]


@pytest.mark.parametrize("pool", [0, 997, 64])
def test_batch_films_equal_single_renders(native, gpu_ctx, pool):
    items = _jobs(native, gpu_ctx)
    single = [r.render(w, h, capture_aovs=a) for r, w, h, a in items]
    batch, total = native.render_batch(items, pool_paths=pool)
    print(f"pool={pool} samples={total.camera_samples} rays={total.rays_closest}", flush=True)
    assert total.camera_samples == sum(w * h * r.settings.samples_per_pixel for r, w, h, _ in items)
    for k, (one, b) in enumerate(zip(single, batch)):
        for p in PLANES + (("albedo_a", "normal_b", "depth_a") if items[k][3] else ()):
            assert np.array_equal(getattr(one.buffers, p), getattr(b.buffers, p)), (k, p)
    assert batch[3].buffers.beauty.max() == 0.0


def test_batch_rejects_caustics(native, gpu_ctx):
    sc = native.Scene(gpu_ctx, scenes.cornell(16, 16))
    r = native.Renderer(sc, 2, 4, enable_caustics=True, caustic_photons=1000)
    with pytest.raises(native.RtError):
        native.render_batch([(r, 16, 16)])


def test_graph_replay_gives_the_same_films(native, gpu_ctx):
    items = _jobs(native, gpu_ctx)[:3]
    direct = [r.render(w, h, capture_aovs=a) for r, w, h, a in items]
    gpu_ctx.tune(bounce_graph=2)
    try:
        graphed = [r.render(w, h, capture_aovs=a) for r, w, h, a in items]
        batch, _ = native.render_batch(items, pool_paths=997)
    finally:
        gpu_ctx.tune(bounce_graph=1)
    for one, g, b in zip(direct, graphed, batch):
        for p in PLANES:
            assert np.array_equal(getattr(one.buffers, p), getattr(g.buffers, p)), p
            assert np.array_equal(getattr(one.buffers, p), getattr(b.buffers, p)), p
