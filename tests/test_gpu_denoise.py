"""GPU denoiser (rt_denoise) against the oracle (oracle/denoise.cpp) on the
same film planes.  Both follow the reference's fp32 operation order, so the
denoised beauty must be bit-identical; the films come from GPU renders with
AOVs (odd sizes, sky pixels for the depth sentinel, several spp)."""
import numpy as np
import pytest

from raytracer import denoise, native, scenes

pytestmark = pytest.mark.gpu


def _film(gpu_ctx, desc, W, H, spp):
    sc = native.Scene(gpu_ctx, desc)
    try:
        return native.Renderer(sc, spp, 6).render(W, H, capture_aovs=True).buffers
    finally:
        sc.close()


def _check(got, want, what):
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))
    print(f"{what}: {bad[0].size} differing values of {got.size}; max |diff| "
          f"{float(np.abs(got.astype(np.float64) - want).max()):.3e}", flush=True)
    assert bad[0].size == 0


@pytest.mark.parametrize("W,H,spp", [(37, 29, 8), (64, 48, 3), (150, 70, 4)])
@pytest.mark.parametrize("kind,quality", [("nlm", "high"), ("nlm", "fast"), ("nfor", "high"), ("nfor", "fast")])
def test_matches_oracle_bit_exact(gpu_ctx, oracle, W, H, spp, kind, quality):
    # 150 x 70 spans several 64-pixel splat tiles in both directions
    buf = _film(gpu_ctx, scenes.cornell(W, H), W, H, spp)
    got = denoise.denoise(buf, denoise.DenoiserOptions(kind, quality), gpu_ctx)
    want = oracle.denoise(buf, kind, quality)
    _check(got, want, f"{kind} {quality} {W}x{H} {spp} spp")


@pytest.mark.parametrize("kind", ["nlm", "nfor"])
def test_with_sky_and_one_sample(gpu_ctx, oracle, kind):
    # an open scene (sky pixels: depth -1) and spp = 1 (samples_b = 0)
    W, H = 40, 24
    desc = scenes.displaced_icosphere_scene(W, H, subdiv=3, sphere_lights=2)
    for spp in (1, 4):
        buf = _film(gpu_ctx, desc, W, H, spp)
        assert (buf.depth_a < 0).any()
        got = denoise.denoise(buf, denoise.DenoiserOptions(kind, "fast"), gpu_ctx)
        _check(got, oracle.denoise(buf, kind, "fast"), f"{kind} sky {spp} spp")


def test_nfor_selection_margin(gpu_ctx, oracle):
    W, H = 33, 31
    buf = _film(gpu_ctx, scenes.cornell(W, H), W, H, 6)
    for margin in (0.0, 0.5, 2.0):
        got = denoise.denoise(buf, denoise.DenoiserOptions("nfor", "high", margin), gpu_ctx)
        _check(got, oracle.denoise(buf, "nfor", "high", margin), f"nfor margin {margin}")


def test_denoise_argument_errors(gpu_ctx):
    buf = _film(gpu_ctx, scenes.cornell(16, 16), 16, 16, 2)
    buf.albedo_a = None
    with pytest.raises(RuntimeError):
        denoise.denoise(buf, denoise.DenoiserOptions("nlm"), gpu_ctx)
    with pytest.raises(ValueError):
        denoise.DenoiserOptions("bilateral")
