"""PRNG sampler mode on the GPU: the same hashed draws and strata as the
oracle (same streams), the strata-grid sample count, agreement with the
Sobol render over independent runs, chunks of a PRNG render, and the
refusals (non-square chunk totals, adjoint, primary-hit hook)."""
import numpy as np
import pytest

from raytracer import abi, scenes

pytestmark = pytest.mark.gpu


def _lum(x):
    x = x.astype(np.float64)
    return 0.2126 * x[0] + 0.7152 * x[1] + 0.0722 * x[2]


@pytest.mark.parametrize("scene,spp", [("cornell", 10), ("icosphere", 9)])
def test_prng_same_streams_as_oracle(native, oracle, gpu_ctx, scene, spp):
    W, H = 64, 48
    desc = scenes.cornell(W, H) if scene == "cornell" else scenes.displaced_icosphere_scene(W, H, subdiv=4)
    r = native.Renderer(native.Scene(gpu_ctx, desc), spp, 8, sampler="prng").render(W, H)
    st = abi.settings_defaults(W, H, spp, 8)
    st.sampler = abi.RT_SAMPLER_PRNG
    o = oracle.OracleScene(desc).render(st)
    q = int(np.ceil(np.sqrt(spp)))
    assert (r.buffers.sample_count == q * q).all() and (o["sample_count"] == q * q).all()
    assert r.buffers.samples_a + r.buffers.samples_b == q * q
    g, c = _lum(r.buffers.beauty), _lum(o["beauty"])
    rel = np.abs(g - c) / np.maximum(np.maximum(g, c), 1e-3)
    print(f"{scene}: within 1e-3 {(rel <= 1e-3).mean():.4f}, means {g.mean():.6f} {c.mean():.6f}", flush=True)
    assert (rel <= 1e-3).mean() >= 0.99
    assert abs(g.mean() - c.mean()) <= 2e-3 * c.mean()
    for half in ("beauty_a", "beauty_b"):
        gh, ch = _lum(getattr(r.buffers, half)), _lum(o[half])
        assert abs(gh.mean() - ch.mean()) <= 2e-3 * ch.mean()


def test_prng_estimates_the_sobol_image(native, gpu_ctx):
    W = H = 48
    sc = native.Scene(gpu_ctx, scenes.cornell(W, H))
    runs = 12
    a = np.stack([native.Renderer(sc, 16, 8, sampler="prng", seed_xor=0x100 + k).render(W, H).buffers.beauty
                  for k in range(runs)]).astype(np.float64)
    b = np.stack([native.Renderer(sc, 16, 8, seed_xor=0x200 + k).render(W, H).buffers.beauty
                  for k in range(runs)]).astype(np.float64)
    se = np.sqrt(a.var(0, ddof=1) / runs + b.var(0, ddof=1) / runs)
    z = np.abs(a.mean(0) - b.mean(0)) / np.maximum(se, 1e-6)
    print(f"within 3 sigma {(z <= 3).mean():.4f}, max z {z.max():.2f}", flush=True)
    assert (z <= 3).mean() >= 0.98
    assert z.max() <= 7.0


def test_prng_chunks_merge_and_match_the_oracle(native, oracle, gpu_ctx):
    from test_chunking import merge
    W, H = 48, 40
    desc = scenes.cornell(W, H)
    sc = native.Scene(gpu_ctx, desc)
    full = native.Renderer(sc, 10, 6, sampler="prng").render(W, H).buffers
    parts, split, off = [], (5, 6, 5), 0
    for n in split:
        b = native.Renderer(sc, n, 6, sampler="prng", sample_offset=off, sample_total=16).render(W, H).buffers
        assert (b.sample_count == n).all()
        parts.append({"beauty": b.beauty})
        off += n
    d = np.abs(merge(parts, "beauty", list(split)) - full.beauty).max()
    print(f"max |merged - whole| {d:.2e}", flush=True)
    assert np.allclose(merge(parts, "beauty", list(split)), full.beauty, rtol=2e-5, atol=1e-6)
    st = abi.settings_defaults(W, H, 6, 6)
    st.sampler, st.sample_offset, st.sample_total = abi.RT_SAMPLER_PRNG, 5, 16
    o = oracle.OracleScene(desc).render(st)
    g, c = _lum(parts[1]["beauty"]), _lum(o["beauty"])
    rel = np.abs(g - c) / np.maximum(np.maximum(g, c), 1e-3)
    print(f"chunk [5, 11): within 1e-3 {(rel <= 1e-3).mean():.4f}", flush=True)
    assert (rel <= 1e-3).mean() >= 0.99
    assert abs(g.mean() - c.mean()) <= 2e-3 * c.mean()


def test_prng_refusals(native, gpu_ctx):
    sc = native.Scene(gpu_ctx, scenes.cornell(16, 16))
    with pytest.raises(RuntimeError):
        native.Renderer(sc, 9, 4, sampler="prng", sample_offset=2).render(16, 16)
    with pytest.raises(RuntimeError):   # a chunk's sample_total must be a square
        native.Renderer(sc, 4, 4, sampler="prng", sample_offset=2, sample_total=10).render(16, 16)
    st = abi.settings_defaults(16, 16, 4, 4)
    st.sampler = abi.RT_SAMPLER_PRNG
    with pytest.raises(RuntimeError):
        sc.trace_primary(st)
