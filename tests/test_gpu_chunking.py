"""Chunked renders on the GPU (rt_settings.sample_offset / sample_total):
chunks merge into the whole render by sample-weighted means, a chunk follows
the oracle's chunk with the same sampler streams, and bad offsets are
refused.  No reference counterpart (resume / chunking)."""
import numpy as np
import pytest

from raytracer import abi, scenes
from test_chunking import halves, merge

pytestmark = pytest.mark.gpu


def _gpu(native, sc, W, H, spp, off=0, total=0):
    r = native.Renderer(sc, spp, 6, sample_offset=off, sample_total=total).render(W, H, capture_aovs=True)
    b = r.buffers
    return {"beauty": b.beauty, "beauty_a": b.beauty_a, "beauty_b": b.beauty_b, "albedo_a": b.albedo_a}


@pytest.mark.parametrize("split", [(5, 11), (8, 8), (1, 15), (3, 6, 7)])
def test_gpu_chunks_merge_to_the_whole_render(native, gpu_ctx, split):
    W, H = 48, 40
    sc = native.Scene(gpu_ctx, scenes.cornell(W, H))
    full = _gpu(native, sc, W, H, 16)
    parts, spps, offs = [], [], []
    off = 0
    for n in split:
        parts.append(_gpu(native, sc, W, H, n, off, 16))
        spps.append(n)
        offs.append(off)
        off += n
    tol = dict(rtol=2e-5, atol=1e-6)
    d = np.abs(merge(parts, "beauty", spps) - full["beauty"]).max()
    print(f"split {split}: max |merged - whole| {d:.2e}", flush=True)
    assert np.allclose(merge(parts, "beauty", spps), full["beauty"], **tol)
    for key, h in (("beauty_a", 0), ("beauty_b", 1), ("albedo_a", 0)):
        w = [halves(o, n)[h] for o, n in zip(offs, spps)]
        keep = [i for i in range(len(parts)) if w[i] > 0]
        assert np.allclose(merge([parts[i] for i in keep], key, [w[i] for i in keep]), full[key], **tol)


def test_gpu_chunk_matches_oracle_chunk(native, oracle, gpu_ctx):
    W, H = 64, 64
    desc = scenes.cornell(W, H)
    g = native.Renderer(native.Scene(gpu_ctx, desc), 11, 8, sample_offset=5, sample_total=16).render(W, H)
    st = abi.settings_defaults(W, H, 11, 8)
    st.sample_offset, st.sample_total = 5, 16
    o = oracle.OracleScene(desc).render(st)
    for key in ("beauty", "beauty_a", "beauty_b"):
        a = getattr(g.buffers, key).astype(np.float64)
        c = o[key].astype(np.float64)
        lum_a = 0.2126 * a[0] + 0.7152 * a[1] + 0.0722 * a[2]
        lum_c = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
        rel = np.abs(lum_a - lum_c) / np.maximum(np.maximum(lum_a, lum_c), 1e-3)
        print(f"{key}: within 1e-3 {(rel <= 1e-3).mean():.4f}", flush=True)
        assert (rel <= 1e-3).mean() >= 0.99
        assert abs(lum_a.mean() - lum_c.mean()) <= 2e-3 * lum_c.mean()


@pytest.mark.parametrize("off,total", [(-1, 0), (4, 10), (0, 3)])
def test_gpu_bad_offsets_refused(native, gpu_ctx, off, total):
    sc = native.Scene(gpu_ctx, scenes.cornell(16, 16))
    with pytest.raises(RuntimeError):
        native.Renderer(sc, 8, 4, sample_offset=off, sample_total=total).render(16, 16)
