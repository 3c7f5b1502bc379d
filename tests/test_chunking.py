"""Chunked renders (rt_settings.sample_offset / sample_total) on the oracle.

A render of samples [0, n) split into chunks [0, n1) and [n1, n) merges back
by sample-weighted means: the beauty over all samples, half A over the even
global sample indices and half B over the odd ones (RenderBuffers.cs:43-46),
the same per-sample values in another summation order (so up to float
re-association).  No reference counterpart (the reference renders in one
call); the GPU side is checked against this in test_gpu_chunking.py.
"""
import numpy as np
import pytest

from raytracer import abi, scenes


def render(oracle, desc, W, H, spp, off=0, total=0, aovs=False):
    st = abi.settings_defaults(W, H, spp, 6)
    st.sample_offset, st.sample_total = off, total
    st.capture_aovs = 1 if aovs else 0
    return oracle.OracleScene(desc).render(st, aovs=aovs)


def halves(off, spp):
    idx = np.arange(off, off + spp)
    return int((idx % 2 == 0).sum()), int((idx % 2 == 1).sum())


def merge(parts, key, weights):
    return sum(w * p[key].astype(np.float64) for p, w in zip(parts, weights)) / sum(weights)


@pytest.mark.parametrize("split", [(0, 5, 11), (0, 8, 8), (0, 1, 15)])
def test_oracle_chunks_merge_to_the_whole_render(oracle, split):
    W = H = 20
    desc = scenes.cornell(W, H)
    full = render(oracle, desc, W, H, 16, aovs=True)
    bounds = np.cumsum(split)
    parts, spps, offs = [], [], []
    for k in range(1, len(split)):
        off, n = int(bounds[k - 1]), split[k]
        parts.append(render(oracle, desc, W, H, n, off, 16, aovs=True))
        spps.append(n)
        offs.append(off)
    tol = dict(rtol=2e-5, atol=1e-6)
    assert np.allclose(merge(parts, "beauty", spps), full["beauty"], **tol)
    na = [halves(o, n)[0] for o, n in zip(offs, spps)]
    nb = [halves(o, n)[1] for o, n in zip(offs, spps)]
    keep = [i for i in range(len(parts)) if na[i] > 0]
    assert np.allclose(merge([parts[i] for i in keep], "beauty_a", [na[i] for i in keep]), full["beauty_a"], **tol)
    keep = [i for i in range(len(parts)) if nb[i] > 0]
    assert np.allclose(merge([parts[i] for i in keep], "beauty_b", [nb[i] for i in keep]), full["beauty_b"], **tol)
    keep = [i for i in range(len(parts)) if na[i] > 0]
    assert np.allclose(merge([parts[i] for i in keep], "albedo_a", [na[i] for i in keep]), full["albedo_a"], **tol)


def test_oracle_single_odd_sample_repeats_its_half(oracle):
    # one odd-indexed sample: half B holds it and half A repeats B, as half B
    # repeats A for a one-sample render from index 0
    W = H = 12
    desc = scenes.cornell(W, H)
    one = render(oracle, desc, W, H, 1, off=3, total=4)
    assert np.array_equal(one["beauty_a"], one["beauty_b"])
    assert np.array_equal(one["beauty"], one["beauty_b"])
    zero = render(oracle, desc, W, H, 1)
    assert np.array_equal(zero["beauty_a"], zero["beauty_b"])
    assert not np.array_equal(zero["beauty"], one["beauty"])
