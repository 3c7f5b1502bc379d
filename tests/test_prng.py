"""PRNG sampler mode (SamplerKind.Prng) on the oracle.

The reference's PRNG mode draws from thread-local System.Random instances
seeded in thread-start order (MathUtils.cs:18-20), so none of its values is
reproducible; this build's PRNG mode hashes (sample index, dimension, pixel
seed) instead (PCG-RXS-M-XS as a hash, Jarzynski and Olano 2020).  What the
reference fixes, and what is checked here:
  * every draw is uniform on [0, 1) with 24-bit resolution, and draws of
    different dimensions / samples / pixels are uncorrelated;
  * the camera jitter is stratified: sample s of a pixel lands in stratum
    (s mod q, s div q) of the q x q grid, q = ceil(sqrt(spp)), and the
    render takes q^2 samples (Renderer.cs:556-558, 677-682);
  * the PRNG render estimates the same image as the Sobol render
    (per-pixel agreement within 3 sigma over independent runs);
  * chunks of a PRNG render (sample_offset / sample_total, no reference
    counterpart) keep the whole render's strata and merge back into it.
"""
import numpy as np
import pytest

from raytracer import abi, scenes


def test_prng_draws_are_uniform_and_uncorrelated(oracle):
    n = 1 << 16
    idx = np.arange(n, dtype=np.uint32)
    u = oracle.prng(idx, 0, 12345)
    assert u.min() >= 0.0 and u.max() < 1.0
    # 24-bit grid: every value is k * 2^-24
    assert np.array_equal(u * 16777216.0, np.floor(u * 16777216.0))
    hist = np.bincount((u * 64).astype(int), minlength=64)
    chi2 = ((hist - n / 64) ** 2 / (n / 64)).sum()
    assert chi2 < 130          # 63 degrees of freedom: P(chi2 > 130) ~ 2e-6
    for other in (oracle.prng(idx, 1, 12345), oracle.prng(idx, 0, 12346), oracle.prng(idx + 1, 0, 12345)):
        r = np.corrcoef(u, other)[0, 1]
        assert abs(r) < 0.02   # ~5 sigma at n = 65536
    # consecutive dimensions of one sample and one dimension across pixels
    dims = oracle.prng(7, np.arange(n, dtype=np.uint32), 99)
    assert abs(np.corrcoef(dims[:-1], dims[1:])[0, 1]) < 0.02
    seeds = oracle.prng(3, 5, np.arange(n, dtype=np.uint32))
    hist = np.bincount((seeds * 64).astype(int), minlength=64)
    assert ((hist - n / 64) ** 2 / (n / 64)).sum() < 130


def _render(oracle, desc, W, H, spp, sampler, seed_xor):
    st = abi.settings_defaults(W, H, spp, 5)
    st.sampler, st.seed_xor = sampler, seed_xor
    return oracle.OracleScene(desc).render(st)


def test_prng_render_takes_the_strata_grid(oracle):
    W = H = 8
    desc = scenes.cornell(W, H)
    o = _render(oracle, desc, W, H, 10, abi.RT_SAMPLER_PRNG, 0)
    assert (o["sample_count"] == 16).all()          # ceil(sqrt(10))^2
    o = _render(oracle, desc, W, H, 9, abi.RT_SAMPLER_PRNG, 0)
    assert (o["sample_count"] == 9).all()
    # a chunk takes exactly its spp samples of the sample_total = q^2 grid;
    # without a square sample_total it is refused
    st = abi.settings_defaults(W, H, 6, 5)
    st.sampler, st.sample_offset, st.sample_total = abi.RT_SAMPLER_PRNG, 3, 16
    assert (oracle.OracleScene(desc).render(st)["sample_count"] == 6).all()
    for off, total in ((3, 0), (3, 10)):
        st.sample_offset, st.sample_total = off, total
        with pytest.raises(RuntimeError):
            oracle.OracleScene(desc).render(st)


@pytest.mark.parametrize("split", [(7, 9), (5, 6, 5), (1, 15)])
def test_prng_chunks_merge_to_the_whole_render(oracle, split):
    # spp 10 renders the 4 x 4 grid; chunks of it, each drawing sample s in
    # stratum (s mod 4, s div 4) of the whole grid, merge back by
    # sample-weighted means (a chunk that restarted the strata at its own
    # first sample would move its camera rays and miss by far more)
    from test_chunking import halves, merge
    W = H = 16
    desc = scenes.cornell(W, H)
    full = _render(oracle, desc, W, H, 10, abi.RT_SAMPLER_PRNG, 0)
    parts, offs = [], []
    off = 0
    for n in split:
        st = abi.settings_defaults(W, H, n, 5)
        st.sampler, st.sample_offset, st.sample_total = abi.RT_SAMPLER_PRNG, off, 16
        parts.append(oracle.OracleScene(desc).render(st))
        offs.append(off)
        off += n
    tol = dict(rtol=2e-5, atol=1e-6)
    assert np.allclose(merge(parts, "beauty", list(split)), full["beauty"], **tol)
    for key, h in (("beauty_a", 0), ("beauty_b", 1)):
        w = [halves(o, n)[h] for o, n in zip(offs, split)]
        keep = [i for i in range(len(parts)) if w[i] > 0]
        assert np.allclose(merge([parts[i] for i in keep], key, [w[i] for i in keep]), full[key], **tol)


def test_prng_render_estimates_the_sobol_image(oracle):
    W = H = 16
    desc = scenes.cornell(W, H)
    runs = 8
    a = np.stack([_render(oracle, desc, W, H, 16, abi.RT_SAMPLER_PRNG, 0x1000 + k)["beauty"] for k in range(runs)])
    b = np.stack([_render(oracle, desc, W, H, 16, abi.RT_SAMPLER_SOBOL, 0x2000 + k)["beauty"] for k in range(runs)])
    a, b = a.astype(np.float64), b.astype(np.float64)
    se = np.sqrt(a.var(0, ddof=1) / runs + b.var(0, ddof=1) / runs)
    z = np.abs(a.mean(0) - b.mean(0)) / np.maximum(se, 1e-6)
    print(f"within 3 sigma {(z <= 3).mean():.4f}, max z {z.max():.2f}")
    # 8 runs a side: ~14 degrees of freedom, two-sided tail beyond 3 is ~1 %
    assert (z <= 3).mean() >= 0.97
    assert z.max() <= 8.0
    # image means: per-run means on each side, 4 sigma
    ma, mb = a.mean(axis=(1, 2, 3)), b.mean(axis=(1, 2, 3))
    sm = np.sqrt(ma.var(ddof=1) / runs + mb.var(ddof=1) / runs)
    print(f"image means {ma.mean():.6f} {mb.mean():.6f}, z {abs(ma.mean() - mb.mean()) / sm:.2f}")
    assert abs(ma.mean() - mb.mean()) <= 4 * sm
