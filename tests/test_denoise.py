"""CPU checks of the denoiser oracle (oracle/denoise.cpp), after the
reference's NlMeansFilterTests.cs, RegressionSolverTests.cs and
DenoiserRegressionTests.cs, plus closed-form pins of the plane operations
(no GPU needed).

The end-to-end gates render with the oracle renderer (Sobol only: the
reference's PRNG sampler is not restated), in display space (x / (1 + x)),
against a 2048-spp render of the same scene."""
import numpy as np
import pytest

from raytracer import abi, scenes


def _gauss_planes(seed, shape, mean, sigma):
    rng = np.random.default_rng(seed)
    return (mean + rng.normal(0.0, sigma, shape)).astype(np.float32)


def test_nlm_flat_noisy_plane(oracle):
    # NlMeansFilterTests.FlatNoisyPlane_VarianceReducedTenfold_WithNegligibleBias
    W = H = 96
    g = _gauss_planes(99, (1, H, W), 1.0, 0.2)
    t = _gauss_planes(100, (1, H, W), 1.0, 0.2)
    v = np.full((1, H, W), 0.04, np.float32)
    f = oracle.nlm_filter(g, v, t, 5, 3, 1.0)[0, 8:-8, 8:-8].astype(np.float64)
    assert abs(f.mean() - 1.0) < 5e-3
    assert f.var() < 0.04 / 10


def test_nlm_step_edge_contrast(oracle):
    # NlMeansFilterTests.StepEdge_ContrastPreserved
    W = H = 96
    base = np.where(np.arange(W) < W // 2, 0.2, 1.0).astype(np.float32)[None, None, :].repeat(H, 1)
    g = base + _gauss_planes(123, (1, H, W), 0.0, 0.05)
    t = base + _gauss_planes(124, (1, H, W), 0.0, 0.05)
    v = np.full((1, H, W), 0.05 ** 2, np.float32)
    f = oracle.nlm_filter(g, v, t, 5, 3, 1.0)[0]
    contrast = f[8:-8, W // 2 + 1].mean() - f[8:-8, W // 2 - 2].mean()
    assert contrast > 0.9 * 0.8


def test_nlm_constant_image_is_fixed_point(oracle):
    g = np.full((3, 20, 27), 0.375, np.float32)
    v = np.full((3, 20, 27), 0.01, np.float32)
    f = oracle.nlm_filter(g, v, g, 4, 3, 0.7, features=np.zeros((2, 20, 27), np.float32),
                          inv_bw=np.array([500.0, 100.0], np.float32))
    assert np.allclose(f, 0.375, rtol=2e-7, atol=0)


def test_binomial_and_box_smooth_closed_forms(oracle):
    c = np.full((13, 17), 0.5, np.float32)
    assert np.array_equal(oracle.binomial_smooth7(c), c)       # renormalised at the borders
    assert np.array_equal(oracle.box_smooth(c, 4), c)
    rng = np.random.default_rng(5)
    x = rng.random((13, 17)).astype(np.float32)
    for r in (2, 4):
        want = np.empty_like(x, dtype=np.float64)
        for yy in range(13):
            for xx in range(17):
                want[yy, xx] = x[max(0, yy - r):yy + r + 1, max(0, xx - r):xx + r + 1].astype(np.float64).mean()
        assert np.allclose(oracle.box_smooth(x, r), want, rtol=1e-5, atol=1e-6)
    k = np.array([1, 6, 15, 20, 15, 6, 1], np.float64)
    want = np.empty_like(x, dtype=np.float64)
    for yy in range(13):
        for xx in range(17):
            ys = [y for y in range(yy - 3, yy + 4) if 0 <= y < 13]
            xs = [z for z in range(xx - 3, xx + 4) if 0 <= z < 17]
            wy = k[[y - yy + 3 for y in ys]]
            wx = k[[z - xx + 3 for z in xs]]
            want[yy, xx] = (wy[:, None] * wx[None, :] * x[np.ix_(ys, xs)]).sum() / (wy.sum() * wx.sum())
    assert np.allclose(oracle.binomial_smooth7(x), want, rtol=1e-5, atol=1e-6)


def test_percentile_of_non_negative(oracle):
    rng = np.random.default_rng(2)
    x = rng.random(1001).astype(np.float32) * 7.0
    x[::5] = -1.0                       # sky sentinel
    v = np.sort(x[x >= 0])
    assert oracle.percentile_nonneg(x, 0.99, 1.0) == v[int(np.float32(0.99) * np.float32(v.size - 1))]
    assert oracle.percentile_nonneg(np.full(10, -1.0, np.float32), 0.99, 1.0) == 1.0


def _random_spd(rng, dim):
    a = rng.random((dim, dim)) * 2 - 1
    return a.T @ a + np.eye(dim)


@pytest.mark.parametrize("seed", [1, 42, 1337])
def test_regression_solver_random_spd(oracle, seed):
    # RegressionSolverTests.Solve_RandomSpdSystem_MatchesDirectVerification
    rng = np.random.default_rng(seed)
    g = _random_spd(rng, 8)
    x = rng.random(8) * 4 - 2
    ok, sol = oracle.regression_solve(g, g @ x, lambda_scale=0.0)
    assert ok and np.abs(sol[0] - x).max() < 1e-8


def test_regression_solver_planted_and_rank_deficient(oracle):
    # Solve_RecoverPlantedCoefficients_FromWeightedObservations
    rng = np.random.default_rng(7)
    beta = rng.random(8) * 2 - 1
    gram, rhs = np.zeros((8, 8)), np.zeros(8)
    for _ in range(200):
        f = np.concatenate([[1.0], rng.random(7) * 2 - 1])
        y = beta @ f
        w = 0.1 + rng.random()
        rhs += w * f * y
        gram += np.tril(w * np.outer(f, f))     # lower triangle authoritative
    ok, sol = oracle.regression_solve(gram, rhs, lambda_scale=1e-9)
    assert ok and np.abs(sol[0] - beta).max() < 1e-4
    # Solve_RankDeficientGram_RegularisationKeepsItFiniteAndStable
    rng = np.random.default_rng(11)
    gram, rhs = np.zeros((8, 8)), np.zeros(8)
    for _ in range(100):
        f = np.concatenate([[1.0], rng.random(6), [0.0]])
        f[7] = f[1]
        y = 3.0 * f[1] + 0.5
        rhs += f * y
        gram += np.tril(np.outer(f, f))
    ok, sol = oracle.regression_solve(gram, rhs)
    assert ok and np.isfinite(sol).all() and np.abs(sol).max() < 100
    # not positive definite even after regularisation -> false (the caller's weighted-mean fallback)
    ok, _ = oracle.regression_solve(-np.eye(8), np.ones(8))
    assert not ok


class _Buffers:
    pass


def _render(osc, W, H, spp, seed=0):
    st = abi.settings_defaults(W, H, spp, 8)
    st.capture_aovs = 1
    st.seed_xor = seed
    r = osc.render(st, aovs=True)
    b = _Buffers()
    for k in ("beauty", "beauty_a", "beauty_b", "albedo_a", "albedo_b", "normal_a", "normal_b", "depth_a",
              "depth_b"):
        setattr(b, k, r[k])
    b.samples_a, b.samples_b = (spp + 1) // 2, spp // 2
    return b


@pytest.fixture(scope="module")
def cornell48(oracle):
    osc = oracle.OracleScene(scenes.cornell(48, 48))
    return osc, _render(osc, 48, 48, 2048).beauty


def _mse(x, ref):
    tm = lambda a: np.clip(a, 0, None) / (1 + np.clip(a, 0, None))   # noqa: E731
    return float(((tm(x) - tm(ref)) ** 2).mean())


def test_denoiser_cuts_error_at_low_spp(oracle, cornell48):
    # DenoiserRegressionTests.Denoise_LowSppRender_CutsMseAgainstConvergedReference
    osc, ref = cornell48
    noisy = _render(osc, 48, 48, 8, seed=7)
    m0 = _mse(noisy.beauty, ref)
    nlm = _mse(oracle.denoise(noisy, "nlm", "high"), ref)
    nfor = _mse(oracle.denoise(noisy, "nfor", "high"), ref)
    fast = _mse(oracle.denoise(noisy, "nfor", "fast"), ref)
    print(f"display MSE noisy {m0:.3e}: nlm {nlm / m0:.3f}x nfor {nfor / m0:.3f}x fast {fast / m0:.3f}x", flush=True)
    assert nlm < 0.80 * m0 and nfor < 0.80 * m0 and fast < 0.80 * m0


def test_nfor_never_regresses_near_converged_sobol(oracle, cornell48):
    # Denoise_NearConvergedSobolRender_NeverRegresses
    osc, ref = cornell48
    noisy = _render(osc, 48, 48, 32, seed=7)
    assert _mse(oracle.denoise(noisy, "nfor", "high"), ref) < 1.05 * _mse(noisy.beauty, ref)
