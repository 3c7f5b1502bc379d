"""Owen-scrambled Sobol sampler (OwenSobol.cs:48-210, Sampler.cs:75-118).

Pins: an independent numpy restatement of the published construction
(Sobol dim 0 = van der Corput, dim 1 from the Bratley–Fox recurrence,
Laine–Karras/Burley nested Owen scrambling), and the reference's own
invariants (SamplerTests.cs).

Note: the reference's comment names p(x) = x^2 + x + 1 for dimension 1, but
its tabulated direction numbers (OwenSobol.cs:132-142: 0x80000000,
0xc0000000, 0xa0000000, 0xf0000000, 0x88000000, ...) are those of the
degree-1 primitive polynomial x + 1 (m_k = 1, 3, 5, 15, 17, ...), the usual
second Sobol dimension.  The table is what the code uses, so that is what is
restated and pinned here.
"""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def _u32(x):
    return np.asarray(x, np.uint64) & M32


def np_reverse(v):
    v = _u32(v)
    v = ((v & 0xaaaaaaaa) >> 1) | ((v & 0x55555555) << 1)
    v = _u32(v)
    v = ((v & 0xcccccccc) >> 2) | ((v & 0x33333333) << 2)
    v = _u32(v)
    v = ((v & 0xf0f0f0f0) >> 4) | ((v & 0x0f0f0f0f) << 4)
    v = _u32(v)
    v = ((v & 0xff00ff00) >> 8) | ((v & 0x00ff00ff) << 8)
    v = _u32(v)
    return _u32((v >> 16) | (v << 16))


def np_hash(x):
    x = _u32(x)
    x ^= x >> 16
    x = _u32(x * 0x7feb352d)
    x ^= x >> 15
    x = _u32(x * 0x846ca68b)
    x ^= x >> 16
    return x


def np_owen(v, seed):
    v = np_reverse(v)
    v = _u32(v + seed)
    for c in (0x6c50b47c, 0xb82f1e52, 0xc7afe638, 0x8d22f6e6):
        v = _u32(v ^ _u32(v * c))
    return np_reverse(v)


def joe_kuo_dim1_directions():
    """Direction numbers v_k = m_k << (32 - k) of the primitive polynomial
    x + 1 (degree s = 1) from the Bratley–Fox recurrence
    m_k = 2 m_{k-1} ^ m_{k-1}, m_1 = 1."""
    m = [0, 1]
    for k in range(2, 33):
        m.append((2 * m[k - 1]) ^ m[k - 1])
    return [m[k] << (32 - k) for k in range(1, 33)]


def np_sobol(index, dim, seed):
    index = _u32(index)
    dim = _u32(dim)
    seed = _u32(seed)
    dirs = np.array(joe_kuo_dim1_directions(), np.uint64)
    v1 = np.zeros_like(index)
    for k in range(32):
        bit = (index >> np.uint64(k)) & np.uint64(1)
        v1 = v1 ^ (bit * dirs[k])
    v0 = np_reverse(index)
    vh = np_reverse(np_owen(index, np_hash(_u32(dim * 0x68bc21eb + 0xc7afe638))))
    v = np.where(dim == 0, v0, np.where(dim == 1, v1, vh))
    scr = np_hash(seed ^ np_hash(_u32(dim * 0xa511e9b3 + 0x736f6c34)))
    v = np_owen(v, scr)
    return (v >> np.uint64(8)).astype(np.float64) * (1.0 / 16777216.0)


def test_recurrence_matches_pascal_closed_form():
    # the dim-1 Sobol matrix is the Pascal triangle mod 2 (d ^= d >> 1)
    d = 0x80000000
    closed = []
    for _ in range(32):
        closed.append(d)
        d ^= d >> 1
    assert joe_kuo_dim1_directions() == closed
    # first entries of the reference's table (OwenSobol.cs:134-135)
    assert closed[:8] == [0x80000000, 0xc0000000, 0xa0000000, 0xf0000000,
                          0x88000000, 0xcc000000, 0xaa000000, 0xff000000]


def test_oracle_matches_independent_numpy_restatement(oracle):
    rng = np.random.default_rng(7)
    idx = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64)
    idx[:64] = np.arange(64)
    dim = rng.integers(0, 300, 20000, dtype=np.uint64)
    dim[:300] = np.arange(300)
    seed = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64)
    got = oracle.sobol(idx.astype(np.uint32), dim.astype(np.uint32), seed.astype(np.uint32)).astype(np.float64)
    want = np_sobol(idx, dim, seed)
    assert np.array_equal(got, want)


def test_draws_stay_in_unit_interval(oracle):
    # SamplerTests.Sobol_DrawsStayInUnitInterval
    pix, smp, dim = np.meshgrid(np.arange(16), np.arange(256), np.arange(32), indexing="ij")
    seed = (pix.astype(np.uint64) * 0xdeadbeef) & 0xFFFFFFFF
    u = oracle.sobol(smp.ravel().astype(np.uint32), dim.ravel().astype(np.uint32), seed.ravel().astype(np.uint32))
    assert u.min() >= 0.0 and u.max() <= 1.0 - 1e-7


def test_pixel_seed_decorrelates(oracle):
    a = oracle.sobol(np.arange(64, dtype=np.uint32), 0, 0xa5a5a5a5)
    b = oracle.sobol(np.arange(64, dtype=np.uint32), 0, 0x5a5a5a5a)
    assert int((a == b).sum()) <= 1


def test_dims_0_1_are_01_nets_at_16(oracle):
    # SamplerTests.Sobol_StratifiesEachDimension1D
    for dim in (0, 1):
        u = oracle.sobol(np.arange(16, dtype=np.uint32), dim, 0)
        cells = np.minimum((u * 16).astype(int), 15)
        assert sorted(cells) == list(range(16))


def test_dims_0_1_cover_2d_cells(oracle):
    # SamplerTests.Sobol_StratifiesUnitSquare_BetterThanPrng
    u = oracle.sobol(np.arange(64, dtype=np.uint32), 0, 0)
    v = oracle.sobol(np.arange(64, dtype=np.uint32), 1, 0)
    cells = set(zip(np.minimum((u * 8).astype(int), 7), np.minimum((v * 8).astype(int), 7)))
    assert len(cells) >= 45


def test_dims_0_1_form_a_022_net(oracle):
    # Sobol (0,2,2)-net: every 2^a x 2^b grid with a+b = 6 holds one point of 64
    for seed in (0, 12345, 0xdeadbeef):
        u = oracle.sobol(np.arange(64, dtype=np.uint32), 0, seed)
        v = oracle.sobol(np.arange(64, dtype=np.uint32), 1, seed)
        for a in range(7):
            b = 6 - a
            cu = np.minimum((u * 2 ** a).astype(int), 2 ** a - 1)
            cv = np.minimum((v * 2 ** b).astype(int), 2 ** b - 1)
            assert len(set(zip(cu, cv))) == 64, (seed, a, b)


def test_high_dims_are_01_nets_at_64(oracle):
    # SamplerTests.Sobol_StratifiesEachDimension1D_AtHighDims
    for dim in range(2, 16):
        u = oracle.sobol(np.arange(64, dtype=np.uint32), dim, 0)
        cells = np.minimum((u * 64).astype(int), 63)
        assert sorted(cells) == list(range(64)), dim
