"""Photon-map oracle checks (CausticPhotonTracer.cs, PhotonMap.cs).

The hash grid is pinned against an independent numpy restatement of the
Teschner hash and a stable counting sort; the k-nearest gather against brute
force (PhotonMapTests' exactness claim); the emission normalisation and the
specular walk against a closed form (a white mirror turning a sun beam onto a
wall deposits E x projected mirror area).
"""
import numpy as np
import pytest

from raytracer import abi, scenes


def _np_keys(pos, origin, cell, nbuckets):
    inv = np.float32(1.0) / np.float32(cell)
    c = np.floor((pos - origin.astype(np.float32)) * inv).astype(np.int64)
    u = (c & 0xFFFFFFFF).astype(np.uint64)
    h = ((u[:, 0] * 73856093) ^ (u[:, 1] * 19349663) ^ (u[:, 2] * 83492791)) & 0xFFFFFFFF
    return (h & (nbuckets - 1)).astype(np.int32)


def _cloud(rng, n):
    a = rng.normal(0.0, 0.05, (n // 2, 3)) + np.array([0.3, 0.0, -0.2])
    b = rng.uniform(-1.0, 1.0, (n - n // 2, 3))
    return np.concatenate([a, b]).astype(np.float32)


@pytest.mark.parametrize("n,cell", [(1, 0.1), (1000, 0.05), (20000, 0.01), (4097, 0.2)])
def test_grid_is_teschner_hash_with_stable_counting_sort(oracle, n, cell):
    rng = np.random.default_rng(n)
    pos = _cloud(rng, n)
    m = oracle.PhotonMap.from_positions(pos, cell)
    nb = 1
    while nb < n:
        nb <<= 1
    assert m.n_buckets == nb
    origin = pos.min(0)
    assert np.array_equal(m.grid[:3], origin)
    keys = _np_keys(pos, origin, m.grid[3], nb)
    assert np.array_equal(m.key, keys)
    # CSR = stable sort by bucket (ties keep photon order)
    order = np.argsort(keys, kind="stable").astype(np.int32)
    assert np.array_equal(m.idx, order)
    counts = np.bincount(keys, minlength=nb)
    assert np.array_equal(m.start, np.concatenate([[0], np.cumsum(counts)]).astype(np.int32))


def test_knn_gather_matches_brute_force(oracle):
    rng = np.random.default_rng(7)
    pos = _cloud(rng, 30000)
    cell, max_r, k = 0.02, 0.04, 40
    m = oracle.PhotonMap.from_positions(pos, cell)
    q = np.concatenate([pos[rng.integers(0, len(pos), 300)] + rng.normal(0, 0.01, (300, 3)),
                        rng.uniform(-1.2, 1.2, (300, 3))]).astype(np.float32)
    cnt, rad, idx = m.gather(q, k, max_r)
    full = 0
    for i in range(len(q)):
        d = pos - q[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        cand = np.nonzero(d2 <= np.float32(max_r) * np.float32(max_r))[0]
        best = cand[np.argsort(d2[cand], kind="stable")][:k]
        assert cnt[i] == len(best)
        assert set(idx[i, :cnt[i]].tolist()) == set(best.tolist())
        if len(best):
            assert rad[i] == np.float32(np.sqrt(d2[best].max()))
        full += cnt[i] == k
    assert full > 100   # dense queries fill the heap


def _mirror_wall_scene(budget_light_intensity=2.0):
    b = scenes.SceneBuilder()
    mirror = b.disney((1.0, 1.0, 1.0), metallic=1.0, roughness=0.0, specular=0.5)
    wall = b.lambertian((0.5, 0.5, 0.5))
    # 1x1 mirror at 45 degrees around (0,1,0): a downward beam leaves along +x
    s = np.float32(1.0 / np.sqrt(2.0))
    b.quad((-0.5 * s, 1.0 + 0.5 * s, -0.5), (s, -s, 0.0), (0.0, 0.0, 1.0), mirror)
    b.quad((3.0, 0.0, -1.0), (0.0, 2.0, 0.0), (0.0, 0.0, 2.0), wall)
    b.directional_light((0.0, -1.0, 0.0), intensity=budget_light_intensity)
    cam = scenes.make_camera((0.0, 1.0, 6.0), (0.0, 1.0, 0.0), 40.0, 1.0)
    return b.build(cam)


def test_mirror_turns_the_sun_beam_into_wall_photons(oracle):
    E = 2.0
    sc = oracle.OracleScene(_mirror_wall_scene(E))
    m = sc.photon_map(200000, 0.05)
    assert m.emitted == 200000
    assert m.n > 5000
    # every photon went light -> mirror -> wall
    assert np.allclose(m.pos[:, 0], 3.0, atol=1e-3)
    # the near-delta GGX lobe (alpha at the delta threshold) keeps a thin tail
    dev = np.abs(m.inc - np.array([-1.0, 0.0, 0.0], np.float32))
    assert np.median(dev) < 1e-3 and np.all(m.inc[:, 0] < -0.97)
    total = m.power.sum(0)
    proj = 1.0 / np.sqrt(2.0)   # mirror area seen along the beam
    assert np.allclose(total, E * proj, rtol=0.03), total


def test_budget_split_by_light_power(oracle):
    b = scenes.SceneBuilder()
    glass = b.dielectric(1.5)
    floor = b.lambertian((0.7, 0.7, 0.7))
    b.sphere((0.0, 1.0, 0.0), 0.5, glass)
    b.quad((-3.0, 0.0, -3.0), (6.0, 0.0, 0.0), (0.0, 0.0, 6.0), floor)
    b.sphere_light((0.0, 3.0, 0.0), 0.1, intensity=30.0, shadow_samples=1)
    b.sphere_light((1.0, 3.0, 0.0), 0.1, intensity=10.0, shadow_samples=1)
    desc = b.build(scenes.make_camera((0.0, 1.0, 5.0), (0.0, 1.0, 0.0), 40.0, 1.0))
    sc = oracle.OracleScene(desc)
    m = sc.photon_map(10001, 0.05)
    # luminous powers 4*pi*30 : 4*pi*10 -> round(7500.75) + round(2500.25)
    assert m.emitted == 7501 + 2500
    assert m.n > 0
    # glass refracts: photons land on the floor, arriving from above
    assert np.all(np.abs(m.pos[:, 1]) < 1e-3)
    assert np.all(m.inc[:, 1] > 0.0)
    assert np.all(m.power >= 0.0)
