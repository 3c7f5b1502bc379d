"""Closed-form checks of the oracle's light estimators, the medium's free-
flight sampler and the caustic density estimate.

Lights (SphereLight / AreaLight / DirectionalLight / GeometryLight): the
light sample's colour is the per-sample estimate of the incident radiance
integral, so the mean of colour * cos(n, dir) over a fine stratified grid
of random numbers must equal the irradiance of the emitter, and the mean
of 1 / pdf_solid_angle must equal the solid angle the sampler covers:
  * sphere of radius r at distance d, normal towards it: E = pi L (r/d)^2,
    solid angle 2 pi (1 - sqrt(1 - (r/d)^2));
  * quad / triangle emitter: E = L * (1/2) sum_i gamma_i n . unit(R_i x R_(i+1))
    (Lambert's polygon formula), solid angle by Van Oosterom-Strackee;
  * sun of angular radius a: directions uniform in the cone, so the mean
    cosine is (1 + cos a) / 2 and pdf = 1 / (2 pi (1 - cos a)).
Medium (HomogeneousMedium.Sample, spectral one-sample MIS): E[beta 1{no
scatter}] = exp(-sigma_t t_max) and E[beta 1{scatter before s}] =
sigma_s / sigma_t (1 - exp(-sigma_t s)) per channel.
Caustics (GatherCaustics): photons of equal power spread uniformly over a
plane give, through the cone filter (k_f = 1, normalisation 1/3), the
density rho = Phi / A times the BRDF, away from the edges.
"""
import numpy as np
import pytest

from raytracer import scenes

GRID = 160   # stratified grid per random dimension


def _grid():
    u = (np.arange(GRID) + 0.5) / GRID
    a, b = np.meshgrid(u, u, indexing="ij")
    return np.stack([a.ravel(), b.ravel()], 1).astype(np.float32)


def _estimate(osc, p, n):
    e = np.zeros(3)
    inv_pdf = 0.0
    for r in _grid():
        # light sample 0 of 1 (one shadow sample: the stratum is the unit square)
        shadow, col, d, _, pdf = osc.light_sample(0, p, n, 0, r)
        assert not shadow
        e += col * max(0.0, float(np.dot(n, d)))
        inv_pdf += 1.0 / pdf if pdf > 0 else 0.0
    return e / len(_grid()), inv_pdf / len(_grid())


def _lambert_polygon(p, n, verts):
    """Integral of cos(n, w) over the solid angle of a convex polygon."""
    R = [(v - p) / np.linalg.norm(v - p) for v in verts]
    s = 0.0
    for i in range(len(R)):
        a, b = R[i], R[(i + 1) % len(R)]
        c = np.cross(a, b)
        gamma = np.arccos(np.clip(np.dot(a, b), -1.0, 1.0))
        s += gamma * np.dot(n, c / np.linalg.norm(c))
    return abs(0.5 * s)


def _solid_angle_triangle(p, a, b, c):
    """Van Oosterom & Strackee (1983)."""
    r1, r2, r3 = a - p, b - p, c - p
    l1, l2, l3 = (np.linalg.norm(x) for x in (r1, r2, r3))
    num = abs(np.dot(r1, np.cross(r2, r3)))
    den = l1 * l2 * l3 + np.dot(r1, r2) * l3 + np.dot(r1, r3) * l2 + np.dot(r2, r3) * l1
    return 2.0 * np.arctan2(num, den)


def _cam():
    return scenes.make_camera((0, 0, 5), (0, 0, 0), 40.0, 1.0)


def _without_proxies(desc):
    """The same lights over a world of one far-away triangle: a light's own
    proxy would shadow samples at its silhouette from the offset shadow-ray
    origin (a reference behaviour that biases the rim, not the estimator)."""
    prims = np.zeros(1, scenes.PRIM_DTYPE)
    prims["kind"] = 0
    prims["mesh"] = -1
    prims["v"][0, :9] = [100, 100, 100, 101, 100, 100, 100, 101, 100]
    lights = desc.lights.copy()
    lights["primitive"] = -1
    return scenes.SceneDesc(prims, desc.mats, lights, desc.camera)


def test_sphere_light_irradiance_and_solid_angle(oracle):
    b = scenes.SceneBuilder()
    L, r, dist = 7.0, 0.3, 2.0
    b.sphere_light((0.0, dist, 0.0), r, (1.0, 1.0, 1.0), L, shadow_samples=1)
    osc = oracle.OracleScene(_without_proxies(b.build(_cam())))
    p, n = np.zeros(3, np.float32), np.array([0, 1, 0], np.float32)
    e, omega = _estimate(osc, p, n)
    s2 = (r / dist) ** 2
    assert np.allclose(e, np.pi * L * s2, rtol=2e-3)
    assert abs(omega - 2 * np.pi * (1 - np.sqrt(1 - s2))) <= 1e-4 * omega


def test_area_light_irradiance_and_solid_angle(oracle):
    b = scenes.SceneBuilder()
    L = 5.0
    corner, u, v = np.array([-0.4, 1.5, -0.3]), np.array([0.8, 0, 0]), np.array([0, 0, 0.6])
    b.area_light(tuple(corner), tuple(u), tuple(v), (1.0, 1.0, 1.0), L, shadow_samples=1)
    osc = oracle.OracleScene(_without_proxies(b.build(_cam())))
    p, n = np.array([0.1, 0.0, 0.05], np.float32), np.array([0, 1, 0], np.float32)
    e, omega = _estimate(osc, p, n)
    verts = [corner, corner + u, corner + u + v, corner + v]
    want = L * _lambert_polygon(p.astype(np.float64), n.astype(np.float64), verts)
    assert np.allclose(e, want, rtol=2e-3)
    sa = _solid_angle_triangle(p, verts[0], verts[1], verts[2]) + _solid_angle_triangle(p, verts[0], verts[2], verts[3])
    assert abs(omega - sa) <= 2e-3 * sa


def test_geometry_light_triangle_irradiance(oracle):
    b = scenes.SceneBuilder()
    b.shadow_samples_override = 1
    Le = 3.0
    m = b.emissive((1.0, 1.0, 1.0), Le)
    tri = np.array([[[-0.5, 1.2, -0.4], [0.6, 1.3, -0.2], [0.0, 1.1, 0.7]]], np.float32)
    b.triangles(tri, m)
    osc = oracle.OracleScene(b.build(_cam()))
    p, n = np.array([0.0, 0.0, 0.0], np.float32), np.array([0, 1, 0], np.float32)
    e, omega = _estimate(osc, p, n)
    verts = [tri[0, k].astype(np.float64) for k in range(3)]
    want = Le * _lambert_polygon(p.astype(np.float64), n.astype(np.float64), verts)
    assert np.allclose(e, want, rtol=2e-3)
    assert abs(omega - _solid_angle_triangle(p, *verts)) <= 2e-3 * omega


@pytest.mark.parametrize("angle", [0.0, 2.5])
def test_directional_light_cone(oracle, angle):
    b = scenes.SceneBuilder()
    b.shadow_samples_override = 1   # the estimate of one sample (the sun divides by its sample count)
    E0 = 3.0
    b.directional_light((0.0, -1.0, 0.0), (1.0, 1.0, 1.0), E0, angular_radius_deg=angle)
    osc = oracle.OracleScene(b.build(_cam()))
    p, n = np.zeros(3, np.float32), np.array([0, 1, 0], np.float32)
    e = np.zeros(3)
    for r in _grid():
        shadow, col, d, _, pdf = osc.light_sample(0, p, n, 0, r)
        assert not shadow
        e += col * float(np.dot(n, d))
        if angle > 0:
            ca = np.cos(np.radians(angle))
            assert abs(pdf - 1.0 / (2 * np.pi * (1 - ca))) <= 1e-3 * pdf
            assert float(np.dot(-np.array([0, -1, 0]), d)) >= ca - 1e-6
    e /= len(_grid())
    ca = np.cos(np.radians(angle))
    assert np.allclose(e, E0 * (1 + ca) / 2, rtol=1e-4)


@pytest.mark.parametrize("sig_a, sig_s", [((0.05, 0.1, 0.2), (0.3, 0.2, 0.05)), ((0.0, 0.5, 1.0), (0.45, 0.1, 0.0))])
def test_medium_free_flight_is_unbiased(oracle, sig_a, sig_s):
    sig_t = np.array(sig_a) + np.array(sig_s)
    tmax, s_cut = 2.0, 0.7
    n_d = 4000
    no_scat = np.zeros(3)
    scat_before = np.zeros(3)
    for ch in range(3):
        uc = (ch + 0.5) / 3.0
        for k in range(n_d):
            t, beta, sc = oracle.medium_sample(sig_t, sig_s, tmax, uc, (k + 0.5) / n_d)
            if not sc:
                no_scat += beta
            elif t < s_cut:
                scat_before += beta
    no_scat /= 3 * n_d
    scat_before /= 3 * n_d
    assert np.allclose(no_scat, np.exp(-sig_t * tmax), rtol=2e-3, atol=1e-6)
    with np.errstate(divide="ignore", invalid="ignore"):
        want = np.where(sig_t > 0, np.array(sig_s) / sig_t * (1 - np.exp(-sig_t * s_cut)), 0.0)
    assert np.allclose(scat_before, want, rtol=5e-3, atol=2e-4)


def test_caustic_estimate_of_uniform_photons(oracle):
    rng = np.random.default_rng(4)
    n, area = 200000, 4.0
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = rng.uniform(-1, 1, n)
    pos[:, 2] = rng.uniform(-1, 1, n)
    phi = np.array([2.0, 1.0, 0.5])
    m = oracle.PhotonMap.from_positions(pos, 0.01)
    m.set_all(phi / n, (0.0, 1.0, 0.0))
    q = np.zeros((400, 3), np.float32)
    q[:, 0] = rng.uniform(-0.6, 0.6, 400)
    q[:, 2] = rng.uniform(-0.6, 0.6, 400)
    nrm = np.tile([0.0, 1.0, 0.0], (400, 1))
    albedo = np.array([0.8, 0.5, 0.2])
    est = oracle.caustic_estimate(m, q, nrm, nrm, albedo, 0.2)
    want = albedo / np.pi * phi / area
    got = est.mean(0)
    print("caustic estimate", got, "want", want, flush=True)
    assert np.allclose(got, want, rtol=0.02)
