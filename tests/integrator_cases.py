"""End-to-end integrator cases with answers fixed by the mathematics.

The same cases run through the oracle (CPU) and through the HIP path: each
takes a `render(desc, settings) -> beauty (3, H, W)` and a
`primary(desc, settings) -> (prim, t, u, v)` callable.

Direct lighting through the whole render (Renderer.TraceRay / ShadeSurface /
ComputeDirectLighting / WeightEmission / MisWeight, Renderer.cs:1008-1228,
1534-1565, 1894-2036): a Lambertian floor quad of albedo rho under
emitters, max depth 2, black sky, clamps and roulette off.  The camera
vertex on the floor adds NEE and its bounce adds the radiance of the
emitter it hits; the depth-1 vertex lies on an emitter or the sky, so
nothing else contributes.  The expected sample value at a floor point x is

  rho / pi * sum_e  integral over the directions w reaching emitter e of
      L_e cos(n, w) [ w_nee(pPick_e p_e(w), p_bsdf(w)) V_e(x, w)
                      + w_em(p_bsdf(w), p_e(w)) ] dw

with p_bsdf = cos / pi (Lambertian.Pdf), p_e the emitter's solid-angle pdf
(sphere cone 1 / Omega, SphereLight.PdfSolidAngle; quad d^2 / (A cos_l),
AreaLight / GeometryLight.PdfSolidAngle), MisWeight(p, q) = p / (p + q)
(balance) or p^2 / (p^2 + q^2) (power).  Two reference rules make this differ
from the textbook rho / pi * E(x), and both are part of the pin:
  * WeightEmission weights a BSDF-sampled emitter hit with the light pdf
    without the pick probability (Renderer.cs:1549-1565) while NEE uses
    pPick * p_e (Renderer.cs:1956-1959): with power / uniform picking the two
    weights sum to less than one;
  * the shadow ray leaves from OffsetOrigin(x, n) = x + 1e-4 n towards the
    light sample with tMax = |sample - origin| - 1e-4 (SphereLight.cs:205-213,
    AreaLight.cs:150-156, GeometryLight.cs:169-177): a sample the shifted
    ray meets on the emitter's own proxy before tMax is shadowed (V = 0).
    For a planar emitter parallel to the floor that is every direction with
    cos < (sqrt(5) - 1) / 2; for a sphere a band of its rim.
With pPick = 1 and V = 1 the bracket is 1 and the integral is the closed
form E(x): sphere of radius r, centre c entirely above the floor,
E = pi L (r / d)^2 cos(theta); planar emitter E = L * (1/2) |sum_i gamma_i
n . unit(R_i x R_(i+1))| (Lambert's polygon formula).  The expected value is
computed as rho / pi * (E(x) - C(x)) with the correction C =
integral L cos [(1 - w_nee - w_em) + w_nee (1 - V)] dw by quadrature in
float64 (uniform in solid angle over the sphere's cone, uniform in area over
a quad).  The floor is a quad, so a primary hit's (u, v) = (alpha, beta)
gives x = Q + alpha U + beta V exactly; E is averaged over each pixel's own
samples and C taken at their mean point.  A dropped term (NEE or the
emission complement), a weight with or without the pick probability on the
wrong side, a missing 1 / pPick, a wrong stratum energy or a shadow origin
without the offset biases the image sum by a percent or more, far above its
standard error.
"""
import numpy as np

from raytracer import abi, scenes

RHO = 0.6
EPS = 1e-4
FLOOR_Q = np.array([-1.75, 0.0, -1.75])
FLOOR_U = np.array([0.0, 0.0, 3.5])   # U x V points up
FLOOR_V = np.array([3.5, 0.0, 0.0])

# emitters at one height, well apart: none occludes another's samples or
# BSDF directions for any floor point.  Every direction from the floor quad
# to a planar emitter has cos >= 0.70 (min_planar_cos), above the
# (sqrt(5) - 1) / 2 below which their own proxies shadow it, so the planar
# emitters are free of that rule (whose edge the float32 rounding of the
# shadow ray's t and tMax would move); the sphere keeps its rim band.
SPHERE = dict(c=np.array([0.6, 3.5, -0.5]), r=0.25, L=50.0)
QUAD = dict(q=np.array([-0.8, 3.5, -0.7]), u=np.array([0.5, 0.0, 0.0]), v=np.array([0.0, 0.0, 0.4]), L=25.0)
AREA = dict(q=np.array([-0.3, 3.5, 0.4]), u=np.array([0.6, 0.0, 0.0]), v=np.array([0.0, 0.0, 0.3]), L=20.0)


def min_planar_cos():
    """Smallest cosine to the floor normal of a segment from a floor corner
    to a planar-emitter corner (the extreme pairs of both convex sets)."""
    fl = [FLOOR_Q + a * FLOOR_U + b * FLOOR_V for a in (0, 1) for b in (0, 1)]
    c = 1.0
    for em in (QUAD, AREA):
        for a in (0, 1):
            for b in (0, 1):
                p = em["q"] + a * em["u"] + b * em["v"]
                for f in fl:
                    d = p - f
                    c = min(c, d[1] / np.linalg.norm(d))
    return c


MIS = {"balance": abi.RT_MIS_BALANCE, "power": abi.RT_MIS_POWER}
PICK = {"all": abi.RT_LIGHTS_ALL, "power": abi.RT_LIGHTS_POWER, "uniform": abi.RT_LIGHTS_UNIFORM}


def lambert_polygon(P, n, verts):
    """Integral of cos(n, w) over the solid angle of a convex polygon seen
    from each point of P (k, 3)."""
    R = [(v - P) / np.linalg.norm(v - P, axis=1, keepdims=True) for v in verts]
    s = np.zeros(len(P))
    for i in range(len(R)):
        a, b = R[i], R[(i + 1) % len(R)]
        c = np.cross(a, b)
        gamma = np.arccos(np.clip((a * b).sum(1), -1.0, 1.0))
        s += gamma * (c @ n) / np.linalg.norm(c, axis=1)
    return np.abs(0.5 * s)


def floor_scene(emitters, W, H):
    """Floor quad (primitive 0) + the chosen emitters: 'sphere' (SphereLight
    with its visible proxy), 'area' (AreaLight with its back-face-culled
    proxy quad), 'quad' (emissive quad: a GeometryLight, last in the light
    list).  Planar emitters face down (U x V points down)."""
    b = scenes.SceneBuilder()
    floor = b.lambertian((RHO, RHO, RHO))
    b.quad(tuple(FLOOR_Q), tuple(FLOOR_U), tuple(FLOOR_V), floor)
    if "sphere" in emitters:
        b.sphere_light(tuple(SPHERE["c"]), SPHERE["r"], (1.0, 1.0, 1.0), SPHERE["L"], shadow_samples=4)
    if "area" in emitters:
        b.area_light(tuple(AREA["q"]), tuple(AREA["u"]), tuple(AREA["v"]), (1.0, 1.0, 1.0), AREA["L"],
                     shadow_samples=4)
    if "quad" in emitters:
        em = b.emissive((1.0, 1.0, 1.0), QUAD["L"])
        b.quad(tuple(QUAD["q"]), tuple(QUAD["u"]), tuple(QUAD["v"]), em)
    cam = scenes.make_camera((0.0, 7.5, 3.5), (0.0, 0.0, 0.0), 36.0, W / H)
    return b.build(cam)


def light_order(emitters):
    """Light-list order of floor_scene: explicit lights as added, geometry
    lights last (SceneLoader: CreateLight, then ExtractGeometryLights)."""
    return [e for e in ("sphere", "area", "quad") if e in emitters]


def pick_probs(emitters, pick):
    """LightDistribution pdf (LightDistribution.cs:41-98): power shares
    (SphereLight 4 pi I, AreaLight / GeometryLight pi A L, luminance of a
    white colour = 1) or uniform."""
    order = light_order(emitters)
    if pick == "all":
        return {e: 1.0 for e in order}
    if pick == "uniform":
        return {e: 1.0 / len(order) for e in order}
    pw = {"sphere": 4 * np.pi * SPHERE["L"],
          "area": np.pi * np.linalg.norm(np.cross(AREA["u"], AREA["v"])) * AREA["L"],
          "quad": np.pi * np.linalg.norm(np.cross(QUAD["u"], QUAD["v"])) * QUAD["L"]}
    tot = sum(pw[e] for e in order)
    return {e: pw[e] / tot for e in order}


def mis_weight(p, q, mis):
    if mis == "power":
        return p * p / (p * p + q * q + 1e-30)
    return p / (p + q + 1e-30)


def closed_form_E(P, emitters):
    """Irradiance of the emitters at floor points P (k, 3), float64."""
    n = np.array([0.0, 1.0, 0.0])
    E = np.zeros(len(P))
    if "sphere" in emitters:
        d = SPHERE["c"] - P
        dist = np.linalg.norm(d, axis=1)
        E += np.pi * SPHERE["L"] * (SPHERE["r"] / dist) ** 2 * (d[:, 1] / dist)
    for name, em in (("quad", QUAD), ("area", AREA)):
        if name in emitters:
            q, u, v = em["q"], em["u"], em["v"]
            E += em["L"] * lambert_polygon(P, n, [q, q + u, q + u + v, q + v])
    return E


# ── shifted shadow rays against the emitter shapes (float64) ───────────────
def _hit_sphere(o, w, tmax, c, r):
    oc = o - c
    b = (oc * w).sum(-1)
    cc = (oc * oc).sum(-1) - r * r
    disc = b * b - cc
    sq = np.sqrt(np.maximum(disc, 0.0))
    t1, t2 = -b - sq, -b + sq
    ok = disc >= 0
    return ok & (((t1 >= EPS) & (t1 <= tmax)) | ((t2 >= EPS) & (t2 <= tmax)))


def _hit_quad(o, w, tmax, q, u, v, culled):
    nrm = np.cross(u, v)
    nn = nrm / np.linalg.norm(nrm)
    denom = w @ nn
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (q @ nn - o @ nn) / denom
    p = o + t[..., None] * w
    wv = nrm / (nrm @ nrm)
    rel = p - q
    alpha = np.cross(rel, v) @ wv
    beta = np.cross(u, rel) @ wv
    hit = (np.abs(denom) >= 1e-8) & (t >= EPS) & (t <= tmax) & (alpha >= 0) & (alpha <= 1) & (beta >= 0) & (beta <= 1)
    if culled:
        hit &= denom < 0   # BackFaceCulledHittable: front-face hits only
    return hit


def _blocked(o, w, tmax, emitters):
    b = np.zeros(w.shape[:-1], bool)
    if "sphere" in emitters:
        b |= _hit_sphere(o, w, tmax, SPHERE["c"], SPHERE["r"])
    if "area" in emitters:
        b |= _hit_quad(o, w, tmax, AREA["q"], AREA["u"], AREA["v"], True)
    if "quad" in emitters:
        b |= _hit_quad(o, w, tmax, QUAD["q"], QUAD["u"], QUAD["v"], False)
    return b


def correction(P, emitters, mis, pick, m=64):
    """C(x) = sum_e integral L cos [(1 - w_nee - w_em) + w_nee (1 - V)] dw
    at floor points P (k, 3), by quadrature (m x m per emitter).  Returns
    (C, the sphere's self-shadow part of C).  That part is known only to a
    few per cent of itself: where the shifted shadow ray grazes the sphere
    the reference decides with the float32 quadratic of Sphere.Hit, whose
    rounding moves the edge of the shadowed rim band (float64 here)."""
    n = np.array([0.0, 1.0, 0.0])
    pp = pick_probs(emitters, pick)
    g = (np.arange(m) + 0.5) / m
    ga, gb = [x.ravel() for x in np.meshgrid(g, g, indexing="ij")]
    C = np.zeros(len(P))
    Cs = np.zeros(len(P))
    for k0 in range(0, len(P), 64):
        X = P[k0:k0 + 64][:, None, :]           # (k, 1, 3)
        so = X + EPS * n                          # OffsetOrigin(x, n)
        for e in light_order(emitters):
            if e == "sphere":
                c, r, L = SPHERE["c"], SPHERE["r"], SPHERE["L"]
                tc = c - X[:, 0]
                dist = np.linalg.norm(tc, axis=1)
                wz = tc / dist[:, None]
                ctmax = np.sqrt(1.0 - (r / dist) ** 2)
                omega = 2 * np.pi * (1 - ctmax)
                ct = 1.0 - ga[None] * (1 - ctmax[:, None])
                st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
                ph = 2 * np.pi * gb[None]
                ax = np.cross(wz, [1.0, 0.0, 0.0])
                ax /= np.linalg.norm(ax, axis=1, keepdims=True)
                ay = np.cross(wz, ax)
                w = (ct[..., None] * wz[:, None] + st[..., None] * (np.cos(ph)[..., None] * ax[:, None] +
                                                                    np.sin(ph)[..., None] * ay[:, None]))
                oc = X - c
                bq = (oc * w).sum(-1)
                cq = (oc * oc).sum(-1) - r * r
                tn = -bq - np.sqrt(np.maximum(bq * bq - cq, 0.0))
                sp = X + tn[..., None] * w
                p_e = np.broadcast_to((1.0 / omega)[:, None], tn.shape)
                meas = omega[:, None] / (m * m)          # solid angle per quadrature point
                cosf = w @ n
            else:
                em = QUAD if e == "quad" else AREA
                q, u, v, L = em["q"], em["u"], em["v"], em["L"]
                A = np.linalg.norm(np.cross(u, v))
                nl = -np.cross(u, v) / A                  # emitting side faces the floor
                sp = (q + ga[:, None] * u + gb[:, None] * v)[None]   # (1, M, 3)
                dv = sp - X
                d2 = (dv * dv).sum(-1)
                w = dv / np.sqrt(d2)[..., None]
                cosl = w @ nl
                p_e = d2 / (A * cosl)
                meas = A * cosl / d2 / (m * m)
                cosf = w @ n
                sp = np.broadcast_to(sp, w.shape)
            pb = cosf / np.pi
            wn = mis_weight(pp[e] * p_e, pb, mis)
            wb = mis_weight(pb, p_e, mis)
            tmax = np.linalg.norm(sp - so, axis=-1) - EPS
            V = ~_blocked(so, w, tmax, emitters)
            integrand = L * cosf * ((1.0 - wn - wb) + wn * (~V))
            C[k0:k0 + 64] += (integrand * meas).sum(1)
            if e == "sphere":
                Cs[k0:k0 + 64] += (L * cosf * wn * (~V) * meas).sum(1)
    return C, Cs


def unclamped(W, H, spp, depth, mis="balance", pick="all", seed_xor=0, roulette=False):
    st = abi.settings_defaults(W, H, spp, depth)
    st.mis = MIS[mis]
    st.light_sampling = PICK[pick]
    st.max_sample_radiance = 1e35
    st.indirect_clamp_factor = 1.0
    st.seed_xor = seed_xor
    if not roulette:
        st.rr_min_bounces, st.rr_min_survival = 1000, 0.1
    return st


def direct_lighting_case(render, primary, emitters, mis, pick, W=32, H=32, spp=32, runs=6, model=True):
    """Per independent run k: D_k = sum over floor pixels of (rendered -
    expected).  Returns (mean D, standard error of the mean, sum of
    expected, per-pixel (rendered, expected) stacks).  model=False uses the
    textbook rho / pi * E(x) (no correction)."""
    desc = floor_scene(emitters, W, H)
    D, tot, R, E, U = [], [], [], [], []
    for k in range(runs):
        st = unclamped(W, H, spp, 2, mis, pick, seed_xor=(0x9E3779B9 * (k + 1)) & 0xFFFFFFFF)
        prim, _, u, v = primary(desc, st)
        prim = prim.reshape(spp, H, W)
        uu, vv = u.reshape(spp, H, W).astype(np.float64), v.reshape(spp, H, W).astype(np.float64)
        on_floor = np.all(prim == 0, axis=0)
        pts = FLOOR_Q + uu[..., None] * FLOOR_U + vv[..., None] * FLOOR_V    # (spp, H, W, 3)
        exp = closed_form_E(pts.reshape(-1, 3), emitters).reshape(spp, H, W).mean(0)
        unc = np.zeros((H, W))
        if model:
            centre = pts.mean(0)[on_floor]
            corr = np.zeros((H, W))
            corr[on_floor], unc[on_floor] = correction(centre, emitters, mis, pick)
            exp = exp - corr
        exp = RHO / np.pi * exp
        unc = RHO / np.pi * unc
        beauty = render(desc, st).astype(np.float64)
        got = beauty.mean(0)   # the floor is grey: every channel carries the same estimate
        D.append(float((got - exp)[on_floor].sum()))
        tot.append(float(exp[on_floor].sum()))
        U.append(np.where(on_floor, unc, 0.0))
        # the block comparison takes the pixels where NEE reaches at least half
        # of the light (elsewhere the image is the rare BSDF hits of a shadowed
        # emitter)
        lit = on_floor & (exp >= 0.5 * RHO / np.pi * closed_form_E(pts.reshape(-1, 3), emitters).reshape(spp, H, W).mean(0))
        R.append(np.where(lit, got, np.nan))
        E.append(np.where(lit, exp, np.nan))
    D = np.array(D)
    out = D.mean(), D.std(ddof=1) / np.sqrt(runs), float(np.mean(tot)), np.array(R), np.array(E)
    return out + (np.mean(U, 0),) if model else out


def check_direct_lighting(render, primary, emitters, mis, pick, **kw):
    mean_d, se, total, R, E, shadow = direct_lighting_case(render, primary, emitters, mis, pick, **kw)
    rel = mean_d / total
    # tolerance: 3 standard errors, 2e-4 of the image for the quadrature,
    # and 10 % of the sphere's self-shadowed share (float32 rim edge)
    tol = 3.0 * se + 2e-4 * total + 0.1 * float(shadow.sum())
    print(f"{emitters} mis={mis} pick={pick}: image sum {total:.4f}, bias {mean_d:+.5f} "
          f"({rel:+.2e}), tolerance {tol:.5f} (3 se {3 * se:.5f}, sphere self-shadow {shadow.sum():.4f})", flush=True)
    assert total > 0
    assert abs(mean_d) <= tol, (mean_d, se, total, shadow)
    # by region: 8 x 8-pixel blocks of the lit floor, the mean over the
    # independent runs of each block's sum within 4 sigma (run-to-run spread)
    # of its expected sum, after the block's rim-edge allowance.  Blocks, not
    # pixels: a pixel's few BSDF-sampled emitter hits are too rare for a
    # per-pixel spread estimate.
    H, W = R.shape[1:]
    lit = np.all(np.isfinite(R), 0)
    nblk = 0
    for by in range(0, H, 8):
        for bx in range(0, W, 8):
            m = lit[by:by + 8, bx:bx + 8]
            if m.sum() < 16:
                continue
            nblk += 1
            rs = np.array([r[by:by + 8, bx:bx + 8][m].sum() for r in R])
            es = np.array([e[by:by + 8, bx:bx + 8][m].sum() for e in E]).mean()
            allow = 0.1 * shadow[by:by + 8, bx:bx + 8][m].sum()
            sd = rs.std(ddof=1) / np.sqrt(len(rs))
            dev = max(abs(rs.mean() - es) - allow, 0.0)
            assert dev <= 4.0 * sd + 2e-4 * es, (by, bx, rs.mean(), es, sd, allow)
    assert nblk >= 4, nblk


def cornell_multi(W, H):
    """The plain-method Cornell box plus two sphere lights of different power
    (three lights, so power and uniform picking differ)."""
    b = scenes.SceneBuilder()
    white = b.lambertian((0.73, 0.73, 0.73))
    red = b.lambertian((0.65, 0.05, 0.05))
    green = b.lambertian((0.12, 0.45, 0.15))
    P = lambda x, y, z: np.array([x, y, z], np.float32)
    walls = [
        (P(-1, 0, -1), P(1, 0, -1), P(1, 0, 1), white), (P(-1, 0, -1), P(1, 0, 1), P(-1, 0, 1), white),
        (P(-1, 2, -1), P(-1, 2, 1), P(1, 2, 1), white), (P(-1, 2, -1), P(1, 2, 1), P(1, 2, -1), white),
        (P(-1, 0, -1), P(-1, 2, -1), P(1, 2, -1), white), (P(-1, 0, -1), P(1, 2, -1), P(1, 0, -1), white),
        (P(-1, 0, -1), P(-1, 0, 1), P(-1, 2, 1), red), (P(-1, 0, -1), P(-1, 2, 1), P(-1, 2, -1), red),
        (P(1, 0, -1), P(1, 2, -1), P(1, 2, 1), green), (P(1, 0, -1), P(1, 2, 1), P(1, 0, 1), green),
    ]
    for a, c, d, m in walls:
        b.triangle(a, c, d, m)
    b.sphere((-0.45, 0.3, -0.3), 0.3, white)
    glass = b.disney((1.0, 1.0, 1.0), spec_trans=1.0, ior=1.5, roughness=0.0)
    b.sphere((0.45, 0.3, 0.2), 0.3, glass)
    em = b.emissive((1.0, 1.0, 1.0), 15.0)
    b.quad((-0.125, 1.999, -0.125), (0.25, 0.0, 0.0), (0.0, 0.0, 0.25), em)
    b.sphere_light((-0.6, 1.5, 0.4), 0.08, (1.0, 0.8, 0.6), 30.0, shadow_samples=1)
    b.sphere_light((0.55, 1.2, -0.6), 0.12, (0.6, 0.8, 1.0), 6.0, shadow_samples=2)
    cam = scenes.make_camera((0.0, 1.0, 3.9), (0.0, 1.0, 0.0), 39.0, W / H)
    return b.build(cam)


def estimator_means(render, W=24, H=24, spp=32, runs=6):
    """Image-mean luminance of the balance and power MIS heuristics with
    all-light sampling over independent runs on the multi-light Cornell box
    (depth 8, roulette on, clamps off): {mis: (mean, standard error)}.
    Both are unbiased estimators of the same image (the weights of every
    light sum to one with all-light sampling), so their means must agree."""
    desc = cornell_multi(W, H)
    out = {}
    for mis in ("balance", "power"):
        vals = []
        for k in range(runs):
            st = unclamped(W, H, spp, 8, mis, "all", seed_xor=(0x85EBCA6B * (k + 3)) & 0xFFFFFFFF, roulette=True)
            b = render(desc, st).astype(np.float64)
            vals.append(float((0.2126 * b[0] + 0.7152 * b[1] + 0.0722 * b[2]).mean()))
        vals = np.array(vals)
        out[mis] = (vals.mean(), vals.std(ddof=1) / np.sqrt(runs))
    return out


def check_estimator_invariance(means):
    (mb, sb), (mp, sp) = means["balance"], means["power"]
    print(f"Cornell image mean: balance {mb:.5f} +- {sb:.5f}, power {mp:.5f} +- {sp:.5f}", flush=True)
    assert abs(mb - mp) <= 3.5 * np.hypot(sb, sp), (mb, mp, sb, sp)


SKY = (0.4, 0.6, 0.9)


def sky_scene(W, H, sky=SKY, extra_light=False):
    """An infinite Lambertian ground (albedo RHO) under a flat sky, which the
    scene builder turns into an EnvironmentLight; optionally a sphere light."""
    b = scenes.SceneBuilder()
    b.plane((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), b.lambertian((RHO, RHO, RHO)))
    if extra_light:
        b.sphere_light((0.5, 1.5, -0.5), 0.2, (1.0, 1.0, 1.0), 20.0, shadow_samples=2)
    cam = scenes.make_camera((0.0, 1.5, 3.0), (0.0, 0.3, 0.0), 60.0, W / H)
    return b.build(cam, sky=sky)


def check_sky_furnace(render, primary, mis, pick, W=32, H=32, spp=16, runs=6, depth=3):
    """Uniform sky radiance L over an unoccluded ground: the irradiance is
    pi L, so every ground sample's expected value is rho L per channel (NEE
    towards the environment light with w_nee, escaping BSDF bounces with the
    complementary weight; a ground bounce can only escape); sky pixels are L
    exactly.  A missing MIS complement, a wrong sphere pdf or sample weight,
    or a horizon test on the wrong side moves the ground by percent."""
    desc = sky_scene(W, H)
    L = np.array(SKY)
    sums = []
    for k in range(runs):
        st = unclamped(W, H, spp, depth, mis, pick, seed_xor=(0x27D4EB2F * (k + 1)) & 0xFFFFFFFF)
        prim = primary(desc, st)[0].reshape(spp, H, W)
        ground = np.all(prim == 0, 0)
        sky_px = np.all(prim < 0, 0)
        img = render(desc, st).astype(np.float64)
        assert ground.sum() > 100 and sky_px.sum() > 50
        assert np.allclose(img[:, sky_px].T, L.astype(np.float32), rtol=1e-6), "sky pixels must be the sky"
        sums.append(img[:, ground].sum(1) / (RHO * L * ground.sum()))
    sums = np.array(sums)            # (runs, 3): ground mean / (rho L)
    m, se = sums.mean(0), sums.std(0, ddof=1) / np.sqrt(runs)
    print(f"sky furnace mis={mis} pick={pick}: ground / (rho L) = {m} +- {se}", flush=True)
    assert np.all(np.abs(m - 1.0) <= 3.0 * se + 2e-4), (m, se)
