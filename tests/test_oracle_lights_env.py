"""Oracle pins for light sampling (inc/light/*.glsl), light intersection
(lights.rint.glsl) and the environment map (inc/env_map.glsl,
env_map_eval.rcall.glsl) with its alias-table sampler: closed-form solid
angles, unbiasedness identities, and the alias table's exact reconstruction."""
import math

import numpy as np
import pytest
from scipy import stats

import hpt
import hpt_scenes
import pyoracle


def _light(kind, **kw):
    l = np.zeros(1, hpt.LIGHT_DTYPE)
    l["type"] = kind
    for k, v in kw.items():
        l[k] = v
    return l


def _rngs(n, seed):
    return np.random.default_rng(seed).integers(0, 2 ** 32, (n, 4), dtype=np.uint64).astype(np.uint32)


def _tri_solid_angle(a, b, c):
    # Van Oosterom & Strackee, IEEE Trans. Biomed. Eng. 30(2), 1983.
    la, lb, lc = (np.linalg.norm(x) for x in (a, b, c))
    num = abs(np.dot(a, np.cross(b, c)))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2 * math.atan2(num, den)


def test_point_and_spot_closed_forms():
    p = np.array([0.3, -0.2, 0.1])
    lp = np.array([1.0, 2.0, 3.0])
    l = _light(hpt.LIGHT_POINT, position=lp, intensity=(2, 3, 4))
    d, n, e, dist, pdf, _ = pyoracle.light_sample(l, 3, p, [1, 2, 3, 4])
    r = lp - p
    assert np.allclose(d, r / np.linalg.norm(r), atol=1e-6) and np.allclose(n, -d)
    assert math.isclose(dist, np.linalg.norm(r), rel_tol=1e-6) and pdf == 1.0
    assert np.allclose(e, np.array([2, 3, 4]) * 3 / np.dot(r, r), rtol=1e-5)  # times num_of_lights
    # spot: smoothstep(cos_outer, cos_inner, cos) between 30 and 20 degrees
    ci, co = math.cos(math.radians(20)), math.cos(math.radians(30))
    axis = np.array([0.0, 0.0, -1.0])
    for ang, expect in ((0.0, 1.0), (25.0, 0.5), (40.0, 0.0), (20.0, 1.0), (30.0, 0.0)):
        a = math.radians(ang)
        q = np.array([2 * math.sin(a), 0.0, 2 - 2 * math.cos(a)])  # distance 2 from the apex at (0,0,2)
        l = _light(hpt.LIGHT_SPOT, position=(0, 0, 2), u=axis, v=(ci, co, 0), intensity=(1, 1, 1))
        d, n, e, dist, pdf, _ = pyoracle.light_sample(l, 1, q, [0, 0, 0, 0])
        t = min(max((math.cos(a) - co) / (ci - co), 0), 1)
        assert math.isclose(float(e[0]) * 4, t * t * (3 - 2 * t), abs_tol=2e-5)
        assert math.isclose(float(e[0]) * 4, expect, abs_tol=0.1)


def test_directional_cone_is_uniform_in_solid_angle():
    cos_a = math.cos(math.radians(10))
    u = np.array([0.0, 0.6, -0.8])
    l = _light(hpt.LIGHT_DIRECTIONAL, u=u, v=(cos_a, 0, 0), intensity=(1, 1, 1))
    zs = []
    for r in _rngs(4000, 1):
        d, n, e, dist, pdf, _ = pyoracle.light_sample(l, 2, [0, 0, 0], r)
        assert pdf == 1.0 and dist == hpt.WORLD_SIZE and np.allclose(e, 2.0)
        zs.append(float(np.dot(d, -u)))
    zs = np.array(zs)
    assert zs.min() >= cos_a - 1e-6
    assert stats.kstest((zs - cos_a) / (1 - cos_a), "uniform").pvalue > 1e-3
    # v.x = 0: a delta light straight along -u
    l = _light(hpt.LIGHT_DIRECTIONAL, u=u, v=(0, 0, 0), intensity=(1, 1, 1))
    d, *_ = pyoracle.light_sample(l, 1, [0, 0, 0], [1, 2, 3, 4])
    assert np.allclose(d, -u)


@pytest.mark.parametrize("p", [(0.0, 0.0, 0.0), (0.7, -0.4, 0.3), (-1.5, 0.2, 1.2)])
def test_quad_light_pdf_keeps_reference_cosine(p):
    # quad_sample (inc/light/quad.glsl:20-30) takes cos_theta from the light
    # direction before it is normalised, so its pdf is dist/(area*cos), not
    # dist^2/(area*cos): E[1/pdf] = integral of cos/d dA over the quad, which
    # is `dist` times the solid angle.  Kept, and pinned here against both.
    pos, u, v = np.array([-0.25, -0.25, 2.0]), np.array([0.5, 0, 0]), np.array([0, 0.5, 0])
    l = _light(hpt.LIGHT_QUAD, position=pos, u=u, v=v, area=0.25, intensity=(1, 1, 1))
    p = np.array(p)
    inv = []
    for r in _rngs(20000, 2):
        d, n, e, dist, pdf, _ = pyoracle.light_sample(l, 1, p, r)
        assert np.allclose(n, [0, 0, -1])
        inv.append(1.0 / pdf)
    inv = np.array(inv)
    g = (np.arange(200) + 0.5) / 200
    S, T = np.meshgrid(g, g, indexing="ij")
    q = pos + S[..., None] * u + T[..., None] * v - p
    dd = np.linalg.norm(q, axis=-1)
    cos = q[..., 2] / dd
    integral = (cos / dd).mean() * 0.25
    c = [pos - p, pos + u - p, pos + u + v - p, pos + v - p]
    omega = _tri_solid_angle(c[0], c[1], c[2]) + _tri_solid_angle(c[0], c[2], c[3])
    assert abs(inv.mean() - integral) < 4 * inv.std() / math.sqrt(len(inv)) + 1e-4 * integral
    assert math.isclose((cos / dd ** 2).mean() * 0.25, omega, rel_tol=1e-3)  # quadrature vs closed form
    assert inv.mean() > 1.5 * omega


def test_quad_light_back_face_and_intersection_pdf():
    pos, u, v = np.array([-0.25, -0.25, 2.0]), np.array([0.5, 0, 0]), np.array([0, 0.5, 0])
    l = _light(hpt.LIGHT_QUAD, position=pos, u=u, v=v, area=0.25, intensity=(1, 1, 1))
    d, n, e, dist, pdf, _ = pyoracle.light_sample(l, 1, [0, 0, 3], [5, 6, 7, 8])
    assert pdf == 0.0 and np.all(e == 0)  # seen from behind
    t, _ = pyoracle.light_intersect(l, [0, 0, 3], [0, 0, -1])
    assert t == 1e6  # back faces are not hit
    p = np.array([0.3, 0.1, 0.2])
    for r in _rngs(50, 3):
        d, n, e, dist, pdf, _ = pyoracle.light_sample(l, 1, p, r)
        t, ipdf = pyoracle.light_intersect(l, p, d)
        # quad_intersect (quad.glsl:42-60) uses the unit ray direction: its pdf
        # is t^2/(area*cos), the sampler's times the distance.
        assert math.isclose(t, dist, rel_tol=1e-4) and math.isclose(ipdf, pdf * dist, rel_tol=1e-3)
        assert math.isclose(ipdf, t * t / (0.25 * abs(d[2])), rel_tol=1e-4)


def test_sphere_light_pdf_matches_hemisphere_area_measure():
    c, R = np.array([0.0, 0.0, 2.0]), 0.3
    l = _light(hpt.LIGHT_SPHERE, position=c, radius=R, area=4 * math.pi * R * R, intensity=(1, 1, 1))
    p = np.array([0.4, -0.3, 0.0])
    inv = np.array([1.0 / pyoracle.light_sample(l, 1, p, r)[4] for r in _rngs(20000, 4)])
    # The sampler draws the hemisphere facing p uniformly by area:
    # E[1/pdf] = integral over that hemisphere of |cos| / d^2 dA.
    w = (p - c) / np.linalg.norm(p - c)
    t = np.cross(w, [1.0, 0, 0])
    t /= np.linalg.norm(t)
    b = np.cross(w, t)
    th = (np.arange(400) + 0.5) / 400 * (math.pi / 2)
    ph = (np.arange(800) + 0.5) / 800 * 2 * math.pi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    nrm = (np.sin(TH) * np.cos(PH))[..., None] * t + (np.sin(TH) * np.sin(PH))[..., None] * b + np.cos(TH)[..., None] * w
    s = c + R * nrm
    dv = s - p
    d2 = (dv ** 2).sum(-1)
    cos = np.abs((nrm * dv).sum(-1)) / np.sqrt(d2)
    integral = (cos / d2 * R * R * np.sin(TH)).sum() * (math.pi / 2 / 400) * (2 * math.pi / 800)
    assert abs(inv.mean() - integral) < 4 * inv.std() / math.sqrt(len(inv)) + 1e-3 * integral
    # inside the sphere: no light
    assert pyoracle.light_sample(l, 1, c + [0, 0, 0.1], [1, 2, 3, 4])[4] == 0.0


def _env_scene(w=32, h=16, seed=0, rotation=0.0):
    sc = hpt_scenes.cornell_box()
    rng = np.random.default_rng(seed)
    env = rng.lognormal(0, 1.0, (h, w, 3)).astype(np.float32)
    env[3, 7] = 500.0
    sc.env_rgb = env
    sc.env_rotation = rotation
    return sc, env


def test_alias_table_reconstructs_luminance_exactly():
    lum = np.random.default_rng(5).lognormal(0, 2.0, 4096).astype(np.float32)
    lum[10] = 0.0
    prob, alias, total = pyoracle.build_alias(lum)
    n = len(lum)
    recon = prob.astype(np.float64).copy()
    np.add.at(recon, alias, 1.0 - prob.astype(np.float64))
    assert np.allclose(recon / n, lum.astype(np.float64) / total, rtol=1e-5, atol=1e-12)
    assert math.isclose(total, lum.astype(np.float64).sum(), rel_tol=1e-12)


def test_env_sampler_histogram_and_unbiasedness():
    sc, env = _env_scene()
    o = pyoracle.OracleScene(sc)
    h, w = env.shape[:2]
    counts = np.zeros((h, w))
    inv = []
    for r in _rngs(60000, 6):
        d, col, pdf, _ = o.env_sample(r)
        theta = math.acos(max(-1.0, min(1.0, float(d[2]))))
        u = (-(math.pi + math.atan2(d[1], d[0])) / (2 * math.pi)) % 1.0
        counts[min(int(theta / math.pi * h), h - 1), min(int(u * w), w - 1)] += 1
        if pdf > 0:
            inv.append(1.0 / pdf)
        else:
            inv.append(0.0)
    lum = 0.212671 * env[..., 0] + 0.715160 * env[..., 1] + 0.072169 * env[..., 2]
    expect = lum / lum.sum() * 60000
    ok = expect > 10
    chi2 = ((counts - expect) ** 2 / expect)[ok].sum()
    assert stats.chi2.sf(chi2, ok.sum() - 1) > 1e-4
    # pdf is the exact solid-angle density: E[1/pdf] = 4 pi
    inv = np.array(inv)
    assert abs(inv.mean() - 4 * math.pi) < 4 * inv.std() / math.sqrt(len(inv))


def test_env_eval_and_pdf_agree_with_sampler():
    sc, env = _env_scene(w=128, h=64, rotation=0.3)
    o = pyoracle.OracleScene(sc)
    h, w = env.shape[:2]
    total = float((0.212671 * env[..., 0] + 0.715160 * env[..., 1] + 0.072169 * env[..., 2]).astype(np.float64).sum())
    for r in _rngs(300, 7):
        d, col, pdf, _ = o.env_sample(r)
        ev = o.env_eval(d, depth=0)
        assert np.allclose(ev, col, rtol=2e-3, atol=1e-4)
        # env_pdf keeps the reference's threshold: zero unless lum/total > EPS
        ep = o.env_pdf(d)
        if pdf > 0 and ep > 0:
            assert math.isclose(ep, pdf, rel_tol=2e-3)
    # the bright texel is above the threshold, a dim one below it
    prob, alias, lum, tot = o.env_alias(w * h)
    assert math.isclose(tot, total, rel_tol=1e-6)
    assert lum.max() / tot > 3e-4 and np.median(lum) / tot < 3e-4


def test_env_mis_weight_at_depth():
    sc, env = _env_scene()
    o = pyoracle.OracleScene(sc)
    d, col, pdf, _ = o.env_sample(np.array([9, 9, 9, 9], np.uint32))
    e0 = o.env_eval(d, depth=0, bsdf_pdf=pdf)
    e1 = o.env_eval(d, depth=1, bsdf_pdf=pdf)
    ep = o.env_pdf(d)
    wgt = pdf * pdf / (pdf * pdf + ep * ep) if ep > 0 else 1.0
    assert np.allclose(e1, e0 * wgt, rtol=1e-4)


def test_env_eval_is_finite_past_the_pole():
    # normalize() can return z = 1 + 2^-23, where acos is undefined in GLSL
    # (a path of the 1080p grid scene met it at frame 66581); the argument is
    # clamped, so the look-up equals the one at z = 1 instead of a NaN that
    # would poison the pixel's running mean.
    sc, env = _env_scene()
    o = pyoracle.OracleScene(sc)
    z = float(np.nextafter(np.float32(1.0), np.float32(2.0)))
    for depth in (0, 1):
        e = o.env_eval([1.6e-4, 1.6e-4, z], depth=depth, bsdf_pdf=6.5)
        assert np.isfinite(e).all()
        assert np.array_equal(e, o.env_eval([1.6e-4, 1.6e-4, 1.0], depth=depth, bsdf_pdf=6.5))
