"""Oracle pins for the BxDFs (inc/bxdf/{lambert,oren_nayar,disney}.glsl and
common/sampling.glsl): textbook reductions, sampler-vs-pdf consistency by a
chi-square test over the sphere, the mirror limit, and the two reference
behaviours that are not self-consistent (kept verbatim and pinned as such)."""
import math

import numpy as np
import pytest
from scipy import stats

import hpt
import hpt_scenes
import pyoracle

N = np.array([0.0, 0.0, 1.0])
T = np.array([1.0, 0.0, 0.0])
B = np.array([0.0, 1.0, 0.0])


def _mat(kind="diffuse", **kw):
    m = hpt.new_materials(1)
    if kind == "diffuse":
        hpt_scenes.diffuse_material(m[0], kw.pop("color", (0.6, 0.5, 0.4)), **kw)
    else:
        hpt_scenes.disney_material(m[0], kw.pop("color", (0.6, 0.5, 0.4)), **kw)
    return m


def _dir(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def test_lambert_closed_form():
    m = _mat(roughness=0.0)
    fr = pyoracle.make_frame(N, T, B)
    rng = np.random.default_rng(0)
    for _ in range(50):
        V = _dir(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        L = _dir(rng.uniform(0, 3.1), rng.uniform(0, 6.28))
        f, pdf = pyoracle.bxdf_eval(m, fr, V, L)
        c = L[2]
        assert np.allclose(f, np.array([0.6, 0.5, 0.4]) * c / math.pi, rtol=1e-5, atol=1e-7)
        assert math.isclose(pdf, c / math.pi, rel_tol=1e-5, abs_tol=1e-7)  # no clamp (lambert.glsl:38-41)


def test_oren_nayar_reduces_to_scaled_lambert():
    m = _mat(roughness=0.5)
    A = float(m["ax"][0])
    m["ay"] = 0.0  # B = 0
    fr = pyoracle.make_frame(N, T, B)
    rng = np.random.default_rng(1)
    for _ in range(50):
        V = _dir(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        L = _dir(rng.uniform(0, 1.5), rng.uniform(0, 6.28))
        f, pdf = pyoracle.bxdf_eval(m, fr, V, L)
        assert np.allclose(f, A * np.array([0.6, 0.5, 0.4]) * L[2] / math.pi, rtol=1e-5)
    # sigma = roughness*pi/2: A = 1 - s^2/(2(s^2+0.33)), B = 0.45 s^2/(s^2+0.09)
    s2 = (0.5 * math.pi / 2) ** 2
    m2 = _mat(roughness=0.5)
    assert math.isclose(float(m2["ax"][0]), 1 - s2 / (2 * (s2 + 0.33)), rel_tol=1e-6)
    assert math.isclose(float(m2["ay"][0]), 0.45 * s2 / (s2 + 0.09), rel_tol=1e-6)


def test_oren_nayar_retro_term():
    # wo, wi in the same azimuth, theta_i = theta_o = 60 deg:
    # f = rho/pi (A + B sin^2 / cos) (oren_nayar.glsl:7-51).
    m = _mat(roughness=0.8)
    A, Bc = float(m["ax"][0]), float(m["ay"][0])
    fr = pyoracle.make_frame(N, T, B)
    th = math.radians(60)
    V = _dir(th, 0.3)
    L = _dir(th, 0.3)
    f, _ = pyoracle.bxdf_eval(m, fr, V, L)
    expect = np.array([0.6, 0.5, 0.4]) / math.pi * (A + Bc * math.sin(th) * math.tan(th)) * math.cos(th)
    assert np.allclose(f, expect, rtol=1e-5)


def _chi_square(mat, V, n=60000, bins=(12, 16), depth=0, seed=0, sub=6, z_ok=None, return_maps=False):
    """Histogram of sampled directions vs the eval pdf integrated per bin."""
    fr = pyoracle.make_frame(N, T, B, eta=1.0 / 1.5 if mat["type"][0] == 1 else 1.0, depth=depth)
    rng = np.random.default_rng(seed)
    nz, nphi = bins
    counts = np.zeros(bins)
    invalid = 0
    state = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
    for i in range(n):
        L, f, pdf, flags, state = pyoracle.bxdf_sample(mat, fr, V, state)
        if not pdf > 0:
            invalid += 1
            continue
        z = min(max(L[2], -1.0), 1.0)
        iz = min(int((z + 1) / 2 * nz), nz - 1)
        ip = min(int((math.atan2(L[1], L[0]) % (2 * math.pi)) / (2 * math.pi) * nphi), nphi - 1)
        counts[iz, ip] += 1
        if i < 200:  # the sampler returns the eval pdf of the direction it drew
            f2, p2 = pyoracle.bxdf_eval(mat, fr, V, L)
            assert math.isclose(p2, pdf, rel_tol=2e-3)
    expect = np.zeros(bins)
    for iz in range(nz):
        for ip in range(nphi):
            acc = 0.0
            for a in range(sub):
                for b in range(sub):
                    z = -1 + 2 * (iz + (a + 0.5) / sub) / nz
                    ph = 2 * math.pi * (ip + (b + 0.5) / sub) / nphi
                    s = math.sqrt(max(0.0, 1 - z * z))
                    L = np.array([s * math.cos(ph), s * math.sin(ph), z])
                    _, p = pyoracle.bxdf_eval(mat, fr, V, L)
                    acc += max(p, 0.0)
            expect[iz, ip] = acc / (sub * sub) * (2.0 / nz) * (2 * math.pi / nphi)
    mass = expect.sum()
    zc = -1 + 2 * (np.arange(nz) + 0.5) / nz
    rows = np.ones(nz, bool) if z_ok is None else z_ok(zc)
    ok = (expect * n > 20) & rows[:, None]
    chi2 = (((counts - expect * n) ** 2) / np.maximum(expect * n, 1e-30))[ok].sum()
    dof = int(ok.sum()) - 1
    p = stats.chi2.sf(chi2, dof)
    if return_maps:
        return p, mass, invalid / n, counts / n, expect, zc
    return p, mass, invalid / n


@pytest.mark.parametrize("kind,params", [
    ("diffuse", dict(roughness=0.0)),
    ("diffuse", dict(roughness=0.7)),
    ("disney", dict(metallic=0.0, roughness=0.6)),
    ("disney", dict(metallic=1.0, roughness=0.5)),
    ("disney", dict(metallic=1.0, roughness=0.5, anisotropic=0.8)),
    ("disney", dict(metallic=0.3, roughness=0.7, sheen=1.0, specular_tint=0.5)),
])
def test_sampler_matches_pdf(kind, params):
    m = _mat(kind, **params)
    V = _dir(math.radians(35), 0.4)
    p, mass, invalid = _chi_square(m, V)
    assert p > 1e-4, (p, mass)
    # directions the sampler draws with zero pdf (below the horizon) are the
    # missing mass of the reflection lobes
    assert abs((1 - invalid) - mass) < 0.02


def test_sampler_matches_pdf_rough_glass():
    # Refraction compresses the lobe into a narrow peak: a rough glass and fine
    # per-bin quadrature keep the expected counts accurate.  Entering with
    # eta = 1/1.5 the sampler cannot leave Snell's window (L.z <= -0.745 for
    # smooth facets), yet eval_microfacet_refraction (disney.glsl:190-219)
    # has no sign test on dot(L, H) and assigns pdf to grazing transmitted
    # directions too, so its pdf integrates above one.  That is the
    # reference's arithmetic, kept; the sampler is checked where it has
    # support and the extra eval mass is pinned where it has none.
    m = _mat("disney", metallic=0.0, roughness=0.8, specular_transmission=1.0)
    V = _dir(math.radians(30), 1.0)
    p, mass, invalid, obs, expect, zc = _chi_square(m, V, bins=(16, 12), sub=14, return_maps=True,
                                                    z_ok=lambda z: (z < -0.8) | (z > 0))
    assert p > 1e-4, (p, mass)
    grazing = (zc > -0.6) & (zc < 0)
    assert expect[grazing].sum() > 0.03  # eval mass the sampler rarely produces
    assert obs[grazing].sum() < 0.5 * expect[grazing].sum()
    assert mass > 1.03


def test_disney_mirror_limit():
    m = _mat("disney", metallic=1.0, roughness=0.0)  # ax = ay = 0.001
    fr = pyoracle.make_frame(N, T, B, eta=1 / 1.5)
    V = _dir(math.radians(40), 2.0)
    state = np.array([1, 2, 3, 4], np.uint32)
    mirror = np.array([-V[0], -V[1], V[2]])
    dev = []
    for _ in range(200):
        L, f, pdf, flags, state = pyoracle.bxdf_sample(m, fr, V, state)
        dev.append(np.linalg.norm(L - mirror))
        assert flags == 0x8
    # the spherical-cap VNDF sampler tilts H by ~alpha*sqrt(2/(1+z)), so a
    # few draws with z near -1 land farther out
    assert np.median(dev) < 3e-3 and np.quantile(dev, 0.9) < 1e-2


def test_roughness_regularization_only_in_eval():
    # ROUGHNESS_REGULARIZE_EVAL (make_shaders.yaml:4): at depth > 0 after a
    # non-specular bounce ax, ay < 0.3 become clamp(2a, 0.1, 0.3).
    m = _mat("disney", metallic=1.0, roughness=math.sqrt(0.04))  # ax = ay = 0.04
    m2 = m.copy()
    m2["ax"] = 0.1
    m2["ay"] = 0.1
    V = _dir(math.radians(30), 0.0)
    L = _dir(math.radians(32), math.pi + 0.05)
    f_reg, p_reg = pyoracle.bxdf_eval(m, pyoracle.make_frame(N, T, B, 1 / 1.5, depth=2, anb=True), V, L)
    f_ref, p_ref = pyoracle.bxdf_eval(m2, pyoracle.make_frame(N, T, B, 1 / 1.5, depth=0), V, L)
    assert np.allclose(f_reg, f_ref, rtol=1e-6) and math.isclose(p_reg, p_ref, rel_tol=1e-6)
    f_d0, _ = pyoracle.bxdf_eval(m, pyoracle.make_frame(N, T, B, 1 / 1.5, depth=0, anb=True), V, L)
    assert not np.allclose(f_d0, f_reg)


def test_sample_gtr1_keeps_reference_y_component():
    # sampling.glsl:43 returns (sin*cos(phi), sin*sin, cos): y = sin^2(theta).
    rng = np.random.default_rng(3)
    for _ in range(100):
        r1, r2 = rng.random(), rng.random()
        h = pyoracle.sample_gtr1(0.3, r1, r2)
        sin2 = 1.0 - float(h[2]) ** 2
        assert math.isclose(float(h[1]), sin2, rel_tol=1e-4, abs_tol=1e-6)
        assert math.isclose(float(h[0]), math.sqrt(max(sin2, 0)) * math.cos(r1 * 2 * math.pi), abs_tol=1e-5)


def test_cosine_and_hg_samplers():
    rng = np.random.default_rng(4)
    zs = []
    for _ in range(20000):
        r1, r2 = rng.random(), rng.random()
        d = pyoracle.cosine_sample_hemisphere(r1, r2)
        assert abs(np.linalg.norm(d) - 1) < 1e-5
        assert abs(d[2] - math.sqrt(1 - r1)) < 1e-5
        zs.append(d[2] ** 2)
    assert stats.kstest(zs, "uniform").pvalue > 1e-3  # cos^2 uniform for cosine weighting
    # Henyey-Greenstein with the reference's sign convention: pdf of
    # mu = dot(V, out) is 2*pi*phase_hg(mu) with denom 1 + g^2 + 2 g mu.
    for g in (-0.5, 0.3, 0.8):
        V = np.array([0.0, 0.6, 0.8])
        mus = np.array([float(np.dot(V, pyoracle.sample_hg(V, g, rng.random(), rng.random()))) for _ in range(20000)])
        grid = np.linspace(-1, 1, 20001)
        dens = (1 - g * g) / (4 * math.pi * (1 + g * g + 2 * g * grid) ** 1.5) * 2 * math.pi
        cdf = np.concatenate([[0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid))])
        assert abs(cdf[-1] - 1) < 1e-3
        assert stats.kstest(mus, lambda x: np.interp(x, grid, cdf)).pvalue > 1e-3
        # phase_hg integrates to one over the sphere
        vals = np.array([pyoracle.phase_hg(float(m), g) for m in grid[::4]])
        assert abs(np.trapezoid(vals * 2 * math.pi, grid[::4]) - 1) < 2e-3
