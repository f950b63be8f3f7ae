"""Closed-form pins of the oracle's media, volume-MIS transmittance and
next-event estimation with MIS.

* Media (pathtracer.rgen.glsl:157-191,233-243): ABSORB and EMISSIVE give
  exact per-path radiance around a card at the centre of a sphere; SCATTER
  matches the escape probability exp(-sigma R) at depth 1, the
  single-scattering double integral at depth 2 (distance sampling, the
  albedo factor and the Henyey-Greenstein direction all enter), and energy
  conservation at albedo 1.
* eval_transmittance (inc/lighting.glsl:7-61): Beer-Lambert over the chord of
  a slab behind an opacity-0 boundary, for an absorbing (1 - c) and a
  scattering (colour 1) medium; without VOL_MIS the slab blocks the shadow
  ray.
* NEE + MIS (inc/lighting.glsl:69-244, env_map_eval.rcall.glsl:23,
  lights.rint.glsl): a Lambert plane under a constant environment map returns
  albedo * L; two sphere lights over a plane return their closed-form
  irradiance.  Both hold only if the light-sampling and BSDF-sampling
  weights of every direction sum to one.
The expected values are in tests/closed_form_scenes.py.
"""
import math

import numpy as np
import pytest

import closed_form_scenes as cf
import hpt
import pyoracle

MED_T = hpt.FEATURE_MEDIUM | hpt.FEATURE_TRANSPARENT
MED_VMIS_T = hpt.FEATURE_MEDIUM | hpt.FEATURE_VOL_MIS | hpt.FEATURE_TRANSPARENT


def _frames(scene, w, h, depth, features, frames, rr=0):
    o = pyoracle.OracleScene(scene)
    st = pyoracle.settings(w, h, depth, rr, features, tonemap=False)
    out = [o.trace_paths(st, f) for f in range(frames)]
    o.close()
    rad = np.stack([r[0] for r in out]).astype(np.float64)
    alb = np.stack([r[1] for r in out])
    return rad, alb


def _mean_within(vals, expect, k=4.0):
    """Mean of independent paths within k standard errors of `expect` (per
    channel); returns the deviations in standard errors."""
    v = vals.reshape(-1, 3)
    se = v.std(axis=0) / math.sqrt(len(v)) + 1e-12
    z = (v.mean(axis=0) - np.asarray(expect, np.float64)) / se
    assert (np.abs(z) < k).all(), (v.mean(axis=0), expect, z)
    return z


# ----------------------------------------------------------------- media

@pytest.mark.parametrize("features", [MED_T, MED_VMIS_T], ids=["medium", "vol_mis"])
@pytest.mark.parametrize("kind", ["absorb", "emissive"])
def test_card_absorbing_and_emissive_medium_per_path(kind, features):
    color, density, radius, rho, sky = (0.2, 0.5, 0.8), 1.3, 1.0, 0.6, 0.8
    mtype = hpt.MEDIUM_ABSORB if kind == "absorb" else hpt.MEDIUM_EMISSIVE
    sc = cf.card_in_sphere(mtype, color, density, radius=radius, card_albedo=rho, sky=sky)
    rad, alb = _frames(sc, 16, 16, 1, features, 2)
    assert np.allclose(alb, rho)  # every pixel sees the card
    f = cf.card_absorb if kind == "absorb" else cf.card_emissive
    expect = f(rho, sky, color, density, radius)
    # tessellation of the sphere (<= 1.5e-4 R) and the card's extent
    # (<= 1e-4) move the chord by a few 1e-4
    assert np.allclose(rad, expect[None, None, None, :], rtol=1e-3), (rad.min(axis=(0, 1, 2)), expect)
    # the medium changes the result: without it the card just reflects the sky
    assert np.abs(expect - rho * sky).min() > 0.05


@pytest.mark.parametrize("features", [MED_T, MED_VMIS_T], ids=["medium", "vol_mis"])
def test_card_scattering_medium_escape_probability(features):
    # depth 1: a path scattered in the medium ends there (the scattered
    # segment would be vertex 2), so only unscattered crossings reach the sky
    sigma, c, g, radius, rho, sky = 1.2, 0.6, 0.3, 1.0, 0.6, 0.8
    sc = cf.card_in_sphere(hpt.MEDIUM_SCATTER, (c,) * 3, sigma, anisotropy=g, radius=radius, card_albedo=rho,
                           sky=sky)
    rad, _ = _frames(sc, 64, 64, 1, features, 8)
    expect = cf.card_scatter(rho, sky, c, sigma, g, radius, 1)
    _mean_within(rad, [expect] * 3)
    assert set(np.unique(np.round(rad, 5))) <= {0.0, round(rho * sky, 5)}


@pytest.mark.parametrize("features", [MED_T, MED_VMIS_T], ids=["medium", "vol_mis"])
@pytest.mark.parametrize("g", [0.3, -0.5])
def test_card_scattering_medium_single_scattering_integral(g, features):
    sigma, c, radius, rho, sky = 1.5, 0.7, 1.0, 0.6, 0.8
    sc = cf.card_in_sphere(hpt.MEDIUM_SCATTER, (c,) * 3, sigma, anisotropy=g, radius=radius, card_albedo=rho,
                           sky=sky)
    rad, _ = _frames(sc, 96, 96, 2, features, 12)
    expect = cf.card_scatter(rho, sky, c, sigma, g, radius, 2)
    z = _mean_within(rad, [expect] * 3)
    # The test has the power to see the phase function's sign, the albedo
    # factor and the density: each alternative is far outside 4 SE.
    se = rad.reshape(-1, 3)[:, 0].std() / math.sqrt(rad[..., 0].size)
    for other in (cf.card_scatter(rho, sky, c, sigma, -g, radius, 2),
                  cf.card_scatter(rho, sky, c * c, sigma, g, radius, 2),
                  cf.card_scatter(rho, sky, c, sigma * 1.1, g, radius, 2)):
        assert abs(other - expect) > 8 * se, (other, expect, se)
    print(f"single scattering g={g}: mean {rad.mean():.5f} expect {expect:.5f} ({z[0]:+.2f} SE)")


def test_card_conservative_medium_conserves_energy():
    # albedo 1: every path eventually leaves the sphere with its throughput
    # (the card's albedo), whatever the distance or phase sampling
    sigma, g, radius, rho, sky = 1.5, 0.4, 1.0, 0.6, 0.8
    sc = cf.card_in_sphere(hpt.MEDIUM_SCATTER, (1.0, 1.0, 1.0), sigma, anisotropy=g, radius=radius,
                           card_albedo=rho, sky=sky)
    rad, _ = _frames(sc, 48, 48, 64, MED_VMIS_T, 4)
    assert np.allclose(rad.reshape(-1, 3).mean(axis=0), rho * sky, rtol=2e-3)
    assert (rad <= rho * sky * (1 + 1e-5)).all()


# ----------------------------------------------------- volume-MIS transmittance

@pytest.mark.parametrize("kind", ["absorb", "scatter"])
def test_slab_transmittance_beer_lambert(kind):
    # (the scattering medium has albedo 0, so the floor's bounce, which
    # enters the slab, adds no light scattered back from inside it)
    color, density = ((0.2, 0.5, 0.8), 1.5) if kind == "absorb" else ((0.0, 0.0, 0.0), 0.8)
    mtype = hpt.MEDIUM_ABSORB if kind == "absorb" else hpt.MEDIUM_SCATTER
    # ABSORB attenuates by 1 - colour, every other medium by colour 1
    k = (1.0 - np.asarray(color)) * density if kind == "absorb" else np.full(3, density)
    sc = cf.slab_point_light(mtype, color, density)
    w = h = 24
    o = pyoracle.OracleScene(sc)
    # Depth 2: the walk may take max_depth closest-hit steps (entry face, exit
    # face); the floor's bounce leaves upward through the slab into a black
    # sky, so the pixel holds the direct term of the first vertex only.
    st = pyoracle.settings(w, h, 2, 0, MED_VMIS_T, tonemap=False)
    ys, xs = np.mgrid[0:h, 0:w]
    for frame in (0, 5):
        rad = o.trace_paths(st, frame)[0].reshape(-1, 3).astype(np.float64)
        p = cf.floor_points(o.camera_rays(st, frame, xs.ravel(), ys.ravel()))
        expect = cf.slab_direct(p, k)
        assert np.allclose(rad, expect, rtol=1e-4, atol=0), np.abs(rad / expect - 1).max()
    # The walk's loop runs `depth < max_depth` times (inc/lighting.glsl:19):
    # at max_depth 0 it traces nothing and returns 1, even through an opaque
    # slab.
    sc.materials[1]["opacity"] = 1.0
    o0 = pyoracle.OracleScene(sc)
    st0 = pyoracle.settings(w, h, 0, 0, MED_VMIS_T, tonemap=False)
    rad = o0.trace_paths(st0, 0)[0].reshape(-1, 3).astype(np.float64)
    p = cf.floor_points(o0.camera_rays(st0, 0, xs.ravel(), ys.ravel()))
    assert np.allclose(rad, cf.slab_direct(p, np.zeros(3)), rtol=1e-4, atol=0)
    # ... and at depth 2 the opaque slab blocks the light
    assert (o0.trace_paths(st, 0)[0] == 0).all()
    o0.close()
    # without volume MIS the shadow ray is an any-hit query and the slab's
    # boundary blocks it
    st2 = pyoracle.settings(w, h, 2, 0, MED_T, tonemap=False)
    assert (o.trace_paths(st2, 0)[0] == 0).all()
    o.close()


# ------------------------------------------------------------------ NEE + MIS

@pytest.mark.parametrize("features", [0, MED_T, MED_VMIS_T], ids=["plain", "medium", "vol_mis"])
def test_env_lit_plane_nee_and_bsdf_weights_sum_to_one(features):
    L, rho = np.array([0.9, 0.7, 0.4]), 0.5
    sc = cf.env_lit_plane(tuple(L), rho)
    rad, _ = _frames(sc, 48, 48, 1, features, 16)
    _mean_within(rad, rho * L)
    # the BSDF half matters: environment NEE alone (depth 0) falls short
    nee, _ = _frames(sc, 48, 48, 0, features, 4)
    assert (nee.reshape(-1, 3).mean(axis=0) < 0.95 * rho * L).all()


def _sphere_lights_expectation_on_grid():
    """The reference estimator's expectation (tests/closed_form_scenes.py,
    numerical quadrature) on a grid over the floor the camera sees, as a
    bilinear interpolant, with the ideal irradiance for comparison."""
    from scipy.interpolate import RegularGridInterpolator

    g = np.linspace(-0.75, 0.75, 31)
    gx, gy = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    vals = cf.sphere_lights_expected(pts, n_radial=128, n_phi=64).reshape(31, 31, 3)
    interp = RegularGridInterpolator((g, g), vals)
    return lambda p: interp(p[:, :2])


@pytest.fixture(scope="module")
def sphere_lights_expectation():
    return _sphere_lights_expectation_on_grid()


@pytest.mark.parametrize("features", [0, MED_T, MED_VMIS_T], ids=["plain", "medium", "vol_mis"])
def test_sphere_lights_nee_and_bsdf_weights_sum_to_one(features, sphere_lights_expectation):
    """Light samples and BSDF hits of two sphere lights (pick probability
    1/2 with the emission doubled, power-heuristic weights on the undivided
    pdfs) sum to the sphere's irradiance integral.  Without volume MIS the
    shadow ray is an any-hit query up to dist - 2 EPS from an origin offset
    along the floor normal, and samples near the light's lower rim meet the
    light's own sphere first: a reference trait (4-6 % darker here) that the
    quadrature of closed_form_scenes.sphere_lights_expected includes.  With
    volume MIS the shadow walk stops at the first emitter it meets
    (inc/lighting.glsl:36-37), the light itself, so the result is the ideal
    pi Le sin^2(alpha) cos(theta)."""
    vmis = bool(features & hpt.FEATURE_VOL_MIS)
    sc = cf.sphere_lights_plane()
    w = h = 32
    frames = 96
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(w, h, 1, 0, features, tonemap=False)
    ys, xs = np.mgrid[0:h, 0:w]
    dev, dev_other = [], []
    # each frame has its own jitter: compare every path with the expectation
    # at its own floor point
    for f in range(frames):
        rad = o.trace_paths(st, f)[0].reshape(-1, 3).astype(np.float64)
        p = cf.floor_points(o.camera_rays(st, f, xs.ravel(), ys.ravel()))
        ideal, trait = cf.sphere_lights_radiance(p), sphere_lights_expectation(p)
        dev.append(rad - (ideal if vmis else trait))
        dev_other.append(rad - (trait if vmis else ideal))
    o.close()
    d = np.stack(dev)  # (frames, pixels, 3)
    _mean_within(d, np.zeros(3))
    # and in each 8x8-pixel block (the deviation does not vary over the floor)
    blocks = d.reshape(frames, h // 8, 8, w // 8, 8, 3).transpose(1, 3, 0, 2, 4, 5).reshape(-1, frames * 64, 3)
    for b in blocks:
        _mean_within(b, np.zeros(3), k=4.5)
    # the other expectation is far outside the estimate's error
    di = np.stack(dev_other).reshape(-1, 3)
    z_other = di.mean(axis=0) / (di.std(axis=0) / math.sqrt(len(di)))
    assert (np.abs(z_other) > 8).all(), z_other
