"""The closed-form scenes of tests/closed_form_scenes.py on the GPU: media
(pathtracer.rgen.glsl:157-191), the volume-MIS shadow walk
(inc/lighting.glsl:7-61) and NEE + MIS (inc/lighting.glsl:69-244) each meet
the same expectation the oracle is pinned to, and agree with the oracle path
by path on the same frames."""
import math

import numpy as np
import pytest

import closed_form_scenes as cf
import hpt
import pyoracle

pytestmark = pytest.mark.gpu

MED_T = hpt.FEATURE_MEDIUM | hpt.FEATURE_TRANSPARENT
MED_VMIS_T = hpt.FEATURE_MEDIUM | hpt.FEATURE_VOL_MIS | hpt.FEATURE_TRANSPARENT


def _renderer(scene, w, h, depth, features):
    print(f"[gpu] closed form {w}x{h} depth {depth} features {features}", flush=True)
    r = hpt.HalaRenderer(w, h, max_depth=depth, rr_depth=0, features=features, enable_tonemap=False)
    r.set_scene(scene)
    r.commit()
    return r


def _frames(r, frames):
    return np.stack([r.render_single_frame(f).astype(np.float64) for f in range(frames)])


def _agreement(gpu, ref, tol=1e-3):
    err = np.abs(gpu - ref) / (np.abs(ref) + 1e-2)
    return float((err.max(axis=-1) < tol).mean())


def _per_sample(scene, r, w, h, depth, features, frame=0):
    o = pyoracle.OracleScene(scene)
    ref = o.trace_paths(pyoracle.settings(w, h, depth, 0, features, tonemap=False), frame)[0]
    o.close()
    gpu = r.render_single_frame(frame)
    agree = _agreement(gpu, ref)
    print(f"[gpu] per-sample agreement {agree:.4f}", flush=True)
    assert agree >= 0.995
    return gpu


def _mean_within(vals, expect, k=4.0):
    v = vals.reshape(-1, 3)
    se = v.std(axis=0) / math.sqrt(len(v)) + 1e-12
    z = (v.mean(axis=0) - np.asarray(expect, np.float64)) / se
    print(f"[gpu] mean {v.mean(axis=0)} expected {np.asarray(expect)} ({z} SE)", flush=True)
    assert (np.abs(z) < k).all()


@pytest.mark.parametrize("features", [MED_T, MED_VMIS_T], ids=["medium", "vol_mis"])
@pytest.mark.parametrize("kind", ["absorb", "emissive"])
def test_card_absorbing_and_emissive_medium(kind, features):
    color, density, radius, rho, sky = (0.2, 0.5, 0.8), 1.3, 1.0, 0.6, 0.8
    mtype = hpt.MEDIUM_ABSORB if kind == "absorb" else hpt.MEDIUM_EMISSIVE
    sc = cf.card_in_sphere(mtype, color, density, radius=radius, card_albedo=rho, sky=sky)
    r = _renderer(sc, 16, 16, 1, features)
    gpu = _per_sample(sc, r, 16, 16, 1, features)
    f = cf.card_absorb if kind == "absorb" else cf.card_emissive
    assert np.allclose(gpu, f(rho, sky, color, density, radius)[None, None, :], rtol=1e-3)
    r.close()


@pytest.mark.parametrize("features", [MED_T, MED_VMIS_T], ids=["medium", "vol_mis"])
def test_card_single_scattering(features):
    sigma, c, g, radius, rho, sky = 1.5, 0.7, 0.3, 1.0, 0.6, 0.8
    sc = cf.card_in_sphere(hpt.MEDIUM_SCATTER, (c,) * 3, sigma, anisotropy=g, radius=radius, card_albedo=rho,
                           sky=sky)
    r = _renderer(sc, 96, 96, 2, features)
    _per_sample(sc, r, 96, 96, 2, features)
    _mean_within(_frames(r, 24), [cf.card_scatter(rho, sky, c, sigma, g, radius, 2)] * 3)
    r.close()


@pytest.mark.parametrize("kind", ["absorb", "scatter"])
def test_slab_transmittance(kind):
    color, density = ((0.2, 0.5, 0.8), 1.5) if kind == "absorb" else ((0.0, 0.0, 0.0), 0.8)
    mtype = hpt.MEDIUM_ABSORB if kind == "absorb" else hpt.MEDIUM_SCATTER
    k = (1.0 - np.asarray(color)) * density if kind == "absorb" else np.full(3, density)
    sc = cf.slab_point_light(mtype, color, density)
    w = h = 24
    r = _renderer(sc, w, h, 2, MED_VMIS_T)
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(w, h, 2, 0, MED_VMIS_T, tonemap=False)
    ys, xs = np.mgrid[0:h, 0:w]
    for frame in (0, 3):
        gpu = r.render_single_frame(frame).reshape(-1, 3).astype(np.float64)
        p = cf.floor_points(o.camera_rays(st, frame, xs.ravel(), ys.ravel()))
        assert np.allclose(gpu, cf.slab_direct(p, k), rtol=2e-4, atol=0)
    r.close()
    o.close()
    # the walk takes at most max_depth steps: none at depth 0
    sc.materials[1]["opacity"] = 1.0
    r0 = _renderer(sc, w, h, 0, MED_VMIS_T)
    o0 = pyoracle.OracleScene(sc)
    st0 = pyoracle.settings(w, h, 0, 0, MED_VMIS_T, tonemap=False)
    gpu = r0.render_single_frame(0).reshape(-1, 3).astype(np.float64)
    p = cf.floor_points(o0.camera_rays(st0, 0, xs.ravel(), ys.ravel()))
    assert np.allclose(gpu, cf.slab_direct(p, np.zeros(3)), rtol=2e-4, atol=0)
    r0.close()
    o0.close()


@pytest.mark.parametrize("features", [0, MED_T, MED_VMIS_T], ids=["plain", "medium", "vol_mis"])
def test_env_lit_plane(features):
    L, rho = np.array([0.9, 0.7, 0.4]), 0.5
    sc = cf.env_lit_plane(tuple(L), rho)
    r = _renderer(sc, 64, 64, 1, features)
    _per_sample(sc, r, 64, 64, 1, features)
    _mean_within(_frames(r, 32), rho * L)
    r.close()


@pytest.mark.parametrize("features", [0, MED_VMIS_T], ids=["plain", "vol_mis"])
def test_sphere_lights(features):
    from scipy.interpolate import RegularGridInterpolator

    vmis = bool(features & hpt.FEATURE_VOL_MIS)
    sc = cf.sphere_lights_plane()
    w = h = 32
    r = _renderer(sc, w, h, 1, features)
    _per_sample(sc, r, w, h, 1, features)
    g = np.linspace(-0.75, 0.75, 31)
    gx, gy = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    trait = RegularGridInterpolator((g, g), cf.sphere_lights_expected(pts, n_radial=128, n_phi=64).reshape(31, 31, 3))
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(w, h, 1, 0, features, tonemap=False)
    ys, xs = np.mgrid[0:h, 0:w]
    dev = []
    for f in range(96):
        gpu = r.render_single_frame(f).reshape(-1, 3).astype(np.float64)
        p = cf.floor_points(o.camera_rays(st, f, xs.ravel(), ys.ravel()))
        dev.append(gpu - (cf.sphere_lights_radiance(p) if vmis else trait(p[:, :2])))
    o.close()
    r.close()
    _mean_within(np.stack(dev), np.zeros(3))
