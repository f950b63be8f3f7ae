"""Oracle pins: RNG (common/wltcRS.glsl), cameras (perspective_ray.rcall.glsl,
orthographic_ray.rcall.glsl), tonemaps and sRGB (common/tonemap.glsl,
common/common.glsl)."""
import json
import math
import os

import numpy as np

import hpt
import hpt_scenes
import pyoracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pcg4d.json")


def test_pcg4d_golden_vectors():
    g = json.load(open(GOLDEN))
    for case in g["pcg4d"]:
        assert pyoracle.pcg4d(case["in"]).tolist() == case["out"]
    chain = g["pcg4d_chain"]
    for a, b in zip(chain, chain[1:]):
        assert pyoracle.pcg4d(a).tolist() == b


def test_rng_initialize_golden():
    for case in json.load(open(GOLDEN))["rng_initialize"]:
        s0, s1 = pyoracle.rng_init(case["pixel"][0], case["pixel"][1], case["frame"])
        assert s0.tolist() == case["s0"] and s1.tolist() == case["s1"]


def test_rand_is_round_to_nearest_over_2_pow_32():
    # float(u) / float(0xffffffff): the divisor rounds to 2^32, so rand() can be 1.0.
    assert pyoracle.u32_to_unit(0) == 0.0
    assert pyoracle.u32_to_unit(0x80000000) == 0.5
    assert pyoracle.u32_to_unit(0xFFFFFFFF) == 1.0
    assert pyoracle.u32_to_unit(0xFFFFFF80) == 1.0  # rounds up to 2^32
    assert pyoracle.u32_to_unit(0xFFFFFF7F) < 1.0
    assert pyoracle.u32_to_unit(12345) == np.float32(12345) / np.float32(2.0 ** 32)


def _pinhole(yfov_deg=39.0, aperture=0.0):
    sc = hpt_scenes.cornell_box()
    sc.cameras = hpt_scenes.look_at_camera((0, -3.85, 1), (0, 0, 1), yfov_deg=yfov_deg, aperture=aperture,
                                           focal_distance=3.85)
    return sc


def test_pinhole_rays_within_one_pixel_tent():
    w, h = 64, 48
    o = pyoracle.OracleScene(_pinhole())
    st = pyoracle.settings(w, h, 8, 3)
    ys, xs = np.mgrid[0:h, 0:w]
    rays = o.camera_rays(st, 3, xs.ravel(), ys.ravel())
    d = rays["direction"].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.allclose(rays["origin"], [0, -3.85, 1], atol=1e-6)
    # Project back to the image plane: d = normalize(right*dx + up*dy + fwd),
    # right=+x, up=+z, fwd=+y; dx, dy in units of tan(fov/2).
    s = math.tan(math.radians(39.0) * 0.5)
    px = d[:, 0] / d[:, 1] / (s * w / h)
    py = d[:, 2] / d[:, 1] / s
    # screen_uv = (x, h - y)/res without a half-pixel offset, jitter in (-1, 1) pixels
    jx = (px + 1) / 2 * w - xs.ravel()
    jy = (py + 1) / 2 * h - (h - ys.ravel())
    assert np.all(np.abs(jx) <= 1.0 + 1e-3) and np.all(np.abs(jy) <= 1.0 + 1e-3)
    # row 0 is the top of the image
    assert py[0] > 0.9 and py[-1] < -0.9


def test_blue_noise_shift_is_per_frame_and_wraps_at_1024():
    # rand4blue reads texel (pixel + shift(frame)) % 1024: pixels 1024 apart
    # in the same frame get the same jitter.
    w, h = 1100, 2
    o = pyoracle.OracleScene(_pinhole())
    st = pyoracle.settings(w, h, 8, 3)
    r = o.camera_rays(st, 5, [10, 1034], [1, 1])
    s = math.tan(math.radians(39.0) * 0.5)
    px = r["direction"][:, 0] / r["direction"][:, 1] / (s * w / h)
    j = (px.astype(np.float64) + 1) / 2 * w - np.array([10, 1034])
    assert abs(j[0] - j[1]) < 1e-3


def test_thin_lens_rays_focus():
    # A thin-lens ray leaves the lens disc and passes through the point the
    # pinhole ray of the same pixel and frame reaches at focal_distance.
    w, h = 32, 32
    lens = pyoracle.OracleScene(_pinhole(aperture=0.05))
    pin = pyoracle.OracleScene(_pinhole(aperture=0.0))
    st = pyoracle.settings(w, h, 8, 3)
    radii = []
    for f in range(32):
        r = lens.camera_rays(st, f, [16], [16])[0]
        p = pin.camera_rays(st, f, [16], [16])[0]
        org, d = r["origin"].astype(np.float64), r["direction"].astype(np.float64)
        focal = np.array([0, -3.85, 1]) + 3.85 * p["direction"].astype(np.float64)
        radii.append(np.linalg.norm(org - [0, -3.85, 1]))
        v = focal - org
        assert np.linalg.norm(v - np.dot(v, d) * d) < 1e-5
    assert max(radii) <= math.sqrt(0.05) + 1e-6  # lens radius sqrt(aperture * u)
    assert max(radii) > 0.3 * math.sqrt(0.05)


def test_orthographic_rays_are_parallel():
    sc = _pinhole()
    cam = sc.cameras.copy()
    cam["type"] = hpt.CAMERA_ORTHOGRAPHIC
    cam["focal_distance_or_xmag"] = 2.0
    cam["aperture_or_ymag"] = 1.5
    sc.cameras = cam
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(16, 16, 8, 3)
    ys, xs = np.mgrid[0:16, 0:16]
    r = o.camera_rays(st, 0, xs.ravel(), ys.ravel())
    assert np.allclose(r["direction"], [0, 1, 0], atol=1e-7)
    assert r["origin"][:, 0].min() >= -2.0 - 0.3 and r["origin"][:, 0].max() <= 2.0 + 0.3
    assert np.allclose(r["origin"][:, 1], -3.85)


def test_srgb_and_tonemaps_closed_forms():
    assert pyoracle.linear_2_srgb(0.0) == 0.0
    assert abs(pyoracle.linear_2_srgb(1.0) - 1.0) < 1e-6
    a = pyoracle.linear_2_srgb(0.0031308)
    b = pyoracle.linear_2_srgb(0.0031309)
    assert abs(a - b) < 1e-4  # the two branches meet
    # Reinhard on luminance with limit 1.5 (tonemap.glsl:45-47), gray input.
    for g in (0.1, 1.0, 10.0):
        out = pyoracle.tonemap("reinhard", [g, g, g])
        assert np.allclose(out, g / (1 + g / 1.5), rtol=1e-6)
    # Narkowicz ACES: f(1) = 2.54/3.16, saturates to 1.
    assert np.allclose(pyoracle.tonemap("aces", [1, 1, 1]), 2.54 / 3.16, rtol=1e-6)
    assert np.allclose(pyoracle.tonemap("aces", [1e4, 1e4, 1e4]), 1.0)
    # Fitted ACES: both matrices have unit row sums, so gray stays gray and
    # equals the RRT/ODT fit; a transposed matrix would tint it.
    for g in (0.05, 0.18, 0.5, 2.0):
        out = pyoracle.tonemap("aces_fitted", [g, g, g])
        fit = (g * (g + 0.0245786) - 0.000090537) / (g * (0.983729 * g + 0.4329510) + 0.238081)
        assert np.allclose(out, min(fit, 1.0), rtol=2e-5)
    assert pyoracle.tonemap("aces_fitted", [1, 0, 0])[0] > pyoracle.tonemap("aces_fitted", [1, 0, 0])[1]
