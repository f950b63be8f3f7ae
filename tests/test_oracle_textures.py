"""Oracle pins for the material maps of pathtracer.rgen.glsl:121-146: the
texture look-up (bilinear, REPEAT, sRGB decode of colour maps) from exact
texel values, and the maps' effect on the integrator from closed forms."""
import math

import numpy as np

import hpt
import hpt_scenes
import pyoracle


def _scene_with(textures):
    sc = hpt_scenes.cornell_box()
    sc.textures = textures
    return pyoracle.OracleScene(sc)


def _srgb(code):
    c = code / 255.0
    return np.float32(c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4)


def test_texel_centres_are_exact_and_srgb_decodes_per_texel():
    t = np.zeros((2, 4, 4), np.uint8)
    t[..., 0] = [[0, 64, 128, 255], [10, 20, 30, 40]]
    t[..., 1] = 200
    t[..., 3] = [[255, 128, 0, 7], [1, 2, 3, 4]]
    o = _scene_with([t])
    for j in range(2):
        for i in range(4):
            u, v = (i + 0.5) / 4, (j + 0.5) / 2
            lin = o.texture_sample(0, u, v, False)
            assert lin[0] == np.float32(t[j, i, 0] / 255.0) and lin[3] == np.float32(t[j, i, 3] / 255.0)
            s = o.texture_sample(0, u, v, True)
            assert s[0] == _srgb(t[j, i, 0]) and s[1] == _srgb(200)
            assert s[3] == lin[3]  # alpha is never sRGB
    assert o.texture_sample(0, 0.875, 0.25, True)[0] == np.float32(1.0)


def test_bilinear_weights_and_repeat_wrap():
    t = np.zeros((2, 2, 4), np.uint8)
    t[..., 0] = [[0, 200], [100, 40]]
    o = _scene_with([t])
    # the centre of the texture mixes all four texels equally
    assert math.isclose(o.texture_sample(0, 0.5, 0.5, False)[0], (0 + 200 + 100 + 40) / 4 / 255, rel_tol=1e-6)
    # u = 0 sits half way between the last and the first column (REPEAT)
    a = o.texture_sample(0, 0.0, 0.25, False)[0]
    assert math.isclose(a, (0 + 200) / 2 / 255, rel_tol=1e-6)
    # one period further is the same look-up
    for u, v in ((0.1, 0.3), (0.77, 0.9)):
        assert np.allclose(o.texture_sample(0, u, v, False), o.texture_sample(0, u + 1.0, v - 1.0, False), atol=1e-6)


def test_constant_base_colour_map_gives_the_furnace_closed_form():
    # A closed emitting cube whose base colour comes from a constant sRGB map
    # returns Le * sum_k rho^k with rho the decoded texel (the texture replaces
    # base_color, pathtracer.rgen.glsl:122-124).
    code = 180
    rho = float(_srgb(code))
    sc = hpt_scenes.furnace_box(albedo=0.1, emission=1.0)
    sc.materials["base_color_map_index"][0] = 0
    t = np.full((4, 4, 4), code, np.uint8)
    sc.textures = [t]
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(32, 32, 3, 0, tonemap=False)
    rad, alb, _ = o.trace_paths(st, 0)
    expect = sum(rho ** k for k in range(4))
    on = np.isclose(rad, expect, rtol=1e-5).all(axis=-1)
    assert on.mean() >= 0.998
    assert np.allclose(alb, rho)


def test_normal_map_tilts_the_shading_normal():
    # A flat-out texel (128,128,255) leaves the normal within 1e-3; a tilted
    # texel turns the first-hit normal AOV toward the tangent.
    sc = hpt_scenes.cornell_box()
    flat = np.zeros((1, 1, 4), np.uint8)
    flat[0, 0] = (128, 128, 255, 255)
    tilt = np.zeros((1, 1, 4), np.uint8)
    tilt[0, 0] = (255, 128, 40, 255)
    mats = np.concatenate([sc.materials, hpt.new_materials(1)])
    hpt_scenes.diffuse_material(mats[3], (0.73, 0.73, 0.73))
    mats[3]["normal_map_index"] = 0
    sc.materials = mats
    sc.primitives[2].material_index = 3  # back wall: shading normal e1 x e2 = +y, tangent e1 = +z
    st = pyoracle.settings(16, 16, 2, 0, tonemap=False)
    normals = []
    for t in (flat, tilt):
        sc.textures = [t]
        _, _, nrm = pyoracle.OracleScene(sc).trace_paths(st, 0, rect=(8, 7, 9, 8))
        normals.append(nrm[0, 0])
    assert np.allclose(normals[0], [0, 1, 0], atol=1e-3)
    n_ts = np.array([1.0, 128 / 255, 40 / 255 * 5]) * 2 - 1
    n_ts /= np.linalg.norm(n_ts)
    # world: tangent +z, bitangent = cross(n, t) = cross(+y, +z) = +x, normal +y
    expect = n_ts[0] * np.array([0, 0, 1.0]) + n_ts[1] * np.array([1.0, 0, 0]) + n_ts[2] * np.array([0, 1.0, 0])
    assert np.allclose(normals[1], expect / np.linalg.norm(expect), atol=2e-3)
