"""Oracle pins for ray queries and the integrator (pathtracer.rgen.glsl).

Ray queries: the watertight FP32 test against an independent FP64
Moller-Trumbore over all triangles, the oracle's BVH2 against its brute
force, and watertightness on shared edges and vertices.
Integrator: closed forms (a closed emitting Lambert sphere), Russian-roulette
unbiasedness, and the reference's depth/transparency/emitter semantics.
"""
import math

import numpy as np
import pytest

import hpt
import hpt_scenes
import pyoracle


def _rays(n, seed, lo=-0.5, hi=1.5):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    t = rng.uniform(0, 1, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return hpt.make_rays(o, d)


def _same(a, b):
    return (a["kind"] == b["kind"]) & (a["instance"] == b["instance"]) & (a["primitive"] == b["primitive"])


def test_fp32_watertight_hits_match_fp64():
    sc = hpt_scenes.triangle_soup(3000, seed=1)
    o = pyoracle.OracleScene(sc)
    rays = _rays(3000, 2)
    h = o.trace_closest(rays, brute=True)
    t64, tri64, gap = o.trace_fp64(rays)
    hit64 = np.isfinite(t64)
    clear = hit64 & (gap > 1e-4)
    assert clear.sum() > 300
    assert (h["kind"][clear] == hpt.HIT_TRIANGLE).all()
    assert np.array_equal(h["primitive"][clear], tri64[clear])  # one primitive: triangle index = global id
    rel = np.abs(h["t"][clear] - t64[clear]) / t64[clear]
    assert rel.max() < 1e-5
    # misses agree except within float noise of an edge
    assert ((h["kind"] == hpt.HIT_TRIANGLE) == hit64).mean() > 0.999


def test_bvh2_matches_brute_force():
    sc = hpt_scenes.triangle_soup(5000, seed=3)
    o = pyoracle.OracleScene(sc)
    rays = _rays(4000, 4)
    a, b = o.trace_closest(rays), o.trace_closest(rays, brute=True)
    assert _same(a, b).all() and np.array_equal(a["t"], b["t"])
    rays["tmax"] = np.random.default_rng(5).uniform(0.01, 1.5, len(rays)).astype(np.float32)
    assert np.array_equal(o.trace_any(rays), o.trace_any(rays, brute=True))


def test_watertight_on_edges_and_vertices():
    sc = hpt_scenes.furnace_sphere(segments=24, rings=12)
    o = pyoracle.OracleScene(sc)
    tri = sc.world_triangles().astype(np.float64)
    tg = np.concatenate([tri.reshape(-1, 3), 0.5 * (tri[:, 0] + tri[:, 1]), 0.5 * (tri[:, 1] + tri[:, 2])])
    org = np.zeros_like(tg) + [0.013, -0.021, 0.034]
    d = tg - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    h = o.trace_closest(hpt.make_rays(org, d), brute=True)
    assert (h["kind"] == hpt.HIT_TRIANGLE).all()


def test_light_ties_go_to_triangles():
    # A triangle exactly in the plane of the quad light: equal t, the
    # triangle wins (the light only replaces a strictly farther hit).
    sc = hpt_scenes.cornell_box()
    v, ix = hpt_scenes.quad_mesh((-0.125, -0.125, 1.999), (0.25, 0, 0), (0, 0.25, 0))
    sc.primitives.append(hpt.Primitive(v, ix, 0))
    o = pyoracle.OracleScene(sc)
    r = hpt.make_rays([[0.01, 0.02, 1.0]], [[0.0, 0.0, 1.0]])
    h = o.trace_closest(r, brute=True)
    assert h["kind"][0] == hpt.HIT_TRIANGLE and h["instance"][0] == len(sc.primitives) - 1


def _paths(sc, w=24, h=24, depth=5, rr=0, features=0, frame=0, tonemap=False):
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(w, h, depth, rr, features, tonemap=tonemap)
    return o.trace_paths(st, frame)


@pytest.mark.parametrize("depth", [0, 1, 2, 7])
def test_furnace_closed_form(depth):
    rho, le = 0.6, 2.0
    rad, alb, nrm = _paths(hpt_scenes.furnace_box(albedo=rho, emission=le), 48, 48, depth=depth)
    expect = le * sum(rho ** k for k in range(depth + 1))
    # exact except a rare path whose offset origin crosses a box edge and
    # escapes (it can only lose energy)
    on = np.isclose(rad, expect, rtol=1e-5).all(axis=-1)
    assert on.mean() >= 0.998 and rad.max() <= expect * (1 + 1e-5)
    assert np.allclose(alb, rho)  # first-hit albedo = base colour


def test_furnace_russian_roulette_unbiased():
    rho, le, depth = 0.7, 1.0, 10
    sc = hpt_scenes.furnace_box(albedo=rho, emission=le)
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(48, 48, depth, 2, tonemap=False)
    vals = np.concatenate([o.trace_paths(st, f)[0].ravel() for f in range(6)])
    expect = le * sum(rho ** k for k in range(depth + 1))
    assert abs(vals.mean() - expect) < 4 * vals.std() / math.sqrt(len(vals))
    assert vals.std() > 0  # roulette made paths differ


def test_emitter_seen_directly_has_unit_mis_weight():
    # Depth-0 hit on the quad light: radiance = intensity exactly
    # (pathtracer.rgen.glsl:96-107).
    sc = hpt_scenes.cornell_box()
    sc.cameras = hpt_scenes.look_at_camera((0.0, 0.0, 1.0), (0.0, 0.0, 2.0), up=(0, 1, 0), yfov_deg=2.0)
    rad, _, _ = _paths(sc, 8, 8)
    assert np.allclose(rad, [17.0, 12.0, 4.0])


def test_transparent_surface_at_depth_zero_ends_the_path():
    # depth is a uint: the pass-through depth-- at depth 0 wraps and the
    # max_depth test ends the path (pathtracer.rgen.glsl:196-199,247), leaving
    # only the surface's own emission.
    sc = hpt_scenes.cornell_box()
    mats = np.concatenate([sc.materials, hpt.new_materials(1)])
    hpt_scenes.diffuse_material(mats[3], (0.5, 0.5, 0.5), emission=(0.25, 0.5, 0.75))
    mats[3]["opacity"] = 0.0
    sc.materials = mats
    # a sheet inside the box, across the whole view, in front of the back wall
    v, ix = hpt_scenes.quad_mesh((-3, 0.5, -2), (6, 0, 0), (0, 0, 6))
    sc.primitives.append(hpt.Primitive(v, ix, 3))
    sc.cameras = hpt_scenes.look_at_camera((0.0, -0.5, 1.6), (0.0, 1.0, 1.6), yfov_deg=20.0)  # above the boxes
    rad, alb, _ = _paths(sc, 16, 16, depth=8, rr=3, features=hpt.FEATURE_TRANSPARENT)
    assert np.allclose(rad, [0.25, 0.5, 0.75])
    assert np.allclose(alb, 0.5)
    # without the TRANSPARENT permutation the sheet is opaque and lit
    rad2, _, _ = _paths(sc, 16, 16, depth=8, rr=3, features=0)
    assert (rad2.mean(axis=(0, 1)) > np.array([0.25, 0.5, 0.75]) + 0.01).all()


def test_invalid_material_paints_magenta():
    sc = hpt_scenes.cornell_box()
    sc.primitives[2] = hpt.Primitive(sc.primitives[2].vertices, sc.primitives[2].indices, hpt.INVALID_INDEX)
    rad, _, _ = _paths(sc, 32, 32, depth=3)
    assert np.isclose(rad, [0.7, 0.0, 0.7]).all(axis=-1).mean() > 0.2  # the back wall


def test_cornell_energy_and_colour_bleeding():
    rad, alb, nrm = _paths(hpt_scenes.cornell_box(), 48, 48, depth=8, rr=3)
    assert np.isfinite(rad).all() and (rad >= 0).all()
    # the left (red) wall is redder than the right (green) wall
    left, right = rad[20:28, 2:6].mean(axis=(0, 1)), rad[20:28, 42:46].mean(axis=(0, 1))
    assert left[0] > left[1] and right[1] > right[0]


def test_accumulation_is_the_running_mean():
    sc = hpt_scenes.cornell_box()
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(16, 16, 4, 2, tonemap=False)
    frames = [o.trace_paths(st, f)[0] for f in range(4)]
    img = o.render(st, 4)["accum"][..., :3]
    assert np.allclose(img, np.mean(frames, axis=0), rtol=1e-5, atol=1e-6)
