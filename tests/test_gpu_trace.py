"""GPU ray-query parity: the HIP BVH walk against the oracle's brute force.

Bar: closest-hit kind / primitive / triangle ids bit-exact, t, u, v
bit-exact (both sides evaluate the same watertight test in the same IEEE
order), any-hit flags identical, and the per-ray node/triangle visit counts
identical to the oracle's restatement of the 8-wide walk.
"""
import numpy as np
import pytest

import hpt
import hpt_scenes
import pyoracle

pytestmark = pytest.mark.gpu


def _renderer(scene, w=64, h=64, **kw):
    print(f"[gpu] scene with {scene.triangle_count} triangles, {w}x{h}", flush=True)
    r = hpt.HalaRenderer(w, h, max_depth=kw.pop("max_depth", 8), rr_depth=kw.pop("rr_depth", 3), **kw)
    r.set_scene(scene)
    r.commit()
    return r


def _same_hits(a, b):
    return (a["kind"] == b["kind"]) & (a["instance"] == b["instance"]) & (a["primitive"] == b["primitive"])


def _random_rays(n, seed, lo=-0.5, hi=1.5, target_lo=0.0, target_hi=1.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    t = rng.uniform(target_lo, target_hi, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return hpt.make_rays(o, d)


def test_cornell_primary_hits_bitexact():
    scene = hpt_scenes.cornell_box()
    w = h = 128
    r = _renderer(scene, w, h)
    o = pyoracle.OracleScene(scene)
    st = pyoracle.settings(w, h, 8, 3)
    ys, xs = np.mgrid[0:h, 0:w]
    for frame in (0, 1, 7):
        rays = o.camera_rays(st, frame, xs.ravel(), ys.ravel())
        gpu = r.trace_closest(rays)
        ref = o.trace_closest(rays, brute=True)
        assert _same_hits(gpu, ref).all()
        assert np.array_equal(gpu["t"], ref["t"])
        assert np.array_equal(gpu["u"], ref["u"]) and np.array_equal(gpu["v"], ref["v"])
        assert (ref["kind"] == hpt.HIT_TRIANGLE).sum() > 0.9 * len(rays)
        assert (ref["kind"] == hpt.HIT_LIGHT).sum() > 0
    r.close()


BENCH_SCENES = {
    "disney_grid_env_1080p": (lambda: hpt_scenes.disney_grid(seed=0), 10, 0),
    "icosphere_lights_medium_1080p": (lambda: hpt_scenes.displaced_icosphere_lights_medium(level=8), 12,
                                      hpt_scenes.MEDIUM_SCENE_FEATURES),
}


@pytest.mark.parametrize("name", sorted(BENCH_SCENES))
def test_benchmark_camera_rays_primary_hits(name):
    """The benchmark scenes' own 1920x1080 cameras (a pinhole over 194 k
    triangles; a thin lens of aperture 0.05 over 1.31 M): primary rays of
    every 13th pixel in two frames, generated by the oracle's camera
    restatement.  Closest-hit ids, t, u and v are bit-exact against the
    oracle's own BVH walk on the identical rays; hit distances lie within
    1e-5 relative of a double-precision brute force on a subset, away from
    near-ties (second-closest triangle within 1e-4 relative)."""
    make, depth, features = BENCH_SCENES[name]
    scene = make()
    r = _renderer(scene, 1920, 1080, max_depth=depth, rr_depth=3, features=features)
    o = pyoracle.OracleScene(scene)
    st = pyoracle.settings(1920, 1080, depth, 3, features)
    ys, xs = np.mgrid[0:1080:13, 0:1920:13]
    for frame in (0, 3):
        rays = o.camera_rays(st, frame, xs.ravel(), ys.ravel())
        gpu = r.trace_closest(rays)
        ref = o.trace_closest(rays)
        assert _same_hits(gpu, ref).all()
        tri = ref["kind"] == hpt.HIT_TRIANGLE
        light = ref["kind"] == hpt.HIT_LIGHT
        bad = gpu["t"] != ref["t"]
        print(f"[gpu] {name} frame {frame}: {tri.sum()} triangle / {light.sum()} light hits; t differs on "
              f"{(bad & tri).sum()} triangle and {(bad & light).sum()} light hits", flush=True)
        # Triangles: the same watertight test in the same IEEE order on both sides.
        assert np.array_equal(gpu["t"][tri], ref["t"][tri])
        assert np.array_equal(gpu["u"][tri], ref["u"][tri]) and np.array_equal(gpu["v"][tri], ref["v"][tri])
        # Analytic lights (lights.rint.glsl), also evaluated without fused multiply-adds.
        if light.any():
            assert (np.abs(gpu["t"][light] - ref["t"][light]) / ref["t"][light]).max() < 1e-6
        assert tri.sum() > 0.3 * len(rays)
    n64 = max(100, int(4e7 / scene.triangle_count))  # a few seconds of double-precision brute force
    sub = rays[:: max(1, len(rays) // n64)]
    hits = r.trace_closest(sub)
    t64, tri64, gap = o.trace_fp64(sub)
    ok = (hits["kind"] == hpt.HIT_TRIANGLE) & np.isfinite(t64) & (gap > 1e-4 * t64)
    assert ok.sum() > 40
    assert (np.abs(hits["t"][ok] - t64[ok]) / t64[ok]).max() < 1e-5
    print(f"[gpu] {name}: {len(rays)} rays per frame bit-exact; {ok.sum()} checked against fp64", flush=True)
    o.close()
    r.close()


@pytest.mark.parametrize("ntri", [1, 7, 1000, 100_000])
def test_soup_closest_bitexact(ntri):
    scene = hpt_scenes.triangle_soup(ntri, seed=ntri)
    r = _renderer(scene)
    o = pyoracle.OracleScene(scene)
    rays = _random_rays(20000, seed=1)
    gpu = r.trace_closest(rays)
    ref = o.trace_closest(rays, brute=ntri <= 1000)
    assert _same_hits(gpu, ref).all(), int((~_same_hits(gpu, ref)).sum())
    assert np.array_equal(gpu["t"], ref["t"])
    if ntri >= 1000:
        assert (ref["kind"] == hpt.HIT_TRIANGLE).sum() > 1000
    r.close()


def test_soup_any_hit_matches():
    scene = hpt_scenes.triangle_soup(50_000, seed=3)
    r = _renderer(scene)
    o = pyoracle.OracleScene(scene)
    rays = _random_rays(20000, seed=2)
    rng = np.random.default_rng(5)
    rays["tmax"] = rng.uniform(0.01, 2.0, len(rays)).astype(np.float32)
    assert np.array_equal(r.trace_any(rays), o.trace_any(rays))
    r.close()


def test_visit_counts_match_oracle_walk():
    scene = hpt_scenes.triangle_soup(30_000, seed=4)
    r = _renderer(scene)
    o = pyoracle.OracleScene(scene)
    nodes, tris = r.export_bvh()
    rays = _random_rays(5000, seed=6)
    for any_hit in (False, True):
        gv, gt = r.trace_counts(rays, any_hit=any_hit)
        hits, ov, ot = o.trace_bvh8(nodes, tris, rays, any_hit=any_hit)
        assert np.array_equal(gv, ov)
        assert np.array_equal(gt, ot)
        assert gv.mean() > 1.0
    # the oracle's walk of the exported tree finds the brute-force closest hits
    hits, _, _ = o.trace_bvh8(nodes, tris, rays)
    assert _same_hits(hits, o.trace_closest(rays)).all()
    r.close()


def test_edge_and_vertex_rays_are_watertight():
    # Rays aimed exactly at shared edges and vertices of a closed cube from
    # inside must never escape.
    scene = hpt_scenes.furnace_sphere(segments=16, rings=8)
    r = _renderer(scene)
    tri = scene.world_triangles()
    targets = np.concatenate([tri.reshape(-1, 3), 0.5 * (tri[:, 0] + tri[:, 1]), 0.5 * (tri[:, 1] + tri[:, 2])])
    o = np.zeros_like(targets) + np.array([0.01, -0.02, 0.03], np.float32)
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    hits = r.trace_closest(hpt.make_rays(o, d))
    assert (hits["kind"] == hpt.HIT_TRIANGLE).all()
    r.close()


def test_hit_distance_vs_fp64():
    scene = hpt_scenes.triangle_soup(2000, seed=9)
    r = _renderer(scene)
    o = pyoracle.OracleScene(scene)
    rays = _random_rays(4000, seed=10)
    gpu = r.trace_closest(rays)
    t64, tri64, gap = o.trace_fp64(rays)
    hit = (gpu["kind"] == hpt.HIT_TRIANGLE) & np.isfinite(t64) & (gap > 1e-4)
    assert hit.sum() > 200
    rel = np.abs(gpu["t"][hit] - t64[hit]) / t64[hit]
    assert rel.max() < 1e-5
    r.close()


def test_lights_in_closest_and_any():
    scene = hpt_scenes.cornell_box()
    r = _renderer(scene)
    o = pyoracle.OracleScene(scene)
    rng = np.random.default_rng(12)
    org = rng.uniform([-0.9, -0.9, 0.1], [0.9, 0.9, 1.9], (4000, 3))
    tgt = rng.uniform([-0.125, -0.125, 1.999], [0.125, 0.125, 1.999], (4000, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = hpt.make_rays(org, d)
    gpu, ref = r.trace_closest(rays), o.trace_closest(rays, brute=True)
    assert _same_hits(gpu, ref).all()
    assert np.array_equal(gpu["t"], ref["t"])
    assert (ref["kind"] == hpt.HIT_LIGHT).sum() > 100
    assert np.array_equal(r.trace_any(rays), o.trace_any(rays, brute=True))
    r.close()


def test_deep_tree_spill_bitexact():
    # Nested triangles shrinking by 0.7 give a 21-level tree: walks push more
    # entries than the LDS short stack holds and use the global spill.
    scene = hpt_scenes.nested_triangles(200, 0.7)
    r = _renderer(scene)
    assert r.bvh_info()["max_depth"] > 12
    o = pyoracle.OracleScene(scene)
    rng = np.random.default_rng(9)
    n = 20000
    tri = scene.world_triangles().astype(np.float64)
    pick = rng.integers(0, len(tri), n)
    b = rng.dirichlet((1, 1, 1), n)
    target = np.einsum("nk,nkd->nd", b, tri[pick])
    org = rng.uniform(-0.5, 1.5, (n, 3))
    d = target - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = hpt.make_rays(org, d)
    gpu = r.trace_closest(rays)
    ref = o.trace_closest(rays, brute=True)
    assert _same_hits(gpu, ref).all()
    assert np.array_equal(gpu["t"], ref["t"])
    assert (ref["kind"] == hpt.HIT_TRIANGLE).sum() > 1000
    rays["tmax"] = rng.uniform(0.01, 2.0, n).astype(np.float32)
    assert np.array_equal(r.trace_any(rays), o.trace_any(rays, brute=True))
    nodes, tris = r.export_bvh()
    for any_hit in (False, True):
        gv, gt = r.trace_counts(rays, any_hit=any_hit)
        _, ov, ot = o.trace_bvh8(nodes, tris, rays, any_hit=any_hit)
        assert np.array_equal(gv, ov) and np.array_equal(gt, ot)
        assert gv.max() > 12
    r.close()
