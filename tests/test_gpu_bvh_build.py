"""The on-GPU BVH build (hpt_set_bvh_builder(HPT_BVH_GPU)) against the host
build: the exported 8-wide nodes and leaf triangles must be the same bytes
(the device build does the host build's arithmetic: binned SAH in the same
order, stable splits, greedy collapse, octant slots, quantisation), over
scenes that take only the one-block subtree path (<= 1024 triangles), the
top-level passes too, and a 1.3 M-triangle mesh; a large node whose
centroids coincide hands over to the host build; a non-finite vertex fails
on both paths."""
import numpy as np
import pytest

import hpt
import hpt_scenes

pytestmark = pytest.mark.gpu


def _build(scene, builder, features=0):
    r = hpt.HalaRenderer(32, 32, max_depth=4, rr_depth=2, features=features)
    r.set_scene(scene)
    r.set_bvh_builder(builder)
    r.commit()
    nodes, tris = r.export_bvh()
    info, stats = r.bvh_info(), r.bvh_build_stats()
    r.close()
    return nodes, tris, info, stats


SCENES = [
    ("cornell", lambda: hpt_scenes.cornell_box(), 0),
    ("nested_triangles", lambda: hpt_scenes.nested_triangles(200), 0),
    ("disney_grid_16x8", lambda: hpt_scenes.disney_grid(seed=1, width=64, height=64, env_width=16, env_height=8,
                                                          segments=16, rings=8), 0),
    ("soup_50k", lambda: hpt_scenes.triangle_soup(50_000, seed=7), 0),
    ("icosphere_l6_medium", lambda: hpt_scenes.displaced_icosphere_lights_medium(level=6),
     hpt_scenes.MEDIUM_SCENE_FEATURES),
]


@pytest.mark.parametrize("name,make,features", SCENES, ids=[s[0] for s in SCENES])
def test_gpu_build_equals_host_build(name, make, features):
    scene = make()
    print(f"[gpu] {name}: {scene.triangle_count} triangles", flush=True)
    hn, ht, hinfo, hstats = _build(scene, hpt.BVH_HOST, features)
    gn, gt, ginfo, gstats = _build(scene, hpt.BVH_GPU, features)
    assert hstats[0] == hpt.BVH_HOST and gstats[0] == hpt.BVH_GPU
    print(f"  host {hstats[1]:.1f} ms, gpu {gstats[1]:.1f} ms, {hinfo}", flush=True)
    assert hinfo == ginfo
    assert np.array_equal(hn, gn), f"{int((hn != gn).sum())} node bytes differ"
    assert np.array_equal(ht, gt), f"{int((ht != gt).sum())} triangle bytes differ"


def test_gpu_build_of_a_large_mesh_equals_host_build():
    scene = hpt_scenes.displaced_icosphere_lights_medium(level=8)
    hn, ht, hinfo, hstats = _build(scene, hpt.BVH_HOST, hpt_scenes.MEDIUM_SCENE_FEATURES)
    gn, gt, ginfo, gstats = _build(scene, hpt.BVH_GPU, hpt_scenes.MEDIUM_SCENE_FEATURES)
    print(f"[gpu] {scene.triangle_count} triangles: host {hstats[1]:.0f} ms, gpu {gstats[1]:.1f} ms", flush=True)
    assert gstats[0] == hpt.BVH_GPU
    assert hinfo == ginfo and np.array_equal(hn, gn) and np.array_equal(ht, gt)


def test_coincident_centroids_fall_back_to_the_host_build():
    """2000 copies of one triangle: no binned split exists, so the device
    build hands the scene to the host build (which halves the range)."""
    v = np.zeros((3, hpt.VERTEX_FLOATS), np.float32)
    v[:, 0:3] = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    v[:, 3:6] = [0, 0, 1]
    sc = hpt_scenes.cornell_box()
    sc.primitives = [hpt.Primitive(v, np.array([[0, 1, 2]] * 2000, np.uint32), 0)]
    hn, ht, hinfo, _ = _build(sc, hpt.BVH_HOST)
    gn, gt, ginfo, gstats = _build(sc, hpt.BVH_GPU)
    assert gstats[0] == hpt.BVH_HOST
    assert hinfo == ginfo and np.array_equal(hn, gn) and np.array_equal(ht, gt)


def test_non_finite_vertex_is_rejected_by_both_builders():
    sc = hpt_scenes.cornell_box()
    bad = sc.primitives[0].vertices.copy()
    bad[1, 0] = np.nan
    sc.primitives[0] = hpt.Primitive(bad, sc.primitives[0].indices, 0)
    for builder in (hpt.BVH_HOST, hpt.BVH_GPU):
        r = hpt.HalaRenderer(16, 16, max_depth=2, rr_depth=1)
        r.set_scene(sc)
        r.set_bvh_builder(builder)
        with pytest.raises(hpt.HptError, match="not finite"):
            r.commit()
        r.close()
