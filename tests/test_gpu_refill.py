"""Lane refill in the traversal kernels (hpt_set_traversal_refill): resident
blocks whose waves take new queue entries for the lanes whose rays have
ended, instead of one queue entry per lane.  Every entry is traced by the
same operations in the same order, so the images must equal the
one-entry-per-lane launches bit for bit, and the node visits, triangle tests
and queries counted over a batch must be the same totals.

Cases: a 640x360 wavefront of 4 frames (more entries per queue segment than
the resident lanes of its blocks: waves refill many times) and a 96x64 one
(segments drain early and blocks move on to the other segments of their
XCD); the medium scene's transmittance walks through a transparent and a
glass boundary, absorbing and scattering media, an environment map (two
walks per scattering event); the Disney grid (extension and shadow rays); eager
and graph launches, sorted shading, the three traversal block sizes, the
tail kernel from bounce 2; group steps of several jobs."""
import numpy as np
import pytest

import hpt
import hpt_scenes

pytestmark = pytest.mark.gpu

IMAGES = (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_NORMAL, hpt.IMAGE_FINAL_RGBA8)
MEDIUM = hpt.FEATURE_MEDIUM | hpt.FEATURE_VOL_MIS | hpt.FEATURE_TRANSPARENT


def _scene(case, w, h):
    if case == "disney":
        return hpt_scenes.disney_grid(seed=2, width=w, height=h, env_width=64, env_height=32, segments=16,
                                      rings=8), 0
    sc = hpt_scenes.displaced_icosphere_lights_medium(level=5, width=w, height=h,
                                                      boundary="glass" if "glass" in case else "transparent")
    if "absorb" in case:
        sc.materials[2]["medium_type"] = hpt.MEDIUM_ABSORB
        sc.materials[2]["medium_color"] = (0.2, 0.5, 0.8)
    if "env" in case:
        sc.env_rgb = hpt_scenes.sky_env_map(64, 32, seed=2, sun_radiance=200.0, sun_diameter_deg=8.0)
    return sc, MEDIUM


def _render(scene, w, h, features, refill, mode=hpt.LAUNCH_EAGER, sort=hpt.SHADE_SORT_OFF, block=128,
            count=False, tail=hpt.TAIL_FROM_NONE):
    r = hpt.HalaRenderer(w, h, max_depth=8, rr_depth=3, features=features)
    r.set_scene(scene)
    r.frames_in_flight = 4
    r.commit()
    r.set_launch_mode(mode)
    r.set_shade_sort(sort)
    r.set_traversal_block(block)
    r.set_traversal_refill(refill)
    r.set_tail_from(tail)
    if count:
        r.set_profiling(hpt.PROFILE_TIMING | hpt.PROFILE_COUNT_VISITS)
    r.render(4)
    out = [r.read_image(im) for im in IMAGES]
    stats = None
    if count:
        stats = {s: r.stage_stats(s) for s in (hpt.STAGE_EXTEND, hpt.STAGE_WALK, hpt.STAGE_SHADOW, hpt.STAGE_TAIL)}
    r.close()
    return out, stats


@pytest.mark.parametrize("size", [(640, 360), (96, 64)], ids=["640x360", "96x64"])
@pytest.mark.parametrize("case", ["medium_transparent", "medium_glass_absorb", "medium_env", "disney"])
def test_refill_gives_the_same_images_and_counts(case, size):
    w, h = size
    scene, features = _scene(case, w, h)
    print(f"[gpu] refill {case} {w}x{h}", flush=True)
    base, base_stats = _render(scene, w, h, features, False, count=True)
    got, got_stats = _render(scene, w, h, features, True, count=True)
    for k, (a, b) in enumerate(zip(base, got)):
        assert np.array_equal(a, b), (case, k)
    for stage in base_stats:
        for key in ("rays", "node_visits", "tri_tests", "queries"):
            assert base_stats[stage][key] == got_stats[stage][key], (stage, key, base_stats[stage], got_stats[stage])
    assert base_stats[hpt.STAGE_EXTEND]["node_visits"] > 0
    if features:
        assert base_stats[hpt.STAGE_WALK]["queries"] > 0
    else:
        assert base_stats[hpt.STAGE_SHADOW]["node_visits"] > 0
    # the plain kernels (no counting) under graph replay, sorted shading and
    # each block size
    plain, _ = _render(scene, w, h, features, False)
    for mode, sort, block in ((hpt.LAUNCH_GRAPH, hpt.SHADE_SORT_OFF, 128), (hpt.LAUNCH_GRAPH, 1, 64),
                              (hpt.LAUNCH_EAGER, 1, 256)):
        out, _ = _render(scene, w, h, features, True, mode, sort, block)
        for k, (a, b) in enumerate(zip(plain, out)):
            assert np.array_equal(a, b), (case, mode, sort, block, k)
    # the tail kernel from bounce 2 (k_tail against k_tail_refill), whose
    # counts must match too
    t_plain, t_stats = _render(scene, w, h, features, False, count=True, tail=2)
    t_got, t_got_stats = _render(scene, w, h, features, True, count=True, tail=2)
    for k, (a, b) in enumerate(zip(plain, t_plain)):
        assert np.array_equal(a, b), (case, "tail", k)
    for k, (a, b) in enumerate(zip(t_plain, t_got)):
        assert np.array_equal(a, b), (case, "tail refill", k)
    for stage in t_stats:
        for key in ("rays", "node_visits", "tri_tests", "queries"):
            assert t_stats[stage][key] == t_got_stats[stage][key], (stage, key, t_stats[stage], t_got_stats[stage])
    assert t_stats[hpt.STAGE_TAIL]["queries"] > 0


def test_refill_in_group_steps():
    scene, features = _scene("medium_transparent", 96, 64)
    t = hpt.HalaRenderer(8, 8, max_depth=1, rr_depth=0, features=features)
    t.frames_in_flight = 1
    t.set_scene(scene)
    t.commit()
    specs = [(160, 96, 8, 3), (96, 64, 6, 2), (64, 40, 10, 0)]

    def run(refill):
        jobs = []
        for w, h, depth, rr in specs:
            r = hpt.HalaRenderer(w, h, max_depth=depth, rr_depth=rr, features=features)
            r.frames_in_flight = 2
            r.share_scene(t)
            jobs.append(r)
        g = hpt.HalaGroup()
        g.set_launch_mode(hpt.LAUNCH_GRAPH)
        g.set_traversal_refill(refill)
        for _ in range(3):
            g.render(jobs, [2, 2, 2])
        out = [[r.read_image(im) for im in IMAGES] for r in jobs]
        g.close()
        for r in jobs:
            r.close()
        return out

    print("[gpu] refill group steps", flush=True)
    base, got = run(False), run(True)
    for j, (a_imgs, b_imgs) in enumerate(zip(base, got)):
        for k, (a, b) in enumerate(zip(a_imgs, b_imgs)):
            assert np.array_equal(a, b), (j, k)
    t.close()
