"""Several jobs in one wavefront (hpt_group_render): each step traces the
frames of every job in the step with one extension, shading and shadow
launch per bounce.  Every job's images and frame counter must equal its solo
hpt_render of the same frames bit for bit -- across resolutions, depths,
Russian-roulette depths and frame counts that differ per job, jobs joining
and leaving between steps, graph replays of repeated steps, sorted shading,
and the medium/transparent permutations.  Jobs of different scenes or a job
listed twice are refused."""
import numpy as np
import pytest

import hpt
import hpt_scenes

pytestmark = pytest.mark.gpu

IMAGES = (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_NORMAL, hpt.IMAGE_FINAL_RGBA8)
# (width, height, max_depth, rr_depth) per job
SPECS = [(96, 64, 8, 3), (64, 64, 6, 2), (40, 24, 10, 0), (80, 48, 4, 3)]
# frames per job and step: job 3 joins at step 2, job 1 leaves after it; steps repeat (graph replays)
PLAN = [{0: 4, 1: 4, 2: 2}, {0: 4, 1: 4, 2: 2}, {0: 4, 1: 4, 2: 2, 3: 3}, {0: 4, 2: 2, 3: 3},
        {0: 4, 2: 2, 3: 3}, {0: 4, 2: 2, 3: 3}]


def _scene(case):
    if case == "disney":
        return hpt_scenes.disney_grid(seed=2, width=96, height=64, env_width=64, env_height=32, segments=16,
                                      rings=8), 0
    return (hpt_scenes.displaced_icosphere_lights_medium(level=3, width=96, height=64),
            hpt.FEATURE_MEDIUM | hpt.FEATURE_VOL_MIS | hpt.FEATURE_TRANSPARENT)


def _template(scene, features):
    t = hpt.HalaRenderer(8, 8, max_depth=1, rr_depth=0, features=features)
    t.frames_in_flight = 1
    t.set_scene(scene)
    t.commit()
    return t


def _job(t, spec, features):
    w, h, depth, rr = spec
    r = hpt.HalaRenderer(w, h, max_depth=depth, rr_depth=rr, features=features, enable_tonemap=True)
    r.frames_in_flight = 2
    r.share_scene(t)
    return r


@pytest.mark.parametrize("mode,sort", [(hpt.LAUNCH_EAGER, hpt.SHADE_SORT_OFF), (hpt.LAUNCH_GRAPH, hpt.SHADE_SORT_OFF),
                                       (hpt.LAUNCH_GRAPH, 1)], ids=["eager", "graph", "graph_sorted"])
@pytest.mark.parametrize("case", ["disney", "medium_vol_mis_transparent"])
def test_group_steps_equal_solo_renders(case, mode, sort):
    scene, features = _scene(case)
    print(f"[gpu] group {case} mode {mode} sort {sort}", flush=True)
    t = _template(scene, features)
    jobs = [_job(t, spec, features) for spec in SPECS]
    g = hpt.HalaGroup()
    g.set_launch_mode(mode)
    g.set_shade_sort(sort)
    totals = [0] * len(SPECS)
    for step in PLAN:
        idx = sorted(step)
        g.render([jobs[i] for i in idx], [step[i] for i in idx])
        for i in idx:
            totals[i] += step[i]
    got = [[r.read_image(im) for im in IMAGES] for r in jobs]  # ordered after the steps by events
    stats = g.stats()
    print(f"[gpu] group stats {stats}", flush=True)
    assert stats["steps"] == len(PLAN)
    if mode == hpt.LAUNCH_GRAPH:  # TRANSPARENT too: its tail runs on the device
        assert stats["graph_launches"] == 3  # steps 1, 4 and 5 repeat their predecessor
    assert g.truncated_paths() == 0
    for i, spec in enumerate(SPECS):
        assert jobs[i].frame_index == totals[i]
        solo = _job(t, spec, features)
        solo.render(totals[i])
        for k, im in enumerate(IMAGES):
            assert np.array_equal(got[i][k], solo.read_image(im)), (i, k)
        solo.close()
    # a job keeps rendering alone after group steps, continuing its frames
    jobs[0].render(2)
    solo = _job(t, SPECS[0], features)
    solo.render(totals[0] + 2)
    assert np.array_equal(jobs[0].read_image(hpt.IMAGE_ACCUM), solo.read_image(hpt.IMAGE_ACCUM))
    solo.close()
    g.close()
    for r in jobs:
        r.close()
    t.close()


def test_group_refuses_mixed_scenes_and_duplicates():
    a = _template(hpt_scenes.cornell_box(64, 64), 0)
    b = _template(hpt_scenes.cornell_box(64, 64), 0)
    ja, jb = _job(a, (32, 32, 4, 2), 0), _job(b, (32, 32, 4, 2), 0)
    g = hpt.HalaGroup()
    with pytest.raises(hpt.HptError, match="one committed scene"):
        g.render([ja, jb], [1, 1])
    with pytest.raises(hpt.HptError, match="twice"):
        g.render([ja, ja], [1, 1])
    g.render([ja], [0])  # nothing to do
    assert ja.frame_index == 0 and g.stats()["steps"] == 0
    g.close()
    for r in (ja, jb, a, b):
        r.close()


def test_handles_of_another_kind_are_refused_on_the_device():
    """A renderer handle given to a group's entry point, or a group handle to
    a renderer's, fails with HPT_ERROR_INVALID_ARGUMENT and leaves both
    handles working (each handle struct starts with a tag of its kind)."""
    lib = hpt.load_library()
    scene, _ = _scene("disney")
    r = hpt.HalaRenderer(32, 24, max_depth=4, rr_depth=2)
    r.set_scene(scene)
    r.commit()
    g = hpt.HalaGroup()
    print("[gpu] wrong-kind handles", flush=True)
    assert lib.hpt_group_set_vmis_phases(r._h, 0) == 1 and b"invalid group" in lib.hpt_last_error()
    assert lib.hpt_set_vmis_phases(g._h, 0) == 1 and b"invalid renderer" in lib.hpt_last_error()
    assert lib.hpt_group_set_traversal_refill(r._h, 0) == 1 and b"invalid group" in lib.hpt_last_error()
    assert lib.hpt_set_traversal_refill(g._h, 0) == 1 and b"invalid renderer" in lib.hpt_last_error()
    assert lib.hpt_set_frames_in_flight(g._h, 4) == 1
    g.set_vmis_phases(True)
    r.set_vmis_phases(True)
    r.render(2)
    solo = r.read_image(hpt.IMAGE_ACCUM)
    assert np.isfinite(solo).all() and r.frame_index == 2
    g.close()
    r.close()
