"""The tail of a TRANSPARENT wavefront: paths that outlive the max_depth + 1
bounces (a pass-through of a transparent surface does not count toward depth,
pathtracer.rgen.glsl:196-200) run to their end in one kernel, each path on one
lane, so a batch needs no host round trip and replays as one graph.

Per-sample against the oracle (whose loop has no bound) on a scene where paths
pass through up to 8 layers after a floor bounce, at max_depth 1 and 3, in the
three TRANSPARENT permutations; eager equals graph replays bit for bit; the
tail ran (stage counters) and cut no path."""
import numpy as np
import pytest

import hpt
import hpt_scenes
import pyoracle

pytestmark = pytest.mark.gpu

PER_SAMPLE_MIN = 0.995  # observed 1.0000 (the oracle and the GPU draw the same streams)
FEATURES = [hpt.FEATURE_TRANSPARENT, hpt.FEATURE_MEDIUM | hpt.FEATURE_TRANSPARENT,
            hpt.FEATURE_MEDIUM | hpt.FEATURE_VOL_MIS | hpt.FEATURE_TRANSPARENT]


def _renderer(scene, w, h, max_depth, rr_depth, features, mode=hpt.LAUNCH_GRAPH, frames_in_flight=0):
    print(f"[gpu] tail render {w}x{h} depth {max_depth} features {features} mode {mode}", flush=True)
    r = hpt.HalaRenderer(w, h, max_depth=max_depth, rr_depth=rr_depth, features=features)
    r.set_launch_mode(mode)
    if frames_in_flight:
        r.frames_in_flight = frames_in_flight
    r.set_scene(scene)
    r.commit()
    return r


def _agreement(gpu, ref, tol=1e-3):
    err = np.abs(gpu - ref) / (np.abs(ref) + 1e-2)
    return float((err.max(axis=-1) < tol).mean())


@pytest.mark.parametrize("max_depth", [1, 3])
@pytest.mark.parametrize("features", FEATURES)
def test_tail_per_sample_against_the_oracle(features, max_depth):
    sc = hpt_scenes.transparent_layers()
    w, h, frames = 64, 40, 4
    r = _renderer(sc, w, h, max_depth, 0, features, frames_in_flight=frames)
    r.render(frames)  # one graph-captured batch of 4 frames
    gpu = r.read_image(hpt.IMAGE_ACCUM)[..., :3]
    assert r.launch_stats()[1] == 1  # captured as a graph
    assert r.truncated_paths() == 0
    r.close()
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(w, h, max_depth, 0, features)
    ref = np.mean([o.trace_paths(st, f)[0] for f in range(frames)], axis=0)
    agree = _agreement(gpu, ref)
    # profiled (eager) run: the tail took paths and traced their bounces
    p = _renderer(sc, w, h, max_depth, 0, features, frames_in_flight=frames)
    p.set_profiling(hpt.PROFILE_TIMING)
    p.render(frames)
    tail = p.stage_stats(hpt.STAGE_TAIL)
    same = np.array_equal(p.read_image(hpt.IMAGE_ACCUM)[..., :3], gpu)
    p.close()
    print(f"[gpu] features {features} depth {max_depth}: per-sample {agree:.4f}, tail paths {tail['rays']}, "
          f"tail bounces {tail['queries']}, profiled run equal {same}", flush=True)
    assert agree >= PER_SAMPLE_MIN
    assert tail["launches"] == 1 and tail["rays"] > 0 and tail["queries"] >= tail["rays"]
    assert same


@pytest.mark.parametrize("case", ["layers", "icosphere_medium"])
def test_tail_graph_replays_equal_eager(case):
    if case == "layers":
        sc, features, w, h, depth = hpt_scenes.transparent_layers(), FEATURES[2], 64, 40, 2
    else:
        sc = hpt_scenes.displaced_icosphere_lights_medium(level=3, width=96, height=54)
        features, w, h, depth = hpt_scenes.MEDIUM_SCENE_FEATURES, 96, 54, 8
    g = _renderer(sc, w, h, depth, min(3, depth), features, mode=hpt.LAUNCH_GRAPH, frames_in_flight=2)
    g.render(6)  # three batches: one capture, two replays
    assert g.launch_stats() == (3, 1)
    e = _renderer(sc, w, h, depth, min(3, depth), features, mode=hpt.LAUNCH_EAGER, frames_in_flight=1)
    for _ in range(6):
        e.render(1)
    for im in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_NORMAL, hpt.IMAGE_FINAL_RGBA8):
        assert np.array_equal(g.read_image(im), e.read_image(im)), im
    assert g.truncated_paths() == 0 and e.truncated_paths() == 0
    g.close()
    e.close()


def test_tail_converges_to_the_oracle():
    """Independent frames: GPU frames [0, 48) against oracle frames [48, 96)."""
    sc = hpt_scenes.transparent_layers()
    w, h, n, depth, f = 48, 30, 48, 3, FEATURES[2]
    r = _renderer(sc, w, h, depth, 2, f)
    r.render(n)
    gpu = r.read_image(hpt.IMAGE_ACCUM)[..., :3].astype(np.float64)
    r.close()
    o = pyoracle.OracleScene(sc)
    st = pyoracle.settings(w, h, depth, 2, f)
    frames = np.array([o.trace_paths(st, k)[0] for k in range(n, 2 * n)], np.float64)
    mean, var = frames.mean(axis=0), frames.var(axis=0, ddof=1)
    se = np.sqrt(2.0 * var.sum() / n) / gpu.size  # the difference of two independent n-frame means
    diff = gpu.mean() - mean.mean()
    print(f"[gpu] tail converged: image mean {gpu.mean():.5f} vs oracle {mean.mean():.5f} (se {se:.5f})", flush=True)
    assert abs(diff) < 4.0 * se + 1e-6


@pytest.mark.parametrize("case", ["disney", "icosphere_medium"])
def test_earlier_tail_gives_the_same_images(case):
    """hpt_set_tail_from: the paths live at bounce 1, 2 or 4 run to their end
    in the tail kernel -- the same images as the wavefront, in eager and
    graph launches; the tail took paths."""
    if case == "disney":
        sc = hpt_scenes.disney_grid(seed=3, width=96, height=64, env_width=64, env_height=32, segments=16, rings=8)
        features, w, h = 0, 96, 64
    else:
        sc = hpt_scenes.displaced_icosphere_lights_medium(level=3, width=96, height=54)
        features, w, h = hpt_scenes.MEDIUM_SCENE_FEATURES, 96, 54
    base = _renderer(sc, w, h, 8, 3, features, frames_in_flight=2)
    base.render(4)
    want = [base.read_image(im) for im in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_FINAL_RGBA8)]
    base.close()
    for tf in (1, 2, 4):
        for mode in (hpt.LAUNCH_EAGER, hpt.LAUNCH_GRAPH):
            r = _renderer(sc, w, h, 8, 3, features, mode=mode, frames_in_flight=2)
            r.set_tail_from(tf)
            r.render(4)
            got = [r.read_image(im) for im in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_FINAL_RGBA8)]
            r.set_profiling(hpt.PROFILE_TIMING)
            r.render(2)
            tail = r.stage_stats(hpt.STAGE_TAIL)
            r.close()
            print(f"[gpu] {case} tail from {tf} mode {mode}: tail paths {tail['rays']}", flush=True)
            for a, b in zip(want, got):
                assert np.array_equal(a, b), (case, tf, mode)
            assert tail["launches"] == 1 and tail["rays"] > 0
    with pytest.raises(hpt.HptError):
        r = _renderer(sc, w, h, 8, 3, features)
        try:
            r.set_tail_from(0)
        finally:
            r.close()


def test_group_with_an_earlier_tail_equals_solo():
    sc = hpt_scenes.displaced_icosphere_lights_medium(level=3, width=96, height=54)
    features = hpt_scenes.MEDIUM_SCENE_FEATURES
    t = _renderer(sc, 8, 8, 1, 0, features)
    specs = [(96, 54, 8, 3), (64, 40, 6, 2), (48, 32, 12, 3)]

    def job(spec):
        w, h, d, rr = spec
        r = hpt.HalaRenderer(w, h, max_depth=d, rr_depth=rr, features=features)
        r.frames_in_flight = 2
        r.share_scene(t)
        return r

    jobs = [job(s) for s in specs]
    g = hpt.HalaGroup()
    g.set_tail_from(3)
    for _ in range(3):  # the second and third steps replay the captured graph
        g.render(jobs, [2, 2, 2])
    assert g.stats()["graph_launches"] == 2
    for r, s in zip(jobs, specs):
        solo = job(s)
        solo.render(6)
        assert np.array_equal(r.read_image(hpt.IMAGE_ACCUM), solo.read_image(hpt.IMAGE_ACCUM))
        solo.close()
        r.close()
    assert g.truncated_paths() == 0
    g.close()
    t.close()


@pytest.mark.parametrize("case", ["disney", "icosphere_medium"])
def test_traversal_block_sizes_give_the_same_images(case):
    """hpt_set_traversal_block: the extension, shadow and walk kernels at 64,
    128 and 256 threads per block give the same images, eager and graph."""
    if case == "disney":
        sc = hpt_scenes.disney_grid(seed=4, width=96, height=64, env_width=64, env_height=32, segments=16, rings=8)
        features, w, h = 0, 96, 64
    else:
        sc = hpt_scenes.displaced_icosphere_lights_medium(level=3, width=96, height=54)
        features, w, h = hpt_scenes.MEDIUM_SCENE_FEATURES, 96, 54
    want = None
    for bs in hpt.TRAVERSAL_BLOCKS:
        for mode in (hpt.LAUNCH_EAGER, hpt.LAUNCH_GRAPH):
            r = _renderer(sc, w, h, 8, 3, features, mode=mode, frames_in_flight=2)
            r.set_traversal_block(bs)
            r.render(4)
            got = [r.read_image(im) for im in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_FINAL_RGBA8)]
            r.close()
            if want is None:
                want = got
            for a, b in zip(want, got):
                assert np.array_equal(a, b), (case, bs, mode)
    r = _renderer(sc, w, h, 8, 3, features)
    with pytest.raises(hpt.HptError):
        r.set_traversal_block(96)
    r.close()
