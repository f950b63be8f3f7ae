"""The serving loop (hpt_server_*): admission by device-memory footprint,
continuous batching of jobs over shared resident scenes, and results that do
not depend on the mix.  The scheduling runs on the CPU as the native loop's
dry run (no device); the GPU tests run real jobs through the C ABI."""
import time

import numpy as np
import pytest

import hpt
import hpt_scenes
import serve


NS_PER_PATH = 100  # dry-run cost: a 128x96 frame takes 1.2 ms of host time


def _dry_scenes():
    return {"a": serve.DryRunScene(), "b": serve.DryRunScene(), "broken": serve.DryRunScene(fail=True)}


def test_admission_keeps_the_footprint_within_budget():
    jobs = serve.poisson_jobs(40, rate=400.0, scene_names=["a", "b"], seed=1,
                              resolutions=[(64, 64), (128, 96), (256, 128)], spp_range=(4, 40))
    budget = 3 * serve.footprint(serve.JobSpec(0, "a", 256, 128, 40, 8))
    srv = serve.Server(_dry_scenes(), budget, keep_images=True, dry_run_ns_per_path=NS_PER_PATH)
    t0 = time.perf_counter()
    res = srv.run(jobs)
    makespan = time.perf_counter() - t0
    assert len(res) == 40 and all(r.error is None for r in res)
    # at every admission instant the admitted, unfinished jobs fit the budget
    for r in res:
        live = [q for q in res if q.admitted <= r.admitted < q.finished]
        assert sum(serve.footprint(q.spec) for q in live) <= budget
    # admitted in arrival order (first come, first served)
    order = sorted(res, key=lambda q: q.admitted)
    assert [q.spec.job_id for q in order] == sorted(q.spec.job_id for q in res)
    for r in res:  # a dry run's image holds the server's job id (1, 2, ... in submission order)
        assert r.image[0, 0, 0] == r.spec.job_id + 1 and r.latency >= 0
        assert r.image.shape == (r.spec.height, r.spec.width, 4)
    assert srv.stats["peak_bytes"] <= budget and srv.stats["finished"] == 40
    s = serve.summarize(res, makespan)
    assert s["jobs"] == 40 and s["failed"] == 0 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_a_failing_job_is_isolated():
    jobs = [serve.JobSpec(i, "broken" if i == 2 else "a", 32, 32, 8, 8) for i in range(5)]
    res = serve.Server(_dry_scenes(), 1 << 40, dry_run_ns_per_path=NS_PER_PATH).run(jobs)
    bad = [r for r in res if r.error]
    assert len(res) == 5 and len(bad) == 1 and bad[0].job_id == 2 and "device lost" in bad[0].error


def test_wavefront_mode_steps_jobs_of_a_scene_together():
    """Continuous batching: jobs of one scene that overlap in time share
    wavefront steps; jobs join and leave between steps; the budget holds;
    every job gets exactly its samples."""
    jobs = serve.poisson_jobs(40, rate=400.0, scene_names=["a", "b"], seed=2,
                              resolutions=[(64, 64), (128, 96)], spp_range=(4, 64))
    srv = serve.Server(_dry_scenes(), 0, max_frames=8, keep_images=True, wavefront=True, step_paths=64 * 64 * 8,
                       dry_run_ns_per_path=NS_PER_PATH)
    budget = srv.budget = 6 * srv.job_footprint(serve.JobSpec(0, "a", 128, 96, 64, 8))
    fp = {j.job_id: srv.job_footprint(j) for j in jobs}
    t0 = time.perf_counter()
    res = srv.run(jobs)
    assert len(res) == 40 and all(r.error is None for r in res)
    for r in res:
        live = [q for q in res if q.admitted <= r.admitted < q.finished]
        assert sum(fp[q.job_id] for q in live) <= budget
        assert r.image[0, 0, 0] == r.spec.job_id + 1
    assert srv.stats["max_jobs_per_step"] > 1  # steps shared by several jobs
    # chunks of about step_paths paths: 8 frames at 64x64, 3 at 128x96
    assert srv.step_frames(serve.JobSpec(0, "a", 64, 64, 64, 8)) == 8
    assert srv.step_frames(serve.JobSpec(0, "a", 128, 96, 64, 8)) == 3
    assert srv.step_frames(serve.JobSpec(0, "a", 128, 96, 2, 8)) == 2
    assert srv.stats["wavefront_steps"] >= max(-(-j.spp // srv.step_frames(j)) for j in jobs)
    # every job's chunks sum to its samples: each step holds a job at most once
    assert srv.stats["jobs_in_steps"] == sum(-(-j.spp // srv.step_frames(j)) for j in jobs)
    assert serve.summarize(res, time.perf_counter() - t0)["failed"] == 0
    bad = serve.Server(_dry_scenes(), 1 << 40, wavefront=True, dry_run_ns_per_path=NS_PER_PATH).run(
        [serve.JobSpec(i, "broken" if i == 1 else "a", 32, 32, 8, 8) for i in range(3)])
    assert sorted(r.job_id for r in bad if r.error) == [1]


def test_job_larger_than_the_budget_is_refused():
    with pytest.raises(ValueError, match="above the budget"):
        serve.Server(_dry_scenes(), 1000, dry_run_ns_per_path=NS_PER_PATH).run([serve.JobSpec(0, "a", 64, 64, 8, 8)])
    srv = hpt.RenderServer(1000, dry_run_ns_per_path=NS_PER_PATH)
    sid = srv.add_scene(None, 0)
    with pytest.raises(hpt.HptError, match="above the budget"):
        srv.submit(srv.job_desc(sid, 64, 64, 8, 8))
    srv.close()


def test_native_footprint_equals_the_model():
    """The native planner charges what serve.footprint models: per path in
    flight 288 B (+5 B with shading sort), per pixel 52 B; the batch is the
    tuned or default frames in flight, never more than the job's samples."""
    import autotune

    tuning = autotune.TuningCache()
    tuning.put("a", 64, 64, 0, autotune.Tuning(8, hpt.LAUNCH_GRAPH, 1.0, {}, 0.1, 1))
    for max_frames in (4, 32):
        for t in (None, tuning):
            srv = serve.Server(_dry_scenes(), 1 << 40, max_frames=max_frames, tuning=t, dry_run_ns_per_path=1)
            for spec in (serve.JobSpec(0, "a", 64, 64, 100, 8), serve.JobSpec(1, "a", 64, 64, 3, 8),
                         serve.JobSpec(2, "b", 1920, 1080, 128, 12), serve.JobSpec(3, "b", 7, 5, 1, 1)):
                assert srv.job_footprint(spec) == serve.footprint(spec, max_frames, t)
                assert srv.step_frames(spec) == serve.frames_in_flight(spec, max_frames, t)


def test_native_server_results_and_validation():
    """Results by id and in completion order; an image is read once; bad
    job descriptions are refused with the reason."""
    srv = hpt.RenderServer(1 << 40, keep_images=True, dry_run_ns_per_path=NS_PER_PATH)
    a = srv.add_scene(None, 0)
    ids = [srv.submit(srv.job_desc(a, 16 + i, 8, 2, 4)) for i in range(4)]
    r = srv.wait(ids[2])
    assert r["status"] == 0 and r["frames"] == 2 and r["width"] == 18 and r["has_image"]
    assert r["submitted"] <= r["admitted"] <= r["finished"]
    srv.drain()
    got = []
    while (x := srv.next_result(0.0)) is not None:
        got.append(x["job_id"])
    assert sorted(got) == ids and srv.next_result(0.01) is None
    img = srv.read_image(ids[0], 16, 8)
    assert img.shape == (8, 16, 4) and (img == ids[0]).all()
    with pytest.raises(hpt.HptError, match="unknown or released"):
        srv.read_image(ids[0], 16, 8)
    with pytest.raises(hpt.HptError, match="16"):
        srv.read_image(ids[1], 8, 8)  # wrong size
    for bad, msg in ((dict(scene=7), "scene"), (dict(spp=0), "spp"), (dict(traversal_block=96), "traversal_block"),
                     (dict(frames_in_flight=65), "frames_in_flight"), (dict(width=0), "size")):
        d = srv.job_desc(a, 16, 8, 2, 4)
        for k, v in bad.items():
            setattr(d, k, v)
        with pytest.raises(hpt.HptError, match=msg):
            srv.submit(d)
    dev = hpt.RenderServer(1 << 40)  # a server on the device takes scenes only as committed handles
    with pytest.raises(hpt.HptError, match="null scene handle"):
        dev.add_scene(None, 0)
    dev.close()
    st = srv.stats()
    assert st["submitted"] == 4 and st["finished"] == 4 and st["failed"] == 0 and st["bytes_in_use"] == 0
    srv.close()


def test_server_close_waits_for_running_jobs():
    srv = hpt.RenderServer(1 << 40, dry_run_ns_per_path=NS_PER_PATH)
    a = srv.add_scene(None, 0)
    for i in range(6):
        srv.submit(srv.job_desc(a, 64, 64, 4, 4))
    srv.close()  # returns after the six jobs (about 10 ms of dry-run work)


@pytest.mark.gpu
@pytest.mark.parametrize("wavefront", [False, True], ids=["streams", "wavefront"])
def test_results_identical_alone_and_in_the_mix(wavefront):
    scenes = {
        "cornell": serve.ResidentScene("cornell", hpt_scenes.cornell_box(), 0),
        "grid": serve.ResidentScene("grid", hpt_scenes.disney_grid(env_width=128, env_height=64, segments=16,
                                                                   rings=8), 0),
        "medium": serve.ResidentScene("medium", hpt_scenes.displaced_icosphere_lights_medium(level=2),
                                      hpt_scenes.MEDIUM_SCENE_FEATURES),
    }
    specs = [serve.JobSpec(0, "cornell", 64, 48, 20, 8), serve.JobSpec(1, "grid", 96, 54, 12, 10),
             serve.JobSpec(2, "medium", 80, 45, 8, 12), serve.JobSpec(3, "cornell", 40, 40, 33, 9),
             serve.JobSpec(4, "grid", 64, 64, 5, 8), serve.JobSpec(5, "medium", 48, 32, 17, 10)]
    print(f"[gpu] serving 6 jobs together (wavefront {wavefront})", flush=True)
    srv = serve.Server(scenes, 1 << 34, max_frames=8, keep_images=True, wavefront=wavefront)
    mixed = {r.job_id: r for r in srv.run(specs)}
    srv.close()
    assert all(r.error is None for r in mixed.values()), [r.error for r in mixed.values()]
    for s in specs:
        print(f"[gpu] job {s.job_id} alone", flush=True)
        alone = serve.Server(scenes, 1 << 34, keep_images=True).run([s])[0]  # streams, 32-frame batches
        assert np.array_equal(alone.image, mixed[s.job_id].image)
    # one of them against a handle that owns its scene (no sharing at all)
    s = specs[1]
    r = hpt.HalaRenderer(s.width, s.height, max_depth=s.max_depth, rr_depth=s.rr_depth)
    r.set_scene(scenes["grid"].scene)
    r.commit()
    r.render(s.spp)
    assert np.array_equal(r.read_image(hpt.IMAGE_ACCUM), mixed[1].image)
    r.close()
    # the shared scene's environment cannot change under its jobs
    j = hpt.HalaRenderer(16, 16, max_depth=4, rr_depth=2)
    j.share_scene(scenes["grid"].template)
    with pytest.raises(hpt.HptError, match="shared"):
        j.set_envmap(np.ones((4, 8, 3), np.float32))
    f, sc = j.memory_usage()
    assert f > 0 and sc == scenes["grid"].scene_bytes
    j.close()
    for rs in scenes.values():
        rs.close()


def test_autotune_candidates_and_cache(tmp_path):
    import autotune

    # memory and 32-bit queue limits prune the candidates
    assert autotune.candidates_for(1920, 1080, max_frames=64) == [2, 4, 8, 16, 32, 64]
    assert autotune.candidates_for(1920, 1080, max_frames=16) == [2, 4, 8, 16]
    assert autotune.candidates_for(1920, 1080, memory_bytes=8 * 1920 * 1080 * 288) == [2, 4, 8]
    assert autotune.candidates_for(65536, 65536) == [1]
    c = autotune.TuningCache()
    c.put("cornell", 256, 256, 0, autotune.Tuning(32, hpt.LAUNCH_GRAPH, 900.0, {"32/graph": 900.0}, 0.5))
    path = str(tmp_path / "tuning.json")
    c.save(path)
    d = autotune.TuningCache.load(path)
    assert d.get("cornell", 256, 256, 0).frames_in_flight == 32
    # the server takes the tuned batch size (never above the job's samples) and its footprint
    spec = serve.JobSpec(0, "cornell", 256, 256, 100, 8)
    assert serve.frames_in_flight(spec, 16, d) == 32
    assert serve.frames_in_flight(serve.JobSpec(1, "cornell", 256, 256, 10, 8), 16, d) == 10
    assert serve.frames_in_flight(serve.JobSpec(2, "other", 256, 256, 100, 8), 16, d) == 16
    assert serve.footprint(spec, 16, d) == 256 * 256 * (32 * serve.PATH_BYTES + serve.IMAGE_BYTES)
    assert serve.shade_sort(spec, d) == hpt.SHADE_SORT_OFF
    # a tuned shading sort costs its key and sorted entry per path; old caches load without it
    c.put("disney", 64, 64, 0, autotune.Tuning(8, hpt.LAUNCH_GRAPH, 500.0, {"8/graph/s1": 500.0}, 0.5, 1))
    c.save(path)
    d = autotune.TuningCache.load(path)
    sspec = serve.JobSpec(3, "disney", 64, 64, 100, 8)
    assert serve.shade_sort(sspec, d) == 1
    assert serve.footprint(sspec, 16, d) == 64 * 64 * (8 * (serve.PATH_BYTES + serve.SORT_BYTES) + serve.IMAGE_BYTES)
    assert autotune.Tuning(**{"frames_in_flight": 4, "launch": 1, "msamples_per_s": 1.0}).shade_sort == \
        hpt.SHADE_SORT_OFF


@pytest.mark.gpu
def test_autotune_picks_a_measured_candidate():
    import autotune

    rs = serve.ResidentScene("cornell", hpt_scenes.cornell_box(), 0)
    try:
        t = autotune.tune(rs.template, 128, 128, 6, max_frames=16, samples=2e6)
    finally:
        rs.close()
    assert set(t.table) >= {"2/graph", "4/graph", "8/graph", "16/graph"}
    assert sum(1 for k in t.table if "/s" in k) == 3  # sort from bounce 0, 1, 2 at the best batch
    assert t.shade_sort in (hpt.SHADE_SORT_OFF, 0, 1, 2)
    best = max(t.table.values())
    assert t.msamples_per_s >= 0.97 * best - 1e-9
    assert t.frames_in_flight in (2, 4, 8, 16) and t.launch in (hpt.LAUNCH_GRAPH, hpt.LAUNCH_EAGER)


@pytest.mark.gpu
def test_server_denoises_finished_jobs():
    """With a denoiser the server returns each finished job's image denoised
    on the device: the same bytes as denoising a solo render of the job."""
    import denoiser

    net = denoiser.UNet.load(denoiser.weights_path())
    rs = serve.ResidentScene("cornell", hpt_scenes.cornell_box(), 0)
    spec = serve.JobSpec(0, "cornell", 64, 48, 8, 8)
    print("[gpu] served job with the denoiser", flush=True)
    for wavefront in (False, True):
        res = serve.Server({"cornell": rs}, 1 << 34, keep_images=True, denoiser=net, wavefront=wavefront).run([spec])[0]
        assert res.error is None, res.error
        r = hpt.HalaRenderer(64, 48, max_depth=8, rr_depth=3, enable_tonemap=True)
        r.share_scene(rs.template)
        r.render(8)
        want = net.denoise_renderer(r).numpy()
        denoiser.synchronize()
        r.close()
        assert res.image.shape == (48, 64, 4) and np.array_equal(res.image, want)
    rs.close()


def test_handles_of_another_kind_are_refused():
    """Every handle starts with a tag of its kind: an entry point given a
    handle of another kind fails with HPT_ERROR_INVALID_ARGUMENT instead of
    writing into a struct it does not own (here a server handle passed to a
    renderer's, a group's and a denoiser's setters)."""
    lib = hpt.load_library()
    srv = hpt.RenderServer(1 << 30, dry_run_ns_per_path=1)
    assert lib.hpt_set_vmis_phases(srv._h, 1) == 1 and b"invalid renderer" in lib.hpt_last_error()
    assert lib.hpt_group_set_vmis_phases(srv._h, 1) == 1 and b"invalid group" in lib.hpt_last_error()
    assert lib.hpt_set_traversal_refill(srv._h, 0) == 1 and b"invalid renderer" in lib.hpt_last_error()
    assert lib.hpt_group_set_traversal_refill(srv._h, 0) == 1 and b"invalid group" in lib.hpt_last_error()
    assert lib.hpt_set_frames_in_flight(srv._h, 4) == 1
    out = np.zeros(4, np.float32)
    assert lib.hpt_denoise(srv._h, srv._h, out.ctypes.data, 16, 0) == 1
    lib.hpt_destroy(srv._h)  # ignored: not a renderer
    assert srv.stats()["submitted"] == 0  # the server is intact
    srv.close()


def test_running_jobs_capped_in_stream_mode():
    """max_running bounds the jobs rendering at once (streams mode) on top of
    the memory budget; admission stays first come, first served."""
    jobs = serve.poisson_jobs(30, rate=1000.0, scene_names=["a"], seed=4, resolutions=[(64, 64)], spp_range=(4, 16))
    srv = serve.Server(_dry_scenes(), 1 << 40, dry_run_ns_per_path=NS_PER_PATH, max_running=3)
    res = srv.run(jobs)
    assert len(res) == 30 and all(r.error is None for r in res)
    for r in res:
        assert sum(1 for q in res if q.admitted <= r.admitted < q.finished) <= 3
    order = sorted(res, key=lambda q: q.admitted)
    assert [q.spec.job_id for q in order] == sorted(q.spec.job_id for q in res)


@pytest.mark.gpu
def test_kept_handles_render_later_jobs_of_their_class():
    """A finished job's handle renders the next job of its class (same scene,
    size, depths, batch and tuning) with its images cleared: every image still
    equals the job's solo render; a job of another class that needs the room
    evicts kept handles, and without reuse nothing is kept."""
    rs = serve.ResidentScene("cornell", hpt_scenes.cornell_box(), 0)
    specs = [serve.JobSpec(i, "cornell", 64, 48, 8 + 4 * (i % 2), 6, 3, 0.05 * i) for i in range(4)]
    specs.append(serve.JobSpec(4, "cornell", 96, 96, 8, 6, 3, 0.25))  # another class
    print("[gpu] kept handles: 5 jobs, 2 classes", flush=True)
    one = serve.footprint(specs[0], 32)
    # room for one 64x48 job and the kept handle of another, not for the 96x96 job beside them
    budget = max(2 * one + serve.footprint(specs[4], 32) // 2, serve.footprint(specs[4], 32))
    srv = serve.Server({"cornell": rs}, budget, keep_images=True, max_running=1)
    res = {r.job_id: r for r in srv.run(specs)}
    assert all(r.error is None for r in res.values()), [r.error for r in res.values()]
    assert srv.stats["handles_reused"] >= 2, srv.stats  # jobs 2 and 3 reuse the handles of 0 and 1
    assert srv.stats["peak_bytes"] <= budget
    for s in specs:
        r = hpt.HalaRenderer(s.width, s.height, max_depth=s.max_depth, rr_depth=s.rr_depth, enable_tonemap=True)
        r.share_scene(rs.template)
        r.render(s.spp)
        assert np.array_equal(r.read_image(hpt.IMAGE_ACCUM), res[s.job_id].image), s
        r.close()
    fresh = serve.Server({"cornell": rs}, 1 << 34, keep_images=True, reuse_handles=False)
    again = {r.job_id: r for r in fresh.run(specs[:2])}
    assert fresh.stats["handles_reused"] == 0 and fresh.stats["cached_bytes"] == 0
    assert all(np.array_equal(again[i].image, res[i].image) for i in range(2))
    srv.close()
    fresh.close()
    rs.close()
