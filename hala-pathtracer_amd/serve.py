"""Render serving on one GPU: progressive jobs of mixed scene, resolution,
samples and depth, admitted by device-memory footprint and run concurrently.

Design (no counterpart in the reference, which renders one scene per
process):

* Resident scenes.  Each scene is committed once into a small template
  handle (BVH build + upload); every job over it is a handle that shares that
  device copy (`hpt_share_scene`), so admitting a job costs only its own path
  state and images.
* Admission.  A job's footprint is its frame state: k * W * H paths of
  288 B (path state, queues, shadow rays) plus 52 B of images per pixel,
  with k frames in flight.  Jobs are admitted in arrival order while the sum
  of admitted footprints fits the budget; a job that does not fit waits (and
  blocks the ones behind it, so large jobs are not starved).
* Execution.  An admitted job runs on its own HIP stream in chunks of k
  frames from a worker thread; the device interleaves the kernels of the
  active jobs (the hardware queues run streams concurrently).  When its last
  chunk ends, the job's image is read back (the latency end point), its
  handle is closed and its memory returns to the budget.
* Wavefront mode (`wavefront=True`).  The jobs of one resident scene share
  one wavefront (`hpt_group_render`): a worker per scene renders, step after
  step, the next chunk of every active job of that scene together -- one
  camera and accumulation launch per job, one extension, shading and shadow
  launch per bounce for all of them -- so their bounces interleave in one
  queue; a job that finishes leaves after its step and jobs admitted
  meanwhile join the next one.  A job's chunk is sized to about `step_paths`
  paths (frames = step_paths / pixels, at least 1), so a step costs every job
  about the same and small jobs are not held back by large ones.  Up to two
  steps are queued per scene: the next step is queued before the jobs that
  ended in the previous one are read back.  Workers of different scenes
  overlap on their own streams.  A step that repeats the previous one (same
  jobs and chunk sizes) replays a captured HIP graph.
* Results do not depend on the mix: a job's image is a function of its own
  scene, size, depths and frame range (the RNG is seeded by pixel and frame),
  so it is the same alone and among other jobs, bit for bit.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

import hpt

PATH_BYTES = 288     # per path in flight: 11 float4 + float2 path fields, 2 queue slots, 2 shadow rays
IMAGE_BYTES = 52     # per pixel: three rgba32f images + the rgba8 final image
SORT_BYTES = 5       # per path while shading sort is on: a key byte and a sorted queue entry


@dataclass
class JobSpec:
    job_id: int
    scene: str
    width: int
    height: int
    spp: int
    max_depth: int
    rr_depth: int = 3
    arrival: float = 0.0  # seconds after the start of the run


@dataclass
class JobResult:
    job_id: int
    spec: JobSpec
    admitted: float
    finished: float
    image: np.ndarray | None = None
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.spec.arrival


def frames_in_flight(spec: JobSpec, max_frames: int = 32, tuning=None) -> int:
    """The job's batch size: the autotuned one for its (scene, resolution)
    class when `tuning` (an autotune.TuningCache) has it, else max_frames;
    never more than the job's samples."""
    k = max_frames
    if tuning is not None:
        t = tuning.get(spec.scene, spec.width, spec.height, 0)
        if t is not None:
            k = t.frames_in_flight
    return max(1, min(k, spec.spp))


def launch_mode(spec: JobSpec, tuning=None, default: int = hpt.LAUNCH_GRAPH) -> int:
    if tuning is not None:
        t = tuning.get(spec.scene, spec.width, spec.height, 0)
        if t is not None:
            return t.launch
    return default


def shade_sort(spec: JobSpec, tuning=None) -> int:
    if tuning is not None:
        t = tuning.get(spec.scene, spec.width, spec.height, 0)
        if t is not None:
            return t.shade_sort
    return hpt.SHADE_SORT_OFF


def tail_from(spec: JobSpec, tuning=None) -> int:
    if tuning is not None:
        t = tuning.get(spec.scene, spec.width, spec.height, 0)
        if t is not None:
            return t.tail_from
    return hpt.TAIL_FROM_NONE


def traversal_block(spec: JobSpec, tuning=None) -> int:
    if tuning is not None:
        t = tuning.get(spec.scene, spec.width, spec.height, 0)
        if t is not None:
            return t.trav_block
    return 128


def footprint(spec: JobSpec, max_frames: int = 32, tuning=None) -> int:
    px = spec.width * spec.height
    per_path = PATH_BYTES + (SORT_BYTES if shade_sort(spec, tuning) != hpt.SHADE_SORT_OFF else 0)
    return px * (frames_in_flight(spec, max_frames, tuning) * per_path + IMAGE_BYTES)


class ResidentScene:
    """A scene committed once; jobs share its device copy."""

    def __init__(self, name: str, scene: hpt.Scene, features: int, device: int = 0):
        self.name = name
        self.features = features
        self.scene = scene
        t0 = time.perf_counter()
        self.template = hpt.HalaRenderer(8, 8, max_depth=1, rr_depth=0, features=features, device=device)
        self.template.frames_in_flight = 1
        self.template.set_scene(scene)
        self.template.commit()
        self.commit_seconds = time.perf_counter() - t0
        self.scene_bytes = self.template.memory_usage()[1]

    def close(self):
        self.template.close()


MAX_STEP_JOBS = 256  # jobs per wavefront step (hpt_group_render)


class _Active:
    """A job in a scene's wavefront."""

    def __init__(self, spec, handle, result, fp, k):
        self.spec, self.handle, self.result, self.fp, self.k = spec, handle, result, fp, k
        self.done = 0


class Server:
    def __init__(self, scenes: dict, budget_bytes: int, max_frames: int = 32, keep_images: bool = False,
                 device: int = 0, tuning=None, launch: int = hpt.LAUNCH_GRAPH, wavefront: bool = False,
                 step_paths: int = 8 << 20, denoiser=None):
        self.scenes = scenes                  # name -> ResidentScene
        self.budget = budget_bytes
        self.max_frames = max_frames
        self.tuning = tuning                  # autotune.TuningCache keyed by (scene, width, height, 0)
        self.launch = launch                  # launch mode of jobs the tuning does not cover
        self.keep_images = keep_images
        self.device = device
        self.wavefront = wavefront            # jobs of a scene share one wavefront per step
        self.step_paths = step_paths          # paths per job and wavefront step
        self.denoiser = denoiser              # denoiser.UNet: a finished job's image is denoised before read-back
        self._net_lock = threading.Lock()     # one network, one set of activation buffers
        self._lock = threading.Condition()
        self._used = 0
        self._pending = {}                    # scene -> jobs admitted but not yet in its wavefront
        self._closed = False
        self.step_sizes = []                  # jobs per wavefront step, in step order

    def _make_job(self, spec: JobSpec):
        rs = self.scenes[spec.scene]
        r = hpt.HalaRenderer(spec.width, spec.height, max_depth=spec.max_depth, rr_depth=spec.rr_depth,
                             enable_tonemap=True, features=rs.features, device=self.device)
        # A wavefront job traces in its scene's group; its own batch stays one frame.
        r.frames_in_flight = 1 if self.wavefront else frames_in_flight(spec, self.max_frames, self.tuning)
        r.share_scene(rs.template)
        r.set_launch_mode(launch_mode(spec, self.tuning, self.launch))
        r.set_shade_sort(shade_sort(spec, self.tuning))
        r.set_tail_from(tail_from(spec, self.tuning))
        r.set_traversal_block(traversal_block(spec, self.tuning))
        return r

    def _run_job(self, spec: JobSpec, t0: float, admitted: float, results: list, fp: int):
        res = JobResult(spec.job_id, spec, admitted, 0.0)
        try:
            r = self._make_job(spec)
            k = frames_in_flight(spec, self.max_frames, self.tuning)
            done = 0
            while done < spec.spp:
                n = min(k, spec.spp - done)
                r.render(n)
                r.wait_idle()
                done += n
            img = self._result_image(r)
            res.finished = time.perf_counter() - t0
            if self.keep_images:
                res.image = img
            r.close()
        except Exception as e:  # a failing job does not take the others down
            res.finished = time.perf_counter() - t0
            res.error = f"{type(e).__name__}: {e}"
        with self._lock:
            self._used -= fp
            results.append(res)
            self._lock.notify_all()

    def step_frames(self, spec: JobSpec) -> int:
        """Frames of `spec` per wavefront step: about step_paths paths."""
        return max(1, min(spec.spp, self.max_frames * 4, round(self.step_paths / (spec.width * spec.height))))

    def job_footprint(self, spec: JobSpec) -> int:
        """Device bytes admitting `spec` takes: its frame state, or in wavefront
        mode its share of the scene's wavefront plus its one-frame handle."""
        if not self.wavefront:
            return footprint(spec, self.max_frames, self.tuning)
        px = spec.width * spec.height
        return px * ((self.step_frames(spec) + 1) * PATH_BYTES + IMAGE_BYTES)

    def _result_image(self, r) -> np.ndarray:
        """A finished job's image in host memory: its accumulated radiance, or
        that image denoised on the device from the job's resident colour,
        albedo and normal images and its sample count."""
        if self.denoiser is None:
            return r.read_image(hpt.IMAGE_ACCUM)
        import denoiser

        with self._net_lock:
            out = self.denoiser.denoise_renderer(r)
            denoiser.synchronize()
            return out.numpy()

    def _new_group(self, scene: str):
        g = hpt.HalaGroup(self.device)
        g.set_launch_mode(self.launch)
        sorts = [t.shade_sort for key, t in (self.tuning.entries.items() if self.tuning else [])
                 if key.startswith(scene + "|")]
        if sorts:  # the choice most classes of this scene were tuned to
            g.set_shade_sort(max(set(sorts), key=sorts.count))
        tails = [t.tail_from for key, t in (self.tuning.entries.items() if self.tuning else [])
                 if key.startswith(scene + "|")]
        if tails:
            g.set_tail_from(max(set(tails), key=tails.count))
        blocks = [t.trav_block for key, t in (self.tuning.entries.items() if self.tuning else [])
                  if key.startswith(scene + "|")]
        if blocks:
            g.set_traversal_block(max(set(blocks), key=blocks.count))
        return g

    def _finish(self, a: _Active, t0: float, results: list, error: str | None = None):
        res = a.result
        try:
            if error is None:
                img = self._result_image(a.handle)
                if self.keep_images:
                    res.image = img
        except Exception as e:
            error = f"{type(e).__name__}: {e}"
        res.finished = time.perf_counter() - t0
        res.error = error
        try:
            if a.handle is not None:
                a.handle.close()
        finally:
            with self._lock:
                self._used -= a.fp
                results.append(res)
                self._lock.notify_all()

    def _scene_worker(self, scene: str, t0: float, results: list):
        """The wavefront of one scene: step after step, the next chunk of every
        active job together; finished jobs leave, newly admitted ones join."""
        g = None
        active, finishing = [], []
        try:
            while True:
                with self._lock:
                    while not self._pending[scene] and not active and not finishing and not self._closed:
                        self._lock.wait()
                    active.extend(self._pending[scene])
                    self._pending[scene].clear()
                    if not active and not finishing:
                        return
                ended = []
                if active:
                    if g is None:
                        g = self._new_group(scene)
                    step = active[:MAX_STEP_JOBS]
                    frames = [min(a.k, a.spec.spp - a.done) for a in step]
                    try:
                        g.render([a.handle for a in step], frames)  # returns with at most two steps queued
                    except Exception as e:  # the step's jobs fail; the server goes on
                        for a in step:
                            active.remove(a)
                            self._finish(a, t0, results, f"{type(e).__name__}: {e}")
                        continue
                    with self._lock:
                        self.step_sizes.append(len(step))
                    for a, n in zip(step, frames):
                        a.done += n
                        if a.done >= a.spec.spp:
                            active.remove(a)
                            ended.append(a)
                # Jobs whose last step was queued before this one: their read-back
                # waits for that step while the device runs the next.
                for a in finishing:
                    self._finish(a, t0, results)
                finishing = ended
        finally:
            if g is not None:
                g.close()

    def _run_wavefront(self, specs: list) -> list:
        results, workers = [], {}
        t0 = time.perf_counter()
        self._closed = False
        for spec in specs:
            wait = spec.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            fp = self.job_footprint(spec)
            with self._lock:
                while self._used + fp > self.budget:
                    self._lock.wait()
                self._used += fp
            admitted = time.perf_counter() - t0
            res = JobResult(spec.job_id, spec, admitted, 0.0)
            k = self.step_frames(spec)
            try:
                a = _Active(spec, self._make_job(spec), res, fp, k)
            except Exception as e:
                self._finish(_Active(spec, None, res, fp, k), t0, results, f"{type(e).__name__}: {e}")
                continue
            with self._lock:
                self._pending.setdefault(spec.scene, []).append(a)
                if spec.scene not in workers:
                    th = threading.Thread(target=self._scene_worker, args=(spec.scene, t0, results), daemon=True)
                    workers[spec.scene] = th
                    th.start()
                self._lock.notify_all()
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for th in workers.values():
            th.join()
        return results

    def run(self, specs: list) -> list:
        """Serves `specs` (arrival times relative to now); returns the results
        in completion order."""
        specs = sorted(specs, key=lambda s: s.arrival)
        for s in specs:
            if self.job_footprint(s) > self.budget:
                raise ValueError(f"job {s.job_id} needs {self.job_footprint(s)} B, above the budget")
        if self.wavefront:
            return self._run_wavefront(specs)
        results, threads = [], []
        t0 = time.perf_counter()
        for spec in specs:
            wait = spec.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            fp = footprint(spec, self.max_frames, self.tuning)
            with self._lock:
                while self._used + fp > self.budget:
                    self._lock.wait()
                self._used += fp
            admitted = time.perf_counter() - t0
            th = threading.Thread(target=self._run_job, args=(spec, t0, admitted, results, fp), daemon=True)
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        return results


def tune_classes(scenes: dict, specs: list, max_frames: int = 16, tolerance: float = 0.05, log=None):
    """Autotunes every (scene, resolution) class of `specs` at depth 10 on
    the resident scenes; returns the TuningCache the Server takes.  For
    serving, the batch is capped (max_frames) and the smallest one within
    `tolerance` of the best single-job rate wins: a job's memory is what
    bounds how many jobs run together, and on a shared GPU concurrency, not
    one job's batch, keeps the device full."""
    import autotune

    cache = autotune.TuningCache()
    for scene, w, h in sorted({(s.scene, s.width, s.height) for s in specs}):
        rs = scenes[scene]
        t = autotune.tune(rs.template, w, h, 10, features=rs.features, max_frames=max_frames, tolerance=tolerance)
        cache.put(scene, w, h, 0, t)
        if log:
            log(f"tuned {scene} {w}x{h}: {t.frames_in_flight} frames, "
                f"{'graph' if t.launch == hpt.LAUNCH_GRAPH else 'eager'}, {t.msamples_per_s:.0f} Msamples/s "
                f"({t.seconds:.2f} s; " + ", ".join(f"{k} {v:.0f}" for k, v in t.table.items()) + ")")
    return cache


def poisson_jobs(n: int, rate: float, scene_names: list, seed: int = 0, resolutions=None, spp_range=(16, 512),
                 depth_range=(8, 12)) -> list:
    """n jobs with exponential inter-arrival times at `rate` per second, the
    scene uniform over `scene_names`, resolution uniform over `resolutions`,
    spp log-uniform over spp_range, depth uniform over depth_range."""
    rng = np.random.default_rng(seed)
    resolutions = resolutions or [(256, 256), (512, 512), (1024, 768), (1920, 1080)]
    t = 0.0
    jobs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        w, h = resolutions[rng.integers(len(resolutions))]
        spp = int(round(math.exp(rng.uniform(math.log(spp_range[0]), math.log(spp_range[1])))))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        jobs.append(JobSpec(i, scene_names[rng.integers(len(scene_names))], w, h, spp, depth, 3, t))
    return jobs


def summarize(results: list, makespan: float) -> dict:
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    wait = np.array([r.admitted - r.spec.arrival for r in ok]) if ok else np.zeros(1)
    samples = sum(r.spec.width * r.spec.height * r.spec.spp for r in ok)
    return {"jobs": len(results), "failed": len(results) - len(ok),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "mean_admission_wait_s": round(float(wait.mean()), 4),
            "makespan_s": round(makespan, 3), "msamples_per_s": round(samples / makespan / 1e6, 2)}
