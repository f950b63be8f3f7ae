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


class Server:
    def __init__(self, scenes: dict, budget_bytes: int, max_frames: int = 32, keep_images: bool = False,
                 device: int = 0, tuning=None, launch: int = hpt.LAUNCH_GRAPH):
        self.scenes = scenes                  # name -> ResidentScene
        self.budget = budget_bytes
        self.max_frames = max_frames
        self.tuning = tuning                  # autotune.TuningCache keyed by (scene, width, height, 0)
        self.launch = launch                  # launch mode of jobs the tuning does not cover
        self.keep_images = keep_images
        self.device = device
        self._lock = threading.Condition()
        self._used = 0

    def _make_job(self, spec: JobSpec):
        rs = self.scenes[spec.scene]
        r = hpt.HalaRenderer(spec.width, spec.height, max_depth=spec.max_depth, rr_depth=spec.rr_depth,
                             enable_tonemap=True, features=rs.features, device=self.device)
        r.frames_in_flight = frames_in_flight(spec, self.max_frames, self.tuning)
        r.share_scene(rs.template)
        r.set_launch_mode(launch_mode(spec, self.tuning, self.launch))
        r.set_shade_sort(shade_sort(spec, self.tuning))
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
            img = r.read_image(hpt.IMAGE_ACCUM)
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

    def run(self, specs: list) -> list:
        """Serves `specs` (arrival times relative to now); returns the results
        in completion order."""
        specs = sorted(specs, key=lambda s: s.arrival)
        for s in specs:
            if footprint(s, self.max_frames, self.tuning) > self.budget:
                raise ValueError(f"job {s.job_id} needs {footprint(s, self.max_frames, self.tuning)} B, above the budget")
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
