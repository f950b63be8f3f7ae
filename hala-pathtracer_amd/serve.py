"""Render serving on one GPU: progressive jobs of mixed scene, resolution,
samples and depth, admitted by device-memory footprint and run concurrently.

The loop itself is native (hpt_server_* in include/hpt.h, csrc/serve.cpp):
admission, the jobs' host threads, wavefront steps and read-back run in C++,
so a host without Python serves the same way.  This module holds the job
specs, the per-class tuning it passes into each job, the Poisson job draw,
the resident-scene wrapper and the latency summary.

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
* Execution.  An admitted job runs on its own HIP stream: a host thread of
  the server renders its batches of k frames one at a time (the streams share
  the process's few in-order hardware queues, so one batch in flight per job
  keeps a long job from holding the others); the device interleaves the
  kernels of the active jobs.  When its last batch ends, the job's image is read back (the
  latency end point), its handle is closed and its memory returns to the
  budget.
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

import io
import math
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


@dataclass
class DryRunScene:
    """A resident scene for dry runs (hpt_server_config.dry_run_ns_per_path):
    no device copy; `fail` makes each of its jobs fail when it starts."""
    features: int = 0
    fail: bool = False
    template = None


class Server:
    """Serves JobSpecs through the native serving loop (hpt_server_*,
    csrc/serve.cpp): admission in arrival order by device footprint, one
    stream per job or one wavefront per scene (`wavefront`), each job's
    image read back (denoised on the device when a denoiser is given).

    `scenes` maps names to ResidentScene (or DryRunScene with
    `dry_run_ns_per_path` > 0); `tuning` (an autotune.TuningCache keyed by
    (scene, width, height, 0)) gives each job its batch size, launch mode,
    shading sort, tail bounce and traversal block; in wavefront mode each
    scene's steps take the choice most of its classes were tuned to.  run()
    submits every spec at its arrival time (seconds after the call) and
    returns the results in completion order; times are seconds after the
    call.  `stats` holds the native server's counters of the last run."""

    def __init__(self, scenes: dict, budget_bytes: int, max_frames: int = 32, keep_images: bool = False,
                 device: int = 0, tuning=None, launch: int = hpt.LAUNCH_GRAPH, wavefront: bool = False,
                 step_paths: int = 8 << 20, denoiser=None, dry_run_ns_per_path: int = 0, max_running: int = 8,
                 reuse_handles: bool = True):
        self.scenes = scenes
        self.budget = budget_bytes
        self.max_frames = max_frames
        self.tuning = tuning
        self.launch = launch
        self.keep_images = keep_images
        self.device = device
        self.wavefront = wavefront
        self.step_paths = step_paths
        self.denoiser = denoiser          # denoiser.UNet or hpt.Denoiser
        self.dry_run_ns_per_path = dry_run_ns_per_path
        self.max_running = max_running    # streams mode: jobs running at once (0: as the budget allows)
        self.reuse_handles = reuse_handles  # a finished job's handle renders the next job of its class
        self.stats = {}
        self._net = None                  # the hpt.Denoiser the native server uses
        self._srv, self._ids = None, {}   # the native server, made at the first run

    def _native(self, budget=None):
        """A native server with this server's settings and scenes, and the
        scene ids by name."""
        srv = hpt.RenderServer(self.budget if budget is None else budget,
                               hpt.SERVE_WAVEFRONT if self.wavefront else hpt.SERVE_STREAMS, self.max_frames,
                               self.keep_images, self.step_paths, self.launch, self.device,
                               self.dry_run_ns_per_path, self.max_running, self.reuse_handles)
        ids = {}
        for name, rs in self.scenes.items():
            ids[name] = srv.add_scene(rs.template, rs.features, *self._step_settings(name),
                                      dry_run_fail=bool(getattr(rs, "fail", False)))
        return srv, ids

    def _step_settings(self, scene: str) -> tuple:
        """Launch mode, shading sort, tail bounce and traversal block of a
        scene's wavefront steps: the choice most of its tuned classes made."""
        entries = [t for key, t in (self.tuning.entries.items() if self.tuning else []) if key.startswith(scene + "|")]

        def most(attr):
            vals = [getattr(t, attr) for t in entries]
            return max(set(vals), key=vals.count) if vals else hpt.SERVE_DEFAULT

        return self.launch, most("shade_sort"), most("tail_from"), most("trav_block")

    def _desc(self, srv, scene_id: int, spec: JobSpec):
        return srv.job_desc(scene_id, spec.width, spec.height, spec.spp, spec.max_depth, spec.rr_depth,
                            frames_in_flight=frames_in_flight(spec, self.max_frames, self.tuning),
                            launch_mode=launch_mode(spec, self.tuning, self.launch),
                            shade_sort=shade_sort(spec, self.tuning), tail_from=tail_from(spec, self.tuning),
                            traversal_block=traversal_block(spec, self.tuning),
                            denoise=int(self.denoiser is not None))

    def _plan(self, spec: JobSpec) -> tuple:
        srv, ids = self._native(budget=1 << 62)
        try:
            return srv.plan(self._desc(srv, ids[spec.scene], spec))
        finally:
            srv.close()

    def step_frames(self, spec: JobSpec) -> int:
        """Frames of `spec` per batch (streams) or per wavefront step (about
        step_paths paths)."""
        return self._plan(spec)[1]

    def job_footprint(self, spec: JobSpec) -> int:
        """Device bytes admitting `spec` takes."""
        return self._plan(spec)[0]

    def _native_denoiser(self):
        if self.denoiser is None or isinstance(self.denoiser, hpt.Denoiser):
            return self.denoiser
        if self._net is None:  # a denoiser.UNet: its weights behind the C ABI
            buf = io.BytesIO()
            self.denoiser.save(buf)
            self._net = hpt.Denoiser(buf.getvalue(), self.device)
        return self._net

    def run(self, specs: list) -> list:
        """Serves `specs` (arrival times relative to now); returns the results
        in completion order.  Successive runs share one native server, so a
        job can render on the kept handle of an earlier job of its class;
        close() frees it."""
        specs = sorted(specs, key=lambda s: s.arrival)
        if self._srv is None:
            self._srv, self._ids = self._native()
            net = self._native_denoiser()
            if net is not None:
                self._srv.set_denoiser(net)
        srv, ids = self._srv, self._ids
        descs = [(s, self._desc(srv, ids[s.scene], s)) for s in specs]
        for s, d in descs:
            fp, _ = srv.plan(d)
            if fp > self.budget:
                raise ValueError(f"job {s.job_id} needs {fp} B, above the budget")
        t0 = srv.clock()
        by_id = {}
        for s, d in descs:
            wait = s.arrival - (srv.clock() - t0)
            if wait > 0:
                time.sleep(wait)
            by_id[srv.submit(d)] = s
        srv.drain()
        results = []
        while (r := srv.next_result(0.0)) is not None:
            s = by_id[r["job_id"]]
            img = srv.read_image(r["job_id"], s.width, s.height) if r["has_image"] else None
            results.append(JobResult(s.job_id, s, r["admitted"] - t0, r["finished"] - t0, img, r["error"]))
        self.stats = srv.stats()
        return results

    def close(self) -> None:
        """Stops the native server (waiting for its jobs) and frees its kept handles."""
        if getattr(self, "_srv", None) is not None:
            self._srv.close()
            self._srv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
