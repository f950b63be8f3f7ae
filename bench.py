"""Benchmark of the path-tracing hot path on one MI355X.

Default workload (the configuration the north-star target is quoted on):
the 1.31 M-triangle fBm-displaced icosphere with a point, spot,
directional, rect and sphere light and a homogeneous scattering medium,
thin-lens camera, 1920x1080, max depth 12, Russian roulette from depth 3,
the MEDIUM|VOL_MIS|TRANSPARENT permutation.  A step is one frame: one
camera path per pixel through the full bounce loop plus accumulation; the
default --steps is the configuration's 128 samples per pixel.  The line
carries
  * value: Msamples/s (pixels x frames / wall seconds, inputs resident in HBM);
  * roofline: the dominant stage's kernel (of extension, shading and the
    queued transmittance walks of the phased volume-MIS shading: on this
    scene the walks), each stage's also as roofline_extend, roofline_shade
    and roofline_walk, from HIP events recorded on the renderer's stream
    around every launch and an algorithmic byte model per unit;
  * time_to_relmse: the plain tracer, the tracer + learned denoiser and the
    served (submission to image in host memory) times to relMSE <= 1e-3
    against a 4096-spp render of the same scene;
  * serving: p50/p99 job latency and Msamples/s of 64 Poisson jobs at 20/s
    over the 8 resident scenes of the batched-job workload;
  * cpu_baseline: the CPU oracle on a bounded sample of the same workload.
--quick skips the relMSE and serving runs.  --workload selects the Disney
grid (1080p, 256 spp, depth 10), the Cornell box (512x512, 64 spp, depth 8),
the serving run alone, the denoiser alone, or `merge`: jobs of one scene
merged into one wavefront per step against the same jobs on their own
streams.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hala-pathtracer_amd"))

import hpt  # noqa: E402
import hpt_scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
NODE_BYTES, TRI_BYTES, RAY_BYTES = 80, 48, 48  # per visited node, tested triangle, ray (in + out)
# Per shading event: the survey's 384 B of scene records (material 144,
# primitive 96, indices 12, three vertices 132) plus the path state the
# shading kernel reads (origin, direction, throughput, hit, RNG, radiance,
# medium: 120 B, and its queue entry) and writes back (104 B, and one queue
# slot).
SHADE_EVENT_BYTES = 384 + 232
HEADLINE = "icosphere_lights_medium_1080p"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(scene, width, height, max_depth, rr_depth, features=0, budget_s=12.0, threads=0):
    """Oracle throughput on the host: centred pixel rows of frame 0, grown
    until the sample takes a quarter of the budget; once the whole image
    fits, whole frames 0, 1, ... until the budget is spent.  `threads` 0:
    all host threads of this process's CPU share (at most 16)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    o = pyoracle.OracleScene(scene)
    st = pyoracle.settings(width, height, max_depth, rr_depth, features)
    threads = threads or max(1, min(16, len(os.sched_getaffinity(0))))
    rows = 4
    while True:
        y0 = height // 2 - rows // 2
        t0 = time.perf_counter()
        o.trace_paths(st, 0, (0, y0, width, y0 + rows), threads=threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or rows >= height:
            break
        rows = min(height, rows * 4)
    frames = 1
    if rows >= height:
        while dt < budget_s:
            t0 = time.perf_counter()
            o.trace_paths(st, frames, threads=threads)
            dt += time.perf_counter() - t0
            frames += 1
    paths = width * rows * frames
    o.close()
    what = (f"frames 0-{frames - 1} of the {width}x{height} image" if frames > 1 else
            f"frame 0, {width}x{rows} centre rows of the {width}x{height} image")
    return {"value": paths / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{what} ({paths} paths, {dt:.1f} s)"}


def profiled_traffic(kernel, workload, batch_frames, tail_from, block, refill=True):
    """HBM bytes per launch of the stage `kernel` (its kernel family) from the
    committed rocprofv3 summary of the workload (FETCH_SIZE and WRITE_SIZE per
    dispatch, each scaled by the reading of a 1 GiB copy), when it was taken
    at the same batch size (frames traced together), tail bounce, traversal
    block size and lane refill setting; else None."""
    d = os.path.join(ROOT, "profiles")
    tail = "none" if tail_from == hpt.TAIL_FROM_NONE else str(tail_from)
    pat = (r"traffic_bytes_per_launch (\w+) \(frames_in_flight (\d+), tail_from (\w+), block (\d+)"
           r"(?:, refill (on|off))?\) = (\d+)")
    for name in sorted(os.listdir(d)) if os.path.isdir(d) else []:
        if not name.endswith(f"_{workload}.md"):
            continue
        for line in open(os.path.join(d, name)):
            m = re.match(pat, line)
            if m and m.group(1) == kernel and int(m.group(2)) == batch_frames and m.group(3) == tail and \
                    int(m.group(4)) == block and (m.group(5) or "off") == ("on" if refill else "off"):
                return int(m.group(6))
    return None


def profiled_conv_traffic(layer):
    """HBM bytes per launch of one denoiser convolution at 1080p from the
    committed rocprofv3 summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated);
    None when absent."""
    path = os.path.join(ROOT, "profiles", "k_conv_denoise_1080p.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"traffic_bytes_per_launch k_conv (\w+) = (\d+)", line)
        if m and m.group(1) == layer:
            return int(m.group(2))
    return None


def device_memory() -> tuple:
    """(free, total) device bytes of the current device, from the HIP runtime
    the library runs on."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return 0, 0
    return free.value, total.value


def copy_calibration(mib=1024, reps=4):
    """Plain device-to-device copies of a known size, for scaling rocprof's
    FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_dispatch": n * 4, "copy_bytes_written_per_dispatch": n * 4,
                      "dispatches": reps}), flush=True)


# The workloads (scene, image, depth, Russian-roulette start, shader permutation,
# samples per pixel): the large-mesh scene with five light types and a medium
# is the headline; the Disney grid and the Cornell box run with --workload.
WORKLOADS = {
    "disney_grid_env_1080p": dict(scene=lambda: hpt_scenes.disney_grid(seed=0), width=1920, height=1080,
                                  depth=10, rr=3, features=0, spp=256),
    "cornell_512": dict(scene=hpt_scenes.cornell_box, width=512, height=512, depth=8, rr=3, features=0, spp=64),
    "icosphere_lights_medium_1080p": dict(scene=lambda: hpt_scenes.displaced_icosphere_lights_medium(level=8),
                                          width=1920, height=1080, depth=12, rr=3,
                                          features=hpt_scenes.MEDIUM_SCENE_FEATURES, spp=128),
}


def relmse(img, ref):
    """mean over pixels and RGB of (I - R)^2 / (R^2 + 0.01)."""
    img = img[..., :3].astype(np.float64)
    ref = ref[..., :3].astype(np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


def time_to_relmse(make_renderer, ref, t_ref, ref_spp, threshold, chunk, max_spp):
    """The plain progressive estimate from frame 0 in chunks of `chunk`
    frames against the reference `ref`; the render time until the estimate's
    relMSE first meets `threshold` (relMSE evaluation excluded)."""
    r = make_renderer()
    r.render(chunk)  # warm-up: graph capture
    r.reset_accumulation()
    curve, t_render, spp, hit = [], 0.0, 0, None
    while spp < max_spp:
        r.wait_idle()
        t0 = time.perf_counter()
        r.render(chunk)
        r.wait_idle()
        t_render += time.perf_counter() - t0
        spp += chunk
        e = relmse(r.read_image(hpt.IMAGE_ACCUM), ref)
        curve.append([spp, round(t_render, 4), e])
        if hit is None and e <= threshold:
            hit = (spp, t_render)
            break
    r.close()
    log(f"relMSE after {spp} spp: {curve[-1][2]:.3e}" + (f"; threshold met at {hit[0]} spp, {hit[1]:.3f} s" if hit else ""))
    return {"threshold": threshold, "ref_spp": ref_spp, "ref_seconds": round(t_ref, 3),
            "time_to_threshold_s": round(hit[1], 4) if hit else None, "spp_at_threshold": hit[0] if hit else None,
            "curve": curve[:: max(1, len(curve) // 16)] + ([curve[-1]] if len(curve) % 16 else [])}


def render_reference(make_renderer, ref_spp):
    """A ref_spp reference from frames [65536, 65536 + ref_spp), so it shares
    no samples with an estimate that starts at frame 0."""
    r = make_renderer()
    base = 1 << 16
    r.frame_index = base
    t0 = time.perf_counter()
    r.render(ref_spp)
    r.wait_idle()
    t_ref = time.perf_counter() - t0
    ref = r.read_image(hpt.IMAGE_ACCUM)[..., :3].astype(np.float64) * ((base + ref_spp) / ref_spp)
    r.close()
    log(f"reference: {ref_spp} spp in {t_ref:.2f} s")
    return ref, t_ref


def time_to_relmse_denoised(make_renderer, ref, net, threshold, max_spp, batch=2):
    """The denoised estimate at k, 2k, 4k, ... spp for batches of k frames
    (progressive accumulation from frame 0): time = render time up to that
    sample count + one denoiser pass (on the renderer's resident images, same
    GPU); the first sample count whose denoised image meets `threshold`
    (relMSE evaluation and the passes of earlier sample counts excluded)."""
    import denoiser

    r = make_renderer(batch)
    r.render(batch)  # warm-up: graph capture
    curve, t_render, spp, hit = [], 0.0, 0, None
    out = denoiser.DeviceArray(r.height * r.width * 4, np.float32, (r.height, r.width, 4))
    net.denoise_renderer(r, out)  # allocate the network's buffers outside the timed region
    denoiser.synchronize()
    r.reset_accumulation()
    while spp < max_spp:
        step = batch if spp == 0 else spp
        r.wait_idle()
        t0 = time.perf_counter()
        r.render(step)
        r.wait_idle()
        t_render += time.perf_counter() - t0
        spp += step
        t0 = time.perf_counter()
        net.denoise_renderer(r, out)
        denoiser.synchronize()
        t_den = time.perf_counter() - t0
        e = relmse(out.numpy(), ref)
        e_plain = relmse(r.read_image(hpt.IMAGE_ACCUM), ref)
        curve.append([spp, round(t_render, 4), round(t_den * 1e3, 3), e_plain, e])
        if hit is None and e <= threshold:
            hit = (spp, t_render + t_den, t_den)
            break
    r.close()
    log(f"denoised relMSE ({batch}-frame batches): " +
        ", ".join(f"{c[0]} spp {c[4]:.2e} (plain {c[3]:.2e})" for c in curve))
    return {"threshold": threshold, "batch_frames": batch, "time_to_threshold_s": round(hit[1], 4) if hit else None,
            "spp_at_threshold": hit[0] if hit else None, "denoise_ms": round(hit[2] * 1e3, 3) if hit else None,
            "curve_spp_render_s_denoise_ms_plain_denoised": curve}


def served_relmse(scene, features, w, h, depth, rr, ref, net, spp_plain, spp_denoised, shade_sort, tails, blocks):
    """The same target through the serving path: jobs submitted to
    serve.Server -- plain at the plain tracer's threshold sample count,
    denoised (the server denoises finished jobs on the device) at the
    denoised threshold count -- timed from submission to the image in host
    memory (admission, tracing, denoising, read-back), then scored against
    the reference.  Each variant runs twice on one server: the first job of
    its class makes its handle and captures its batch graph (cold_latency_s);
    the second renders on the handle the first left (latency_s, the steady
    state of a server that has served that class before)."""
    import autotune
    import serve

    rs = serve.ResidentScene("scene", scene, features)
    out = {}
    try:
        for label, spp, den in (("plain", spp_plain, None), ("denoised", spp_denoised, net)):
            tuning = autotune.TuningCache()
            tuning.put("scene", w, h, 0, autotune.Tuning(32, hpt.LAUNCH_GRAPH, 0.0, {}, 0.0, shade_sort, tails[label],
                                                         blocks[label]))
            srv = serve.Server({"scene": rs}, 1 << 38, max_frames=32, keep_images=True, tuning=tuning, denoiser=den)
            try:
                cold = srv.run([serve.JobSpec(0, "scene", w, h, spp, depth, rr, 0.0)])[0]
                res = srv.run([serve.JobSpec(1, "scene", w, h, spp, depth, rr, 0.0)])[0]
                reused = srv.stats.get("handles_reused", 0)
            finally:
                srv.close()
            if cold.error or res.error:
                raise RuntimeError(cold.error or res.error)
            assert np.array_equal(cold.image, res.image)  # same job: same image on a kept handle
            out[label] = {"spp": spp, "latency_s": round(res.latency, 4), "cold_latency_s": round(cold.latency, 4),
                          "handle_reused": bool(reused), "relmse": relmse(res.image, ref)}
    finally:
        rs.close()
    out["speedup"] = round(out["plain"]["latency_s"] / out["denoised"]["latency_s"], 2)
    log(f"served: plain {out['plain']}, denoised {out['denoised']}, {out['speedup']}x sooner")
    return out


def train_step_time(denoiser, images, W, H, steps):
    """Seconds per training step (batch 8 of 128x128 crops of `images`,
    forward + L1 loss + backward + Adam) and its convolution FLOP."""
    tnet = denoiser.UNet(seed=1)
    bt = tnet._buffers(8, 128, 128, True)
    acc = denoiser.DeviceArray(1, np.float32)
    for i in range(8):
        tnet.make_input(images, W, H, 128 * i, 256, 128, 128, bt["x"].addr(i * 128 * 128 * 16), 16)
        tnet.make_target(images[0], W, H, 128 * i, 256, 128, 128, bt["target"].addr(i * 128 * 128 * 3))
    for _ in range(5):
        tnet.train_step(bt["x"], bt["target"], 8, 128, 128, 1e-4, loss_sum=acc)
    denoiser.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tnet.train_step(bt["x"], bt["target"], 8, 128, 128, 1e-4, loss_sum=acc)
    denoiser.synchronize()
    return (time.perf_counter() - t0) / steps, 3 * tnet.flops_forward(8, 128, 128)


def denoise_bench(args):
    """Inference throughput of the denoiser on a 1080p render (16 spp of the
    large-mesh scene with lights and a medium, resident in HBM), per-layer
    MFMA roofline from HIP events, training-step throughput, and the numpy
    oracle's forward pass on a host sample."""
    import denoiser

    W, H = args.width or 1920, args.height or 1080
    path = args.weights or denoiser.weights_path()
    net = denoiser.UNet.load(path) if os.path.exists(path) else denoiser.UNet(seed=0)
    log(f"weights: {path if os.path.exists(path) else 'untrained (seed 0)'}; {net.param_count} parameters")
    scene = hpt_scenes.displaced_icosphere_lights_medium(level=8)
    r = hpt.HalaRenderer(W, H, max_depth=12, rr_depth=3, features=hpt_scenes.MEDIUM_SCENE_FEATURES)
    r.set_scene(scene)
    r.commit()
    r.render(16)
    r.wait_idle()
    images = tuple(r.image_pointer(k)[0] for k in (hpt.IMAGE_ACCUM, hpt.IMAGE_ALBEDO, hpt.IMAGE_NORMAL))
    out = denoiser.DeviceArray(H * W * 4, np.float32, (H, W, 4))
    for _ in range(max(1, args.warmup)):
        net.denoise(images, W, H, 16, out)
    denoiser.synchronize()
    steps = args.steps or 50
    t0 = time.perf_counter()
    for _ in range(steps):
        net.denoise(images, W, H, 16, out)
    denoiser.synchronize()
    wall = time.perf_counter() - t0
    hp, wp = -(-H // 16) * 16, -(-W // 16) * 16
    flops = net.flops_forward(1, hp, wp)
    prof = net.profile_forward(1, hp, wp)
    conv_ms = sum(ms for ms, _ in prof.values())
    name, (ms, fl) = max(prof.items(), key=lambda kv: kv[1][0])
    log(f"{steps} passes in {wall:.3f} s: {wall / steps * 1e3:.3f} ms/pass, {flops / (wall / steps) / 1e12:.1f} "
        f"TFLOP/s overall; convolutions {conv_ms:.3f} ms; dominant {name}: {ms:.3f} ms, {fl / ms / 1e9:.1f} TFLOP/s")
    log("layers: " + ", ".join(f"{k} {v[0]:.3f} ms {v[1] / v[0] / 1e9:.0f} TF/s" for k, v in prof.items()))

    # Training step throughput: batch 8 of 128x128 crops of the same render.
    t_train, train_flops = None, 0
    if args.train_steps > 0:
        t_train, train_flops = train_step_time(denoiser, images, W, H, args.train_steps)
        log(f"training step (8 x 128x128): {t_train * 1e3:.2f} ms, {train_flops / t_train / 1e12:.1f} TFLOP/s")
    r.close()

    base = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import unet_ref

        ref_net = unet_ref.UNet(net.get_params(), net.levels, bf16_rounding=False)
        x = np.random.default_rng(0).uniform(0, 1, (1, 128, 128, 16)).astype(np.float32)
        ref_net.forward(x)
        t0, n = time.perf_counter(), 0
        while time.perf_counter() - t0 < 10.0:
            ref_net.forward(x)
            n += 1
        dt = time.perf_counter() - t0
        try:
            from threadpoolctl import threadpool_info

            cores = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
        except ImportError:
            cores = os.cpu_count()
        base = {"value": round(n * 128 * 128 / dt / 1e6, 4), "unit": "Mpixels/s", "cores": cores,
                "kind": "port", "sample": f"{n} forward passes of the numpy float32 oracle on 128x128 ({dt:.1f} s)"}
    out_json = {
        "metric": "denoised Mpixels/s", "value": round(W * H * steps / wall / 1e6, 3), "unit": "Mpixels/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "denoise", "width": W, "height": H, "network_size": [hp, wp],
                   "levels": list(net.levels), "parameters": net.param_count, "input_spp": 16,
                   "gflop_per_pass": round(flops / 1e9, 2),
                   "tflops_overall": round(flops / (wall / steps) / 1e12, 2),
                   "train_step_ms_8x128x128": round(t_train * 1e3, 3) if t_train else None,
                   "train_tflops": round(train_flops / t_train / 1e12, 2) if t_train else None,
                   "layer_ms": {k: round(v[0], 4) for k, v in prof.items()}},
        "roofline": {"bound": "mfma", "achieved": round(fl / ms / 1e9, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(fl / ms / 1e9 / 2500.0, 4),
                     "traffic": profiled_conv_traffic(name) if (W, H) == (1920, 1080) else None,
                     "kernel": f"k_conv ({name})",
                     "avg_launch_ms": round(ms, 4)},
        "cpu_baseline": base,
    }
    print(json.dumps(out_json), flush=True)


def run_serving(args, jobs_seed=0):
    """The batched/served workload: 8 resident scenes (hpt_scenes.batched_job_scenes),
    args.jobs Poisson arrivals at args.rate from the job distribution, served
    by serve.Server.  Returns (summary, resident scene sizes, jobs, makespan,
    tuning, tuning seconds)."""
    import serve

    log("building and committing the 8 resident scenes")
    resident = {}
    for name, sc, feat in hpt_scenes.batched_job_scenes():
        rs = serve.ResidentScene(name, sc, feat)
        resident[name] = rs
        log(f"  {name}: {sc.triangle_count} triangles, commit {rs.commit_seconds:.2f} s, "
            f"{rs.scene_bytes / 2**20:.0f} MiB")
    jobs = serve.poisson_jobs(args.jobs, args.rate, sorted(resident), seed=jobs_seed)
    offered = sum(j.width * j.height * j.spp for j in jobs)
    log(f"{len(jobs)} jobs, {offered / 1e9:.2f} Gsamples offered over {jobs[-1].arrival:.2f} s of arrivals")
    tuning, t_tune = None, 0.0
    if args.tune:
        t0 = time.perf_counter()
        tuning = serve.tune_classes(resident, jobs, max_frames=args.job_frames, log=log)
        t_tune = time.perf_counter() - t0
        log(f"autotuned {len(tuning.entries)} (scene, resolution) classes in {t_tune:.1f} s")
    srv = serve.Server(resident, int(args.budget_gib * 2**30), max_frames=args.job_frames, tuning=tuning,
                       launch=hpt.LAUNCH_GRAPH if args.launch == "graph" else hpt.LAUNCH_EAGER,
                       wavefront=args.wavefront, step_paths=int(args.step_mpaths * 2**20),
                       max_running=args.max_running)
    t0 = time.perf_counter()
    res = srv.run(jobs)
    makespan = time.perf_counter() - t0
    srv.close()
    summ = serve.summarize(res, makespan)
    summ["offered_gsamples"] = round(offered / 1e9, 3)
    summ["arrival_span_s"] = round(jobs[-1].arrival, 3)
    summ["arrival_rate_per_s"] = args.rate
    summ["mode"] = "wavefront per scene" if args.wavefront else "stream per job"
    summ["loop"] = "native (hpt_server)"
    summ["max_running"] = args.max_running
    summ["handles_reused"] = srv.stats.get("handles_reused", 0)
    if args.wavefront and srv.stats.get("wavefront_steps"):
        st = srv.stats
        summ["wavefront_steps"] = st["wavefront_steps"]
        summ["mean_jobs_per_step"] = round(st["jobs_in_steps"] / st["wavefront_steps"], 2)
        summ["max_jobs_per_step"] = st["max_jobs_per_step"]
    log(f"served: {summ}")
    sizes = {k: v.scene.triangle_count for k, v in resident.items()}
    for rs in resident.values():
        rs.close()
    return summ, sizes, jobs, makespan, tuning, t_tune


def serving_bench(args):
    summ, sizes, jobs, makespan, tuning, t_tune = run_serving(args)
    out = {
        "metric": "Msamples/s", "value": summ["msamples_per_s"], "unit": "Msamples/s", "n_gpus": 1,
        "steps": len(jobs), "warmup": 0, "ms_per_step": round(makespan / len(jobs) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "serving", "jobs": len(jobs), "arrival_rate_per_s": args.rate,
                   "resident_scenes": sizes,
                   "budget_gib": args.budget_gib, "job_frames_in_flight": args.job_frames,
                   "autotuned": {k: [v.frames_in_flight, "graph" if v.launch == hpt.LAUNCH_GRAPH else "eager",
                                     "sort off" if v.shade_sort == hpt.SHADE_SORT_OFF else f"sort from {v.shade_sort}"]
                                 for k, v in tuning.entries.items()} if tuning else None,
                   "autotune_seconds": round(t_tune, 2), "launch": args.launch,
                   "mode": summ["mode"],
                   "step_mpaths": args.step_mpaths if args.wavefront else None},
        "serving": summ, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def merge_bench(args):
    """J jobs of one resident scene (512x512, 16-frame chunks, depth 10),
    traced either on their own streams (the serving default, one thread per
    job) or merged into one wavefront per step (hpt_group_render), for J in
    1..16 on three of the 8 resident scenes.  Merging jobs of one scene is the
    most a wavefront shared across scenes could gain: the same BVH in every
    wave and no per-path scene lookups."""
    import threading
    import serve

    W, H, F, DEPTH = 512, 512, 16, 10
    scenes = {n: (sc, f) for n, sc, f in hpt_scenes.batched_job_scenes()}
    names = [n for n in scenes if n in ("cornell", "disney_grid")] + [max((n for n in scenes if n.startswith("soup_")),
                                                                        key=lambda n: scenes[n][0].triangle_count)]
    rows = {}
    for name in names:
        sc, feat = scenes[name]
        rs = serve.ResidentScene(name, sc, feat)
        log(f"{name}: {sc.triangle_count} triangles")

        def job(frames):
            r = hpt.HalaRenderer(W, H, max_depth=DEPTH, rr_depth=3, features=feat)
            r.frames_in_flight = frames
            r.share_scene(rs.template)
            r.set_launch_mode(hpt.LAUNCH_GRAPH)
            return r

        def chunks(r, n):
            for _ in range(n):
                r.render(F)
                r.wait_idle()

        for J in (1, 2, 4, 8, 16):
            R = max(2, 16 // J)
            samples = J * R * F * W * H
            jobs = [job(F) for _ in range(J)]
            for r in jobs:
                chunks(r, 1)  # allocation and graph capture
            threads = [threading.Thread(target=chunks, args=(r, R)) for r in jobs]
            t0 = time.perf_counter()
            for t in threads:
                t.start()
            for t in threads:
                t.join()
            streams = samples / (time.perf_counter() - t0) / 1e6
            for r in jobs:
                r.close()
            jobs = [job(1) for _ in range(J)]
            g = hpt.HalaGroup()
            g.set_launch_mode(hpt.LAUNCH_GRAPH)
            g.render(jobs, [F] * J)
            g.wait_idle()
            t0 = time.perf_counter()
            for _ in range(R):
                g.render(jobs, [F] * J)
            g.wait_idle()
            merged = samples / (time.perf_counter() - t0) / 1e6
            g.close()
            for r in jobs:
                r.close()
            log(f"  {J:2d} jobs: streams {streams:7.1f}, merged {merged:7.1f} Msamples/s ({merged / streams:.3f})")
            rows[f"{name}/{J}"] = {"streams": round(streams, 1), "merged": round(merged, 1)}
        rs.close()
    best = max(rows, key=lambda k: rows[k]["merged"])
    out = {"metric": "Msamples/s", "value": rows[best]["merged"], "unit": "Msamples/s", "n_gpus": 1,
           "steps": len(rows), "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "merge", "width": W, "height": H, "frames_per_chunk": F, "max_depth": DEPTH,
                      "scenes": names},
           "merge": rows, "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def stage_roofline(kernel, stats_timed, per_unit_bytes, units_key="rays", traffic=None, extra=None):
    """Roofline of one stage's kernel: algorithmic bytes per launch (units
    per launch in the timed run x bytes per unit) over its mean launch time
    from the HIP events bracketing every launch on the renderer's stream."""
    launches = max(1, stats_timed["launches"])
    ms = stats_timed["total_ms"] / launches
    bytes_per_launch = stats_timed[units_key] / launches * per_unit_bytes
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
           "avg_launch_ms": round(ms, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch),
           "launches": stats_timed["launches"]}
    if extra:
        out.update(extra)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["serving", "denoise", "merge"], default=HEADLINE)
    ap.add_argument("--denoise", action="store_true",
                    help="with --relmse: also the time to the threshold of the denoised estimate (and served)")
    ap.add_argument("--weights", default="", help="denoiser weights (default: the committed network)")
    ap.add_argument("--train-steps", type=int, default=50,
                    help="denoise: timed training steps (0 skips the training measurement)")
    ap.add_argument("--jobs", type=int, default=64, help="serving: number of jobs")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: Poisson arrivals per second")
    ap.add_argument("--budget-gib", type=float, default=192.0, help="serving: device memory for admitted jobs")
    ap.add_argument("--job-frames", type=int, default=16, help="serving: frames in flight per job (at most)")
    ap.add_argument("--wavefront", action="store_true",
                    help="serving: the jobs of each scene share one wavefront per step (hpt_group_render)")
    ap.add_argument("--max-running", type=int, default=8,
                    help="serving: jobs rendering at once in stream-per-job mode (0: as many as the budget admits)")
    ap.add_argument("--step-mpaths", type=float, default=8.0,
                    help="serving --wavefront: paths (in 2^20) per job and step")
    ap.add_argument("--tune", action="store_true",
                    help="serving: autotune frames in flight and launch mode per (scene, resolution) first")
    ap.add_argument("--steps", type=int, default=None, help="frames (default: the workload's spp)")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--width", type=int, default=None)
    ap.add_argument("--height", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads for cpu_baseline (0: all, at most 16)")
    ap.add_argument("--shade-sort", default="auto",
                    help="material-sorted shading from this bounce on ('off': never; 'auto': the autotuner's "
                         "choice among off and bounces 0-2 at this batch size, measured before the timed region)")
    ap.add_argument("--launch", choices=["graph", "eager"], default="graph",
                    help="full batches replayed as captured HIP graphs (default) or every kernel launched")
    ap.add_argument("--traversal-block", default="auto",
                    help="threads per block of the traversal kernels: auto (autotuned with the sort), 64, 128 or 256")
    ap.add_argument("--traversal-refill", choices=["on", "off"], default="on",
                    help="traversal lanes refill from their queues (resident blocks; on) or trace one entry each")
    ap.add_argument("--tail-from", default="auto",
                    help="first bounce run by the tail kernel: auto (autotuned), none, or a bounce index")
    ap.add_argument("--vmis-phases", choices=["on", "off"], default="on",
                    help="volume-MIS shading as phase kernels with queued walks (on) or one kernel with inline walks")
    ap.add_argument("--frames-in-flight", type=int, default=0,
                    help="frames traced together per batch (0: sized by the library)")
    ap.add_argument("--relmse", action="store_true",
                    help="also measure the plain tracer's time to relMSE <= --threshold against a reference "
                         "(on by default for the headline workload)")
    ap.add_argument("--serving", action="store_true",
                    help="also run the served workload (p50/p99 job latency at --rate; on by default for the "
                         "headline workload)")
    ap.add_argument("--quick", action="store_true",
                    help="headline workload without the relMSE and serving measurements")
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--threshold", type=float, default=1e-3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run known-size copies (rocprof counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): frames shard across GPUs by frame "
              f"range with one sum-reduce of the accumulators at the end, which is not built or measured here.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving":
        serving_bench(args)
        return
    if args.workload == "denoise":
        denoise_bench(args)
        return
    if args.workload == "merge":
        merge_bench(args)
        return
    if args.workload == HEADLINE and not args.quick:
        args.relmse = args.denoise = args.serving = True
    wl = WORKLOADS[args.workload]
    W = args.width or wl["width"]
    H = args.height or wl["height"]
    DEPTH, RR, FEATURES = wl["depth"], wl["rr"], wl["features"]
    steps = args.steps or wl["spp"]
    log(f"building scene {args.workload}")
    scene = wl["scene"]()
    log(f"{scene.triangle_count} triangles; creating renderer")

    def make_renderer(frames_in_flight=None):
        """The headline renderer, or one with `frames_in_flight`-frame
        batches and the tail bounce tuned for that size."""
        rr = hpt.HalaRenderer(W, H, max_depth=DEPTH, rr_depth=RR, features=FEATURES)
        rr.set_scene(scene)
        rr.frames_in_flight = frames_in_flight or args.frames_in_flight
        rr.commit()
        rr.set_launch_mode(hpt.LAUNCH_GRAPH if args.launch == "graph" else hpt.LAUNCH_EAGER)
        rr.set_shade_sort(shade_sort)
        rr.set_vmis_phases(args.vmis_phases == "on")
        rr.set_tail_from(tails["denoised"] if frames_in_flight else tails["plain"])
        rr.set_traversal_block(blocks["denoised"] if frames_in_flight else blocks["plain"])
        rr.set_traversal_refill(args.traversal_refill == "on")
        return rr

    shade_sort = hpt.SHADE_SORT_OFF if args.shade_sort in ("off", "auto") else int(args.shade_sort)
    fixed_tail = hpt.TAIL_FROM_NONE if args.tail_from in ("none", "auto") else int(args.tail_from)
    tails = {"plain": fixed_tail, "denoised": fixed_tail}  # big batches / 2-frame batches
    fixed_block = 128 if args.traversal_block == "auto" else int(args.traversal_block)
    blocks = {"plain": fixed_block, "denoised": fixed_block}  # threads per block of the traversal kernels
    t0 = time.perf_counter()
    r = make_renderer()
    log(f"create + commit (BVH build + upload) {time.perf_counter() - t0:.2f} s, bvh {r.bvh_info()}, "
        f"{r.frames_in_flight} frames in flight")
    free_b, total_b = device_memory()
    log(f"device memory: {free_b / 2**30:.1f} GiB free of {total_b / 2**30:.1f} GiB; this handle holds "
        f"{sum(r.memory_usage()) / 2**30:.1f} GiB")
    fif = r.frames_in_flight
    sort_table = None
    if args.shade_sort == "auto":
        import autotune

        t0 = time.perf_counter()
        tuned = autotune.tune(r, W, H, DEPTH, RR, FEATURES, samples=4 * fif * W * H, candidates=(fif,),
                              max_frames=fif, modes=(hpt.LAUNCH_GRAPH if args.launch == "graph" else hpt.LAUNCH_EAGER,),
                              tolerance=0.05, in_place=True)
        shade_sort = tuned.shade_sort
        sort_table = {k: round(v, 1) for k, v in tuned.table.items()}
        r.set_shade_sort(shade_sort)
        if args.tail_from == "auto":
            tails["plain"] = tuned.tail_from
            r.set_tail_from(tuned.tail_from)
        if args.traversal_block == "auto":
            blocks["plain"] = tuned.trav_block
            r.set_traversal_block(tuned.trav_block)
        log(f"shade sort and tail autotuned in {time.perf_counter() - t0:.1f} s: "
            f"{'off' if shade_sort == hpt.SHADE_SORT_OFF else f'from bounce {shade_sort}'}, tail from "
            f"{'none' if tails['plain'] == hpt.TAIL_FROM_NONE else tails['plain']} ({sort_table})")

    # Visit counts per ray (extend), per shading event, and of the volume-MIS
    # transmittance walks (inline in shading, or the walk kernel's when
    # shading is phased) from a counted warm-up (separate kernel variants).
    r.set_profiling(hpt.PROFILE_TIMING | hpt.PROFILE_COUNT_VISITS)
    r.render(max(1, args.warmup))
    ext = r.stage_stats(hpt.STAGE_EXTEND)
    sh = r.stage_stats(hpt.STAGE_SHADE)
    wk = r.stage_stats(hpt.STAGE_WALK)
    nodes_per_ray = ext["node_visits"] / max(1, ext["rays"])
    tris_per_ray = ext["tri_tests"] / max(1, ext["rays"])
    walks = wk if wk["queries"] else sh  # the stage whose kernels ran the walks
    walk_queries = walks["queries"] / max(1, sh["rays"])  # per shading event
    walk_bytes_per_query = (NODE_BYTES * walks["node_visits"] + TRI_BYTES * walks["tri_tests"] +
                            RAY_BYTES * walks["queries"]) / max(1, walks["queries"])
    phased = wk["launches"] > 0
    log(f"warm-up: extend {ext['rays']} rays, {nodes_per_ray:.2f} nodes/ray, {tris_per_ray:.2f} tris/ray; "
        f"shade {sh['rays']} events, {walk_queries:.3f} walk queries per event, {walk_bytes_per_query:.0f} B per "
        f"walk query ({'walk kernel' if phased else 'inline'})")

    # Timed region: `steps` frames from a fresh accumulation, submitted the way
    # the renderer submits them (full batches as graph replays by default).
    r.reset_accumulation()
    r.set_profiling(0)
    r.wait_idle()
    t0 = time.perf_counter()
    r.render(steps)
    r.wait_idle()
    wall = time.perf_counter() - t0
    img = r.read_image(hpt.IMAGE_ACCUM)
    assert np.isfinite(img).all()
    graph_launches = r.launch_stats()[0]
    # Kernel timings: the same frames again with every stage bracketed by HIP
    # events on the renderer's stream (eager launches: an event cannot sit
    # inside a replayed graph).  Same kernels, same work.
    r.reset_accumulation()
    r.set_profiling(hpt.PROFILE_TIMING)
    r.wait_idle()
    t0 = time.perf_counter()
    r.render(steps)
    r.wait_idle()
    profiled_wall = time.perf_counter() - t0
    stats = {s: r.stage_stats(s) for s in range(hpt.STAGE_TAIL + 1)}
    ext_t, shade_t, shadow_t = stats[hpt.STAGE_EXTEND], stats[hpt.STAGE_SHADE], stats[hpt.STAGE_SHADOW]
    samples = W * H * steps
    msps = samples / wall / 1e6
    batch_frames = min(steps, fif)
    walk_t = stats[hpt.STAGE_WALK]
    # rays: closest-hit and shadow queries, the walks' closest-hit queries,
    # and the tail's closest-hit queries (one per bounce it traced)
    rays_total = (ext_t["rays"] + shadow_t["rays"] + int(round(shade_t["rays"] * walk_queries)) +
                  stats[hpt.STAGE_TAIL]["queries"])
    names = ["camera", "extend", "shade", "shadow", "accum", "sort", "walk", "tail"]
    log("stage totals: " + ", ".join(f"{n} {stats[s]['total_ms']:.1f} ms/{stats[s]['launches']}"
                                       for s, n in enumerate(names)))
    log(f"wall {wall:.3f} s ({args.launch}, {graph_launches} graph replays), {msps:.1f} Msamples/s, "
        f"{rays_total / wall / 1e6:.1f} Mrays/s; eager with stage events {profiled_wall:.3f} s")
    r.close()

    # Rooflines: the extension walk (80 B per node, 48 B per triangle, 48 B
    # per ray); shading (616 B per shading event -- material, primitive,
    # indices and vertices as the survey's 384 B plus 232 B of path state read
    # and written -- plus, when the walks run inline, its volume-MIS walks at
    # the extension's rate per closest-hit query); the walk kernel (that rate
    # per query) when shading is phased.
    refill = args.traversal_refill == "on"
    roof_ext = stage_roofline("k_extend_refill" if refill else "k_extend", ext_t,
                              NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray + RAY_BYTES,
                              traffic=profiled_traffic("k_extend", args.workload, batch_frames, tails["plain"],
                                                       blocks["plain"], refill),
                              extra={"nodes_per_ray": round(nodes_per_ray, 3), "tris_per_ray": round(tris_per_ray, 3)})
    shade_bytes = SHADE_EVENT_BYTES + (0.0 if phased else walk_queries * walk_bytes_per_query)
    shade_name = "k_vmis_a/b/c" if phased else "k_shade"
    roof_shade = stage_roofline(shade_name, shade_t, shade_bytes,
                                traffic=profiled_traffic("k_vmis" if phased else "k_shade", args.workload,
                                                         batch_frames, tails["plain"], blocks["plain"], refill),
                                extra={"bytes_per_event": round(shade_bytes, 1),
                                       "walk_queries_per_event": round(walk_queries, 4)})
    rooflines = {"extend": roof_ext, "shade": roof_shade}
    if phased:
        # the walk kernel's units are its closest-hit queries: the timed
        # run's shading events times the warm-up's queries per event
        walk_t = dict(walk_t, rays=int(round(shade_t["rays"] * walk_queries)))
        rooflines["walk"] = stage_roofline("k_walk_refill" if refill else "k_walk", walk_t, walk_bytes_per_query,
                                           traffic=profiled_traffic("k_walk", args.workload, batch_frames,
                                                                    tails["plain"], blocks["plain"], refill),
                                           extra={"bytes_per_query": round(walk_bytes_per_query, 1)})
    stage_of = {"extend": ext_t, "shade": shade_t, "shadow": shadow_t, "walk": walk_t}
    dominant = max((k for k in rooflines), key=lambda k: stage_of[k]["total_ms"])
    roofline = rooflines[dominant]
    log(f"dominant stage: {dominant}; " + ", ".join(f"{v['kernel']} {v['achieved']:.0f} GB/s ({v['frac']:.3f})"
                                                    for v in rooflines.values()))

    conv = None
    if args.relmse:
        ref, t_ref = render_reference(make_renderer, args.ref_spp)
        conv = time_to_relmse(make_renderer, ref, t_ref, args.ref_spp, args.threshold, max(1, fif),
                              args.ref_spp // 2)
        if args.denoise:
            import denoiser

            path = args.weights or denoiser.weights_path()
            # The denoised estimate in batches of 2, 4 and 8 frames (checked after
            # every doubling of the sample count), each batch size with its own
            # tuned tail bounce and traversal block; the fastest is reported.
            net = denoiser.UNet.load(path)
            by_batch, tuned_small = {}, {}
            for k in (2, 4, 8):
                if args.tail_from == "auto":
                    import autotune

                    tmpl = make_renderer(frames_in_flight=k)
                    tk = autotune.tune(tmpl, W, H, DEPTH, RR, FEATURES, samples=4 * k * W * H, candidates=(k,),
                                       modes=(hpt.LAUNCH_GRAPH,), sorts=(), in_place=True)
                    tmpl.close()
                    tails["denoised"] = tk.tail_from
                    if args.traversal_block == "auto":
                        blocks["denoised"] = tk.trav_block
                    log(f"tail for {k}-frame batches: "
                        f"{'none' if tk.tail_from == hpt.TAIL_FROM_NONE else tk.tail_from} "
                        f"({ {key: round(v, 1) for key, v in tk.table.items()} })")
                tuned_small[k] = (tails["denoised"], blocks["denoised"])
                by_batch[k] = time_to_relmse_denoised(make_renderer, ref, net, args.threshold, args.ref_spp // 2, k)
            hits = [k for k in by_batch if by_batch[k]["time_to_threshold_s"]]
            best = min(hits, key=lambda k: by_batch[k]["time_to_threshold_s"]) if hits else 2
            tails["denoised"], blocks["denoised"] = tuned_small[best]
            den = dict(by_batch[best])
            den["by_batch_frames"] = {k: {"time_to_threshold_s": v["time_to_threshold_s"],
                                          "spp_at_threshold": v["spp_at_threshold"],
                                          "tail_from": tuned_small[k][0] if tuned_small[k][0] != hpt.TAIL_FROM_NONE
                                          else None} for k, v in by_batch.items()}
            den["weights"] = os.path.relpath(path, ROOT)
            if den["time_to_threshold_s"] and conv["time_to_threshold_s"]:
                den["speedup_vs_plain"] = round(conv["time_to_threshold_s"] / den["time_to_threshold_s"], 2)
                conv["served"] = served_relmse(scene, FEATURES, W, H, DEPTH, RR, ref, net,
                                               conv["spp_at_threshold"], den["spp_at_threshold"], shade_sort, tails, blocks)
            conv["denoised"] = den
    serving = None
    if args.serving:
        serving, _, _, _, _, _ = run_serving(args)
    base = None if args.no_cpu_baseline else cpu_baseline(scene, W, H, DEPTH, RR, FEATURES, threads=args.cpu_threads)
    config = {"workload": args.workload, "width": W, "height": H, "spp": steps, "max_depth": DEPTH,
              "rr_depth": RR, "features": FEATURES, "triangles": scene.triangle_count,
              "lights": int(len(scene.lights)), "frames_in_flight": fif, "batch_frames": batch_frames,
              "launch": args.launch,
              "shade_sort": "off" if shade_sort == hpt.SHADE_SORT_OFF else shade_sort,
              "shade_sort_tuning": sort_table,
              "tail_from": {k: (None if v == hpt.TAIL_FROM_NONE else v) for k, v in tails.items()},
              "traversal_block": blocks,
              "traversal_refill": args.traversal_refill,
              "vmis_phases": args.vmis_phases if FEATURES & hpt.FEATURE_VOL_MIS else None,
              "eager_profiled_wall_s": round(profiled_wall, 4),
              "stage_ms": {n: round(stats[s]["total_ms"], 3) for s, n in enumerate(names)},
              "mrays_per_s": round(rays_total / wall / 1e6, 2)}
    if scene.env_rgb is not None:
        config["env_map"] = [int(scene.env_rgb.shape[1]), int(scene.env_rgb.shape[0])]
    out = {
        "metric": "Msamples/s",
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "roofline_extend": roof_ext,
        "roofline_shade": roof_shade,
        "roofline_walk": rooflines.get("walk"),
        "cpu_baseline": base,
    }
    if conv is not None:
        tt = {"threshold": args.threshold, "ref_spp": args.ref_spp, "plain_s": conv["time_to_threshold_s"],
              "plain_spp": conv["spp_at_threshold"]}
        den = conv.get("denoised") or {}
        tt["denoised_s"] = den.get("time_to_threshold_s")
        tt["denoised_spp"] = den.get("spp_at_threshold")
        sv = conv.get("served") or {}
        if sv:
            tt["served_plain_s"] = sv["plain"]["latency_s"]
            tt["served_denoised_s"] = sv["denoised"]["latency_s"]
            tt["served_speedup_vs_plain"] = round(conv["time_to_threshold_s"] / sv["denoised"]["latency_s"], 2)
        out["time_to_relmse"] = tt
        out["relmse"] = conv
    if serving is not None:
        out["serving"] = serving
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
