"""Throughput benchmark of the path-tracing hot path on one MI355X.

Workload: the Disney material grid with an importance-sampled sky map -- 7x7
UV spheres (64x32 segments, 194,434 triangles with the ground) at 1920x1080,
max depth 10, Russian roulette from depth 3.  A step is one frame: one camera
path per pixel through the full bounce loop plus accumulation; the default
--steps 256 is the configuration's 256 samples per pixel.

Reports Msamples/s (pixels x frames / wall seconds, inputs resident in HBM),
the roofline of the dominant kernel (BVH traversal of the extension rays,
HBM-bound) from HIP events recorded on the renderer's stream around every
launch, and the CPU oracle's throughput on a bounded sample of the same
workload on this host.
"""
from __future__ import annotations

import argparse
import json
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(scene, width, height, max_depth, rr_depth, budget_s=12.0):
    """Oracle throughput on the host: centred pixel rows of frame 0, grown
    until the sample takes a quarter of the budget; once the whole image
    fits, whole frames 0, 1, ... until the budget is spent.  All host threads
    of this process's CPU share (at most 16)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    o = pyoracle.OracleScene(scene)
    st = pyoracle.settings(width, height, max_depth, rr_depth)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
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


def profiled_traffic(frames_in_flight):
    """HBM bytes per k_extend launch from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the reading of a
    1 GiB copy), when it was taken at the same batch size; else None."""
    path = os.path.join(ROOT, "profiles", "k_extend_disney_grid_1080p.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"traffic_bytes_per_launch k_extend \(frames_in_flight (\d+)\) = (\d+)", line)
        if m and int(m.group(1)) == frames_in_flight:
            return int(m.group(2))
    return None


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--frames-in-flight", type=int, default=0,
                    help="frames traced together per batch (0: sized by the library)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run known-size copies (rocprof counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding of frames is "
              f"described in DESIGN.md and not measured here.", file=sys.stderr)
        sys.exit(2)

    W, H, DEPTH, RR = args.width, args.height, 10, 3
    log("building scene")
    scene = hpt_scenes.disney_grid(seed=0, width=W, height=H)
    log(f"{scene.triangle_count} triangles; creating renderer")
    r = hpt.HalaRenderer(W, H, max_depth=DEPTH, rr_depth=RR)
    r.set_scene(scene)
    r.frames_in_flight = args.frames_in_flight
    t0 = time.perf_counter()
    r.commit()
    log(f"commit (BVH build + upload) {time.perf_counter() - t0:.2f} s, bvh {r.bvh_info()}, "
        f"{r.frames_in_flight} frames in flight")
    fif = r.frames_in_flight

    # Visit counts per ray from a counted warm-up (separate kernel variant).
    r.set_profiling(hpt.PROFILE_TIMING | hpt.PROFILE_COUNT_VISITS)
    r.render(max(1, args.warmup))
    ext = r.stage_stats(hpt.STAGE_EXTEND)
    shd = r.stage_stats(hpt.STAGE_SHADOW)
    nodes_per_ray = ext["node_visits"] / max(1, ext["rays"])
    tris_per_ray = ext["tri_tests"] / max(1, ext["rays"])
    log(f"warm-up: extend {ext['rays']} rays, {nodes_per_ray:.2f} nodes/ray, {tris_per_ray:.2f} tris/ray")

    # Timed region: `steps` frames from a fresh accumulation.
    r.reset_accumulation()
    r.set_profiling(hpt.PROFILE_TIMING)
    r.wait_idle()
    t0 = time.perf_counter()
    r.render(args.steps)
    r.wait_idle()
    wall = time.perf_counter() - t0
    stats = {s: r.stage_stats(s) for s in range(5)}
    ext = stats[hpt.STAGE_EXTEND]
    shadow = stats[hpt.STAGE_SHADOW]
    img = r.read_image(hpt.IMAGE_ACCUM)
    assert np.isfinite(img).all()
    samples = W * H * args.steps
    msps = samples / wall / 1e6
    ext_ms = ext["total_ms"] / max(1, ext["launches"])
    bytes_per_launch = (ext["rays"] / max(1, ext["launches"])) * (
        NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray + RAY_BYTES)
    achieved = bytes_per_launch / (ext_ms * 1e-3) / 1e9
    rays_total = ext["rays"] + shadow["rays"]
    log("stage totals: " + ", ".join(f"{n} {stats[s]['total_ms']:.1f} ms/{stats[s]['launches']}"
                                       for s, n in enumerate(["camera", "extend", "shade", "shadow", "accum"])))
    log(f"wall {wall:.3f} s, {msps:.1f} Msamples/s, {rays_total / wall / 1e6:.1f} Mrays/s")

    base = None if args.no_cpu_baseline else cpu_baseline(scene, W, H, DEPTH, RR)
    r.close()
    out = {
        "metric": "Msamples/s",
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "disney_grid_env_1080p", "width": W, "height": H, "spp": args.steps,
                   "max_depth": DEPTH, "rr_depth": RR, "triangles": scene.triangle_count,
                   "env_map": [2048, 1024], "frames_in_flight": fif,
                   "mrays_per_s": round(rays_total / wall / 1e6, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(fif),
                     "kernel": "k_extend", "avg_launch_ms": round(ext_ms, 4),
                     "nodes_per_ray": round(nodes_per_ray, 3), "tris_per_ray": round(tris_per_ray, 3)},
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
