"""Autotuner of the tracer's launch parameters (no counterpart in the
reference, whose frame loop has no tunables).

What it tunes: the frames traced together per batch (the wavefront and
queue size, `hpt_set_frames_in_flight`), the launch mode (graph replay or
eager launches), the shading sort threshold (the first bounce whose paths
are grouped by material before shading, `hpt_set_shade_sort`, or never) and
the bounce from which the live paths run to their end in the tail kernel
(`hpt_set_tail_from`, or none) and the threads per block of the traversal
kernels (`hpt_set_traversal_block`) for one (scene, resolution, depth) class.  Larger batches
fill the GPU with more paths per launch and amortise the short tail bounces;
they cost k * W * H * 288 bytes of device memory, which the serving loop
spends from its admission budget.  The tuner measures, it does not model:
each candidate renders a warm-up batch and then a timed run on the job's own
scene, and the smallest batch within `tolerance` of the best throughput
wins (less memory and a finer progress granularity for the same speed),
then graph replay, then no sort (it costs 5 bytes per path); the tail
bounce and the traversal block size, which cost no memory, are the fastest
at that choice.

Results are kept in a TuningCache keyed by (scene name, width, height,
max_depth) and can be saved as JSON.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field

import hpt

PATH_BYTES = 288
SORT_BYTES = 5  # key byte and sorted entry per queue slot while shading sort is on
CANDIDATES = (2, 4, 8, 16, 32, 64)
SORT_CANDIDATES = (0, 1, 2)
TAIL_CANDIDATES = (3, 4, 5, 6, 8)
BLOCK_CANDIDATES = (64, 256)  # besides the default 128
# A tail bounce or block size replaces the current choice only when it is this
# much faster: near-equal candidates then do not flip from run to run.
MARGIN = 0.015


@dataclass
class Tuning:
    frames_in_flight: int
    launch: int
    msamples_per_s: float
    table: dict = field(default_factory=dict)  # "k/mode[/sN]" -> Msamples/s
    seconds: float = 0.0                       # time the tuning took
    shade_sort: int = hpt.SHADE_SORT_OFF       # first sorted bounce
    tail_from: int = hpt.TAIL_FROM_NONE        # first bounce run by the tail kernel
    trav_block: int = 128                      # threads per block of the traversal kernels


def candidates_for(width: int, height: int, max_frames: int = 64, memory_bytes: int | None = None,
                   candidates=CANDIDATES) -> list:
    out = []
    for k in candidates:
        if k > max_frames:
            continue
        if memory_bytes is not None and k * width * height * PATH_BYTES > memory_bytes:
            continue
        if k * width * height * 2 > 0xFFFFFFFF:  # 32-bit queue indices
            continue
        out.append(k)
    return out or [1]


def tune(template: "hpt.HalaRenderer", width: int, height: int, max_depth: int, rr_depth: int = 3, features: int = 0,
         samples: float = 32e6, max_frames: int = 64, memory_bytes: int | None = None, tolerance: float = 0.03,
         modes=(hpt.LAUNCH_GRAPH, hpt.LAUNCH_EAGER), sorts=SORT_CANDIDATES, candidates=CANDIDATES,
         tails=TAIL_CANDIDATES, blocks=BLOCK_CANDIDATES, device: int = 0, in_place: bool = False) -> Tuning:
    """Measures every candidate batch size (then launch mode, then shading
    sort threshold, then tail bounce) on a handle that shares `template`'s
    committed scene, each over about `samples` paths (whole batches); returns
    the choice.  `in_place`: measure on `template` itself (it must have the
    given size, depths and features; its accumulation is reset, its batch
    size restored, and the caller sets the other settings it wants)."""
    t_start = time.perf_counter()
    if in_place:
        r, frames0 = template, template.frames_in_flight
    else:
        r = hpt.HalaRenderer(width, height, max_depth=max_depth, rr_depth=rr_depth, features=features, device=device)
    table = {}
    try:
        if not in_place:
            # Sized for the first candidate from the start (not the default batch).
            r.frames_in_flight = candidates_for(width, height, max_frames, memory_bytes, candidates)[0]
            r.share_scene(template)

        def measure(k, mode, sort=hpt.SHADE_SORT_OFF, tail=hpt.TAIL_FROM_NONE, block=128):
            r.frames_in_flight = k
            r.set_launch_mode(mode)
            r.set_shade_sort(sort)
            r.set_tail_from(tail)
            r.set_traversal_block(block)
            n = max(1, -(-int(samples) // (width * height * k))) * k
            r.reset_accumulation()
            r.render(k)  # warm-up: allocation and graph capture
            r.wait_idle()
            t0 = time.perf_counter()
            r.render(n)
            r.wait_idle()
            key = f"{k}/{'graph' if mode == hpt.LAUNCH_GRAPH else 'eager'}"
            if sort != hpt.SHADE_SORT_OFF:
                key += f"/s{sort}"
            if tail != hpt.TAIL_FROM_NONE:
                key += f"/t{tail}"
            if block != 128:
                key += f"/b{block}"
            table[key] = width * height * n / (time.perf_counter() - t0) / 1e6

        # Batch sizes with the first mode, then the other modes at the best size.
        for k in candidates_for(width, height, max_frames, memory_bytes, candidates):
            measure(k, modes[0])
        k_best = int(max(table, key=table.get).split("/")[0])
        for mode in modes[1:]:
            measure(k_best, mode)
        k_best, mode_best = max(table, key=table.get).split("/")[:2]
        for sort in sorts:
            measure(int(k_best), hpt.LAUNCH_GRAPH if mode_best == "graph" else hpt.LAUNCH_EAGER, sort)
        best = max(table.values())
        ok = []
        for key, v in table.items():
            parts = key.split("/")
            if v >= (1 - tolerance) * best and len(parts) <= 3 and not parts[-1].startswith(("t", "b")):
                sort = int(parts[2][1:]) if len(parts) > 2 else hpt.SHADE_SORT_OFF
                ok.append((int(parts[0]), parts[1], sort, v))
        k, mode, sort, v = min(ok, key=lambda t: (t[0], t[1] != "graph", t[2] != hpt.SHADE_SORT_OFF, -t[3]))
        mode = hpt.LAUNCH_GRAPH if mode == "graph" else hpt.LAUNCH_EAGER
        # The tail bounce at that choice: the fastest, none included.
        tail = hpt.TAIL_FROM_NONE
        for tf in tails:
            if tf > max_depth:
                continue
            before = set(table)
            measure(k, mode, sort, tf)
            (key,) = set(table) - before
            if table[key] > v * (1 + MARGIN):
                tail, v = tf, table[key]
        # The traversal block size at that choice.
        block = 128
        for bs in blocks:
            before = set(table)
            measure(k, mode, sort, tail, bs)
            (key,) = set(table) - before
            if table[key] > v * (1 + MARGIN):
                block, v = bs, table[key]
    finally:
        if in_place:
            r.frames_in_flight = frames0
            r.reset_accumulation()
        else:
            r.close()
    return Tuning(k, mode, v, table, time.perf_counter() - t_start, sort, tail, block)


class TuningCache:
    """(scene name, width, height, max_depth) -> Tuning."""

    def __init__(self):
        self.entries = {}

    @staticmethod
    def key(scene: str, width: int, height: int, max_depth: int) -> str:
        return f"{scene}|{width}x{height}|d{max_depth}"

    def get(self, scene: str, width: int, height: int, max_depth: int):
        return self.entries.get(self.key(scene, width, height, max_depth))

    def put(self, scene: str, width: int, height: int, max_depth: int, t: Tuning) -> None:
        self.entries[self.key(scene, width, height, max_depth)] = t

    def save(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump({k: vars(v) for k, v in self.entries.items()}, f, indent=1)

    @classmethod
    def load(cls, path: str) -> "TuningCache":
        c = cls()
        if os.path.exists(path):
            with open(path) as f:
                for k, v in json.load(f).items():
                    c.entries[k] = Tuning(**v)
        return c
