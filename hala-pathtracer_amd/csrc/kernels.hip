// kernels.hip — wavefront path tracing kernels for gfx950.
//
// One frame = camera -> (extend -> shade -> shadow) per bounce -> accumulate.
// The bounce loop of pathtracer.rgen.glsl:55-257 is split at its traceRayEXT
// calls: `extend` answers the closest-hit query for every live path, `shade`
// runs one loop iteration (surface/medium interaction, NEE sampling, BSDF
// sampling, Russian roulette) and appends the surviving paths and the shadow
// rays to compacted queues (wave ballot + one atomic per wave), `shadow`
// answers the any-hit queries and adds the unoccluded NEE terms.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_shading.h"
#include "dev_traverse.h"
#include "kernels.h"

namespace hpt {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Appends `val` for every lane with want == true; one atomic per wave.
__device__ __forceinline__ void wave_append(bool want, uint32_t* count, uint32_t* out, uint32_t val) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (want) {
    const uint64_t below = m & ((1ull << lane_id()) - 1ull);
    out[base + (uint32_t)__popcll(below)] = val;
  }
}

__device__ __forceinline__ uint32_t wave_reserve(bool want, uint32_t* count) {
  const uint64_t m = __ballot(want);
  if (m == 0) return 0;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
}

// Appends `val` to this block's segment of q.
__device__ __forceinline__ void seg_append(bool want, const Queue& q, uint32_t val) {
  const uint32_t s = blockIdx.x % kSeg;
  wave_append(want, q.counts + s * kSegStride, q.items + s * q.seg_cap, val);
}

// Shadow-queue position for lanes with want == true, in this block's segment.
__device__ __forceinline__ uint32_t seg_reserve(bool want, const ShadowQueue& sq) {
  const uint32_t s = blockIdx.x % kSeg;
  return s * sq.seg_cap + wave_reserve(want, sq.counts + s * kSegStride);
}

// This thread's entry in a launch over a segmented queue: entry j of
// segment seg, live while j < n (the segment's count).
struct SegSlot {
  uint32_t seg, j, n;
};
template <uint32_t BS = kBlock>
__device__ __forceinline__ SegSlot seg_slot(const uint32_t* counts) {
  SegSlot t;
  t.seg = blockIdx.x % kSeg;
  t.j = (blockIdx.x / kSeg) * BS + threadIdx.x;
  t.n = counts[t.seg * kSegStride];
  return t;
}

// First frame of the batch: the device counter of a graph replay, else the
// value passed at launch.
__device__ __forceinline__ uint32_t first_frame(const FrameParams& fp) {
  return fp.frame_ptr ? *fp.frame_ptr : fp.frame_index;
}

// ------------------------------------------------------------------ camera
// pathtracer.rgen.glsl:23-54: seed, camera callable, payload initialisation.
// A batch holds batch_frames consecutive frames, one path per pixel each.
__global__ void __launch_bounds__(kBlock) k_camera(SceneView sv, FrameParams fp, PathState ps, uint32_t npaths,
                                                   uint32_t job) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= npaths) return;
  if (ps.job) ps.job[p] = (uint16_t)job;
  const uint32_t j = p / fp.pixels, pix = p - j * fp.pixels;
  const uint32_t x = pix % fp.width, y = pix / fp.width;
  const uint32_t frame = first_frame(fp) + j;
  f3 o, d;
  camera_ray(sv, fp, frame, x, y, o, d);
  ps.rng[p] = make_uint4(x, y, frame, x + y);
  ps.org[p] = make_float4(o.x, o.y, o.z, 1.0f);
  ps.dir[p] = make_float4(d.x, d.y, d.z, __uint_as_float(0u));
  ps.thr[p] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(0u));
  ps.rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ps.albedo[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ps.normal[p] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
  ps.med[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ps.med2[p] = make_float2(1.0f, __uint_as_float(HPT_MEDIUM_NONE));
  ps.nee_env[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ps.nee_light[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Adds per-lane node/triangle counts to the stage counters: a wave sum, then
// one atomic per wave on this block's counter segment.
__device__ __forceinline__ void add_visits(unsigned long long* stats, uint32_t nv, uint32_t nt) {
  unsigned long long a = nv, b = nt;
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  if (lane_id() == 0) {
    const uint32_t s = blockIdx.x % kSeg;
    atomicAdd(stats + (kSeg + s) * kStatStride, a);
    atomicAdd(stats + (2 * kSeg + s) * kStatStride, b);
  }
}

// Path id of this thread for a launch over qin (identity when qin.items is
// null) and the ray count for the stage counters.
__device__ __forceinline__ bool queue_entry(const Queue& qin, uint32_t n_max, unsigned long long* stats,
                                            uint32_t& p) {
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  if (!qin.items) {
    if (stats && gid == 0) atomicAdd(stats, (unsigned long long)n_max);
    p = gid;
    return gid < n_max;
  }
  const SegSlot t = seg_slot(qin.counts);
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + t.seg * kStatStride, (unsigned long long)t.n);
  p = t.j < t.n ? qin.items[t.seg * qin.seg_cap + t.j] : 0u;
  return t.j < t.n;
}

// ------------------------------------------------------------------ extend
// Occupancy target of the traversal kernels (waves per SIMD): the walks are
// bound by memory latency, and 7 waves measured fastest (6 without the
// target; 8 forces register spills).
#ifndef HPT_WALK_WAVES_PER_EU
#define HPT_WALK_WAVES_PER_EU 7
#endif
// Extend and shadow lanes trace up to kRaysPerLane queue entries back to
// back (entries j, j + 64, ... of the wave's range), so a wave steps until
// its lanes' summed walks end instead of idling behind its longest ray:
// per-ray node visits range from 1 to tens.  A lane's next ray (origin,
// direction and, for shadow rays, everything its end needs) is loaded while
// the current one walks, so starting the next walk waits on no memory.
// Entry range of this lane for a launch over a queue: the identity queue's
// grid covers [0, n_max); a segmented queue's block b covers chunk b / kSeg
// of segment b % kSeg.
struct LaneRange {
  uint32_t j;     // first entry (segment-local)
  uint32_t n;     // entries in the segment (or n_max)
  uint32_t base;  // offset of the segment in the entry arrays
};
template <uint32_t BS>
__device__ __forceinline__ LaneRange lane_range(const uint32_t* counts, uint32_t seg_cap, bool identity,
                                                uint32_t n_max, unsigned long long* stats) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  constexpr uint32_t kWaveSpan = 64u * kRaysPerLane, kBlockSpan = BS * kRaysPerLane;
  LaneRange r;
  if (identity) {
    if (stats && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(stats, (unsigned long long)n_max);
    r.j = blockIdx.x * kBlockSpan + wave * kWaveSpan + lane;
    r.n = n_max;
    r.base = 0;
    return r;
  }
  const uint32_t seg = blockIdx.x % kSeg;
  r.n = counts[seg * kSegStride];
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + seg * kStatStride, (unsigned long long)r.n);
  r.j = (blockIdx.x / kSeg) * kBlockSpan + wave * kWaveSpan + lane;
  r.base = seg * seg_cap;
  return r;
}

// Shading sort key of a closest-hit record (see ShadeSort).
__device__ __forceinline__ uint8_t shade_key(const SceneView& sv, uint32_t id) {
  if (id == MISS_ID) return (uint8_t)(kSortKeys - 1);
  if (id & LIGHT_BIT) return (uint8_t)(kSortKeys - 2);
  return sv.tri_key[id];
}

template <bool COUNT, bool SORT, uint32_t BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(HPT_WALK_WAVES_PER_EU))) k_extend(SceneView sv, PathState ps, Queue qin, uint32_t n_max, Queue qnext,
                                                   ShadowQueue sq, Spill sp, unsigned long long* stats,
                                                   uint8_t* keys, uint32_t* phase_counts) {
  __shared__ uint2 lds_stack[kShortStack * BS];
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  if (gid < kSeg) {  // the shade launch after this one appends to these
    qnext.counts[gid * kSegStride] = 0;
    sq.counts[gid * kSegStride] = 0;
    if (phase_counts) {  // the phased volume-MIS shading's queues
      phase_counts[gid * kSegStride] = 0;
      phase_counts[(kSeg + gid) * kSegStride] = 0;
    }
  }
  const LaneRange lr = lane_range<BS>(qin.counts, qin.seg_cap, qin.items == nullptr, n_max, stats);
  const uint32_t* items = qin.items ? qin.items + lr.base : nullptr;
  auto entry = [&](uint32_t j) { return items ? items[j] : j; };
  LaneStack st{lds_stack + threadIdx.x, BS, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
  uint32_t cv = 0, ct = 0;
  auto store = [&](uint32_t p, uint32_t j, const TraceResult& r) {
    ps.hit[p] = make_float4(r.t, r.u, r.v, __uint_as_float(r.id));
    if (SORT) keys[lr.base + j] = shade_key(sv, r.id);
  };
  if (sv.node_count == 0) {  // lights only
    for (uint32_t i = 0, j = lr.j; i < kRaysPerLane && j < lr.n; ++i, j += 64) {
      const uint32_t p = entry(j);
      const float4 o4 = ps.org[p], d4 = ps.dir[p];
      store(p, j, trace_closest(sv, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z), 0.0f, WORLD_SIZE, st));
    }
    return;
  }
  uint32_t j0 = lr.j, p0 = 0u;
  bool act0 = j0 < lr.n;
  if (act0) p0 = entry(j0);
  // Pipeline: current ray (path p), next ray (path pn, origin/direction
  // loaded), the one after (path pnn, id loaded).
  uint32_t i = 0, j = j0;
  bool active = act0;
  uint32_t p = p0;
  bool has_n = active && kRaysPerLane > 1 && j + 64 < lr.n;
  uint32_t pn = has_n ? entry(j + 64) : 0u;
  bool has_nn = has_n && kRaysPerLane > 2 && j + 128 < lr.n;
  uint32_t pnn = has_nn ? entry(j + 128) : 0u;
  Walker<false> w;
  if (active) {
    const float4 o4 = ps.org[p], d4 = ps.dir[p];
    w.begin(prepare_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z)), 0.0f, WORLD_SIZE, st);
  }
  float4 on4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), dn4 = on4;
  if (has_n) {
    on4 = ps.org[pn];
    dn4 = ps.dir[pn];
  }
  while (__ballot(active) != 0) {
    if (active && w.step(sv, st)) {
      TraceResult r{0.0f, 0.0f, 0.0f, MISS_ID};
      if (w.found) r = TraceResult{w.best_t, w.bu, w.bv, w.best_id};
      closest_lights(sv, w.r.org, w.r.dir, 0.0f, WORLD_SIZE, w.found, r);
      store(p, j, r);
      if (COUNT) {
        cv += w.nv;
        ct += w.nt;
      }
      active = has_n;
      if (active) {
        ++i;
        j += 64;
        p = pn;
        w.begin(prepare_ray(mk3(on4.x, on4.y, on4.z), mk3(dn4.x, dn4.y, dn4.z)), 0.0f, WORLD_SIZE, st);
        has_n = has_nn;
        pn = pnn;
        if (has_n) {
          on4 = ps.org[pn];
          dn4 = ps.dir[pn];
        }
        has_nn = has_n && i + 2 < kRaysPerLane && j + 128 < lr.n;
        if (has_nn) pnn = entry(j + 128);
      }
    }
  }
  if (COUNT) add_visits(stats, cv, ct);
}

// Idle lanes a wave collects before it takes new queue entries.
#ifndef HPT_REFILL_IDLE
#define HPT_REFILL_IDLE 32
#endif

// Extension over a compacted queue by resident blocks that refill their
// lanes.  One ray per lane ties a wave to its longest ray (bounce-1 rays on
// the 1080p icosphere: 7.5 node visits on average, 16 at the 90th percentile,
// up to 108), so most lanes of a wave idle for most of its steps.  Here a wave
// keeps stepping and, once HPT_REFILL_IDLE of its lanes have ended, those
// lanes store their hits and take the next entries of the block's segment in
// one atomic, so every node step carries more live lanes.  Block b starts on
// segment b % kSeg and moves on to the other segments of its XCD (b % 8)
// when that one is drained.  Each entry is traced exactly as by k_extend and
// its results go to the same places.  cursors[s * kSegStride] counts the
// entries of segment s handed out, cursors[kSeg * kSegStride] the finished
// blocks; the last block clears them all for the next launch.
template <bool COUNT, bool SORT, uint32_t BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(HPT_WALK_WAVES_PER_EU)))
    k_extend_refill(SceneView sv, PathState ps, Queue qin, uint32_t n_max, Queue qnext, ShadowQueue sq, Spill sp,
                    unsigned long long* stats, uint8_t* keys, uint32_t* phase_counts, uint32_t* cursors) {
  __shared__ uint2 lds_stack[kShortStack * BS];
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  if (gid < kSeg) {  // the shade launch after this one appends to these
    qnext.counts[gid * kSegStride] = 0;
    sq.counts[gid * kSegStride] = 0;
    if (phase_counts) {
      phase_counts[gid * kSegStride] = 0;
      phase_counts[(kSeg + gid) * kSegStride] = 0;
    }
  }
  // The first bounce's identity queue over [0, n_max): segment s holds the
  // kBlock-path chunks c with c % kSeg == s, in order -- the chunks the
  // one-ray-per-lane launch gives block c, so each XCD keeps the same
  // interleaved share of the image.
  const uint32_t nchunks = (n_max + kBlock - 1) / kBlock;
  auto seg_count = [&](uint32_t sg) -> uint32_t {
    if (qin.items) return qin.counts[sg * kSegStride];
    if (sg >= nchunks) return 0u;
    uint32_t c = ((nchunks - 1u - sg) / kSeg + 1u) * kBlock;
    if ((nchunks - 1u) % kSeg == sg) c -= nchunks * kBlock - n_max;
    return c;
  };
  uint32_t seg = blockIdx.x % kSeg;
  uint32_t segn = seg_count(seg);
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + seg * kStatStride, (unsigned long long)segn);
  LaneStack st{lds_stack + threadIdx.x, BS, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
  const uint32_t lane = lane_id();
  uint32_t cv = 0, ct = 0, hops = 0;
  uint32_t p = 0, k = 0;
  bool more = true, active = false, ended = false;
  Walker<false> w;
  auto finish = [&] {
    TraceResult r{0.0f, 0.0f, 0.0f, MISS_ID};
    if (w.found) r = TraceResult{w.best_t, w.bu, w.bv, w.best_id};
    // The direction is read again rather than kept through the walk, which
    // needs only its reciprocal and the triangle test's shear.
    const float4 d4 = ps.dir[p];
    closest_lights(sv, w.r.org, mk3(d4.x, d4.y, d4.z), 0.0f, WORLD_SIZE, w.found, r);
    ps.hit[p] = make_float4(r.t, r.u, r.v, __uint_as_float(r.id));
    if (SORT) keys[k] = shade_key(sv, r.id);
    if (COUNT) {
      cv += w.nv;
      ct += w.nt;
    }
  };
  while (true) {
    const unsigned long long idle = __ballot(!active);
    if (more && __popcll(idle) >= HPT_REFILL_IDLE) {
      if (ended) {
        finish();
        ended = false;
      }
      unsigned long long need = idle;
      while (need != 0ull && more) {
        const uint32_t cnt = (uint32_t)__popcll(need);
        const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
        uint32_t got = 0;
        if (lane == leader) got = atomicAdd(cursors + seg * kSegStride, cnt);
        got = (uint32_t)__builtin_amdgcn_readlane((int)got, (int)leader);
        const uint32_t avail = got < segn ? min(cnt, segn - got) : 0u;
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        const bool take = ((need >> lane) & 1ull) != 0ull && rank < avail;
        if (take) {
          const uint32_t e = got + rank;
          k = qin.items ? seg * qin.seg_cap + e : ((e / kBlock) * kSeg + seg) * kBlock + e % kBlock;
          p = qin.items ? qin.items[k] : k;
          const float4 o4 = ps.org[p], d4 = ps.dir[p];
          w.begin(prepare_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z)), 0.0f, WORLD_SIZE, st);
          active = true;
        }
        need &= ~__ballot(take);
        if (avail < cnt) {  // segment drained: the next one of this XCD
          if (++hops == kSeg / 8) {
            more = false;
          } else {
            seg = (seg + 8) % kSeg;
            segn = seg_count(seg);
          }
        }
      }
    }
    if (__ballot(active) == 0ull) {
      if (!more) break;
      continue;
    }
    if (active && w.step(sv, st)) {
      active = false;
      ended = true;
    }
  }
  if (ended) finish();
  if (COUNT) add_visits(stats, cv, ct);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(cursors + kSeg * kSegStride, 1u) == gridDim.x - 1) {
      for (uint32_t s = 0; s < kSeg; ++s) cursors[s * kSegStride] = 0;
      cursors[kSeg * kSegStride] = 0;
      __threadfence();
    }
  }
}

// ------------------------------------------------------------ shading sort
// Counting sort of each queue segment by the keys extend wrote: per-block
// key histograms in LDS summed into per-segment counts, a scan of each
// segment's kSortKeys counts into cursors, then a scatter that reserves one
// range per (block, key).  Order within a key is not kept (shading results
// do not depend on it).

// Entries of segment s of the sort's input: the segment count of a compacted
// queue, or the slice [s * seg_cap, (s + 1) * seg_cap) of the first bounce's
// identity queue over [0, n_max).
__device__ __forceinline__ uint32_t sort_seg_count(const Queue& qin, uint32_t n_max, uint32_t s) {
  if (qin.items) return qin.counts[s * kSegStride];
  const uint32_t lo = s * qin.seg_cap;
  return n_max > lo ? min(qin.seg_cap, n_max - lo) : 0u;
}

__global__ void __launch_bounds__(kBlock) k_sort_count(Queue qin, uint32_t n_max, ShadeSort ss) {
  __shared__ uint32_t h[kSortKeys];
  const uint32_t s = blockIdx.x % kSeg, j0 = (blockIdx.x / kSeg) * kBlock, j = j0 + threadIdx.x;
  const uint32_t n = sort_seg_count(qin, n_max, s);
  if (j0 >= n) return;  // the whole block (n and j0 are block-uniform)
  if (threadIdx.x < kSortKeys) h[threadIdx.x] = 0u;
  __syncthreads();
  if (j < n) atomicAdd(&h[ss.keys[s * qin.seg_cap + j]], 1u);
  __syncthreads();
  if (threadIdx.x < kSortKeys && h[threadIdx.x])
    atomicAdd(&ss.bucket[s * kSortKeys + threadIdx.x], h[threadIdx.x]);
}

// One wave per segment, one lane per key: exclusive prefix of the counts.
__global__ void __launch_bounds__(kSortKeys) k_sort_scan(ShadeSort ss, uint32_t seg_cap) {
  static_assert(kSortKeys == 64, "one key per lane of a wave");
  const uint32_t s = blockIdx.x, k = threadIdx.x;
  uint32_t* cnt = ss.bucket + s * kSortKeys;
  uint32_t* cur = ss.bucket + (kSeg + s) * kSortKeys;
  const uint32_t c = cnt[k];
  uint32_t x = c;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if ((int)k >= off) x += y;
  }
  cur[k] = s * seg_cap + (x - c);
  cnt[k] = 0u;  // the next sort counts from zero
  if (k == kSortKeys - 1) ss.counts[s * kSegStride] = x;
}

__global__ void __launch_bounds__(kBlock) k_sort_scatter(Queue qin, uint32_t n_max, ShadeSort ss) {
  __shared__ uint32_t h[kSortKeys], base[kSortKeys];
  const uint32_t s = blockIdx.x % kSeg, j0 = (blockIdx.x / kSeg) * kBlock, j = j0 + threadIdx.x;
  const uint32_t n = sort_seg_count(qin, n_max, s);
  if (j0 >= n) return;
  if (threadIdx.x < kSortKeys) h[threadIdx.x] = 0u;
  __syncthreads();
  const bool live = j < n;
  const uint32_t slot = s * qin.seg_cap + j;
  uint32_t key = 0u, rank = 0u, item = 0u;
  if (live) {
    key = ss.keys[slot];
    item = qin.items ? qin.items[slot] : slot;
    rank = atomicAdd(&h[key], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kSortKeys && h[threadIdx.x])
    base[threadIdx.x] = atomicAdd(&ss.bucket[(kSeg + s) * kSortKeys + threadIdx.x], h[threadIdx.x]);
  __syncthreads();
  if (live) ss.items[base[key] + rank] = item;
}

// ------------------------------------------------------------------ shading
struct HitState {  // the State fields triangles.rchit.glsl:66-75 writes
  f3 pos, normal, ffnormal, tangent, bitangent;
  uint32_t material;
};

// triangles.rchit.glsl:13-76.
__device__ __forceinline__ HitState triangle_state(const SceneView& sv, uint32_t id, float t, float u, float v,
                                                   f3 org, f3 dir) {
  HitState h;
  const float4* rec = reinterpret_cast<const float4*>(sv.tri_shade + id);
  const float4 q0 = rec[0], q1 = rec[1], q2 = rec[2], q3 = rec[3], q4 = rec[4], q6 = rec[6];
  const PrimInfo& pi = sv.prims[__float_as_uint(q4.z)];
  const f3 na = mk3(q0.x, q0.y, q0.z), nb = mk3(q0.w, q1.x, q1.y), nc = mk3(q1.z, q1.w, q2.x);
  const f3 ta = mk3(q2.y, q2.z, q2.w), tb = mk3(q3.x, q3.y, q3.z), tc = mk3(q3.w, q4.x, q4.y);
  const float bx = 1.0f - u - v, by = u, bz = v;
  const f3 c0 = mk3(pi.it[0], pi.it[1], pi.it[2]), c1 = mk3(pi.it[3], pi.it[4], pi.it[5]),
           c2 = mk3(pi.it[6], pi.it[7], pi.it[8]);
  f3 n = na * bx + nb * by + nc * bz;
  n = normalize(c0 * n.x + c1 * n.y + c2 * n.z);
  f3 tg = ta * bx + tb * by + tc * bz;
  tg = normalize(c0 * tg.x + c1 * tg.y + c2 * tg.z);
  h.normal = n;
  h.tangent = tg;
  h.bitangent = cross(n, tg);
  h.pos = org + dir * t;
  h.ffnormal = dot(n, dir) <= 0.0f ? n : -n;
  h.material = __float_as_uint(q6.y);  // TriShade::material
  return h;
}

// Texture coordinate of a triangle hit (triangles.rchit.glsl, tex_coord):
// the vertex uvs of the triangle's shading record at the barycentrics.
__device__ __forceinline__ void hit_uv(const SceneView& sv, uint32_t id, float u, float v, float& tu, float& tv) {
  const float4* rec = reinterpret_cast<const float4*>(sv.tri_shade + id);
  const float4 q4 = rec[4], q5 = rec[5], q6 = rec[6];
  const float bx = 1.0f - u - v, by = u, bz = v;
  tu = q4.w * bx + q5.y * by + q5.w * bz;
  tv = q5.x * bx + q5.z * by + q6.x * bz;
}

// The material maps of pathtracer.rgen.glsl:122-146: base colour,
// metallic (blue) and roughness (green), emission, and the normal map
// (z scaled by 5 before the [0,1] -> [-1,1] remap; the bitangent follows the
// new normal, the tangent is kept).
__device__ __forceinline__ void apply_textures(const SceneView& sv, uint32_t id, float u, float v, f3 dir, Mat& m,
                                               HitState& h) {
  float tu, tv;
  hit_uv(sv, id, u, v, tu, tv);
  if (m.base_color_map != INVALID) {
    const float4 c = tex_sample(sv, m.base_color_map, tu, tv, true);
    m.base_color = mk3(c.x, c.y, c.z);
  }
  if (m.metallic_roughness_map != INVALID) {
    const float4 mr = tex_sample(sv, m.metallic_roughness_map, tu, tv, false);
    m.metallic = mr.z;
    m.roughness = mr.y;
  }
  if (m.emission_map != INVALID) {
    const float4 c = tex_sample(sv, m.emission_map, tu, tv, true);
    m.emission = mk3(c.x, c.y, c.z);
  }
  if (m.normal_map != INVALID) {
    const float4 c = tex_sample(sv, m.normal_map, tu, tv, false);
    f3 nts = mk3(c.x, c.y, c.z * 5.0f);
    nts = normalize(nts * 2.0f - 1.0f);
    const f3 original = h.normal;
    h.normal = normalize(h.tangent * nts.x + h.bitangent * nts.y + h.normal * nts.z);
    h.ffnormal = dot(original, dir) <= 0.0f ? h.normal : -h.normal;
    h.bitangent = cross(h.normal, h.tangent);
  }
}

// inc/lighting.glsl:7-61 (MEDIUM + VOL_MIS): closest-hit steps through
// transparent or refractive surfaces, Beer-Lambert on exit faces.
// Closest-hit queries, node visits and triangle tests of one lane's walks
// (counted by the profiling variant of the shading kernel).
struct WalkCounts {
  uint32_t q = 0, nv = 0, nt = 0;
};

// Adds per-lane walk counts to the stage counters (node visits, triangle
// tests, queries), one atomic per wave and counter.
__device__ __forceinline__ void add_walks(unsigned long long* stats, const WalkCounts& wc) {
  unsigned long long a = wc.nv, b = wc.nt, c = wc.q;
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
    c += __shfl_xor(c, off);
  }
  if (lane_id() == 0) {
    const uint32_t s = blockIdx.x % kSeg;
    atomicAdd(stats + (kSeg + s) * kStatStride, a);
    atomicAdd(stats + (2 * kSeg + s) * kStatStride, b);
    atomicAdd(stats + (3 * kSeg + s) * kStatStride, c);
  }
}

template <bool TRANSP, bool COUNT>
__device__ f3 eval_transmittance(const SceneView& sv, uint32_t max_depth, uint4& rng, f3 o, f3 d, LaneStack& st,
                                 WalkCounts& wc) {
  f3 tr = mk3(1.0f);
  for (uint32_t depth = 0; depth < max_depth; depth++) {
    uint32_t nv = 0, nt = 0;
    const TraceResult r = trace_closest(sv, o, d, 0.0f, WORLD_SIZE, st, COUNT ? &nv : nullptr, COUNT ? &nt : nullptr);
    if (COUNT) {
      wc.q++;
      wc.nv += nv;
      wc.nt += nt;
    }
    if (r.id == MISS_ID || (r.id & LIGHT_BIT)) break;
    const HitState h = triangle_state(sv, r.id, r.t, r.u, r.v, o, d);
    if (h.material == INVALID) return mk3(0.0f);
    const Mat m = load_material(sv.materials + h.material);
    const bool is_transparent = TRANSP ? (m.opacity == 0.0f || rnd(rng) > m.opacity) : false;
    const bool is_refractive = (1.0f - m.metallic) * m.specular_transmission > 0.0f;
    if (!(is_transparent || is_refractive)) return mk3(0.0f);
    if (dot(d, h.normal) > 0.0f && m.medium_type != HPT_MEDIUM_NONE) {
      const f3 color = m.medium_type == HPT_MEDIUM_ABSORB ? mk3(1.0f) - m.medium_color : mk3(1.0f);
      const f3 e = -color * m.medium_density * r.t;
      tr *= mk3(expf(e.x), expf(e.y), expf(e.z));
    }
    o = h.pos + d * EPS;
  }
  return tr;
}

struct ShadowOut {
  bool valid;
  f3 o, d, c;
  float tmax;
};

// inc/lighting.glsl:69-244.  Without VOL_MIS the shadow queries are deferred:
// each unoccluded-candidate term is returned with its ray.  With VOL_MIS the
// transmittance walk runs inline (it draws from the path RNG between the
// environment and light samples) and the sum is returned directly.
template <bool MED, bool VMIS, bool TRANSP, bool COUNT>
__device__ f3 direct_lighting(const SceneView& sv, const FrameParams& fp, uint32_t max_depth, uint4& rng,
                              const Mat& mat, const Frame& fr, f3 hit_pos, f3 ffn, f3 V, float aniso,
                              bool is_surface, bool anb, f3 weight, ShadowOut& s_env, ShadowOut& s_light,
                              LaneStack& st, WalkCounts& wc) {
  f3 Ld = mk3(0.0f);
  const f3 scatter_pos = hit_pos + ffn * EPS;
  s_env.valid = false;
  s_light.valid = false;
  if (fp.env_type != HPT_ENV_SKY) {
    f3 Li;
    float light_pdf;
    const f3 light_dir = sample_env_map(rng, sv, fp, Li, light_pdf);
    if (light_pdf > 0.0f) {
      if (VMIS) {
        const f3 trm = eval_transmittance<TRANSP, COUNT>(sv, max_depth, rng, scatter_pos, light_dir, st, wc);
        if (luminance(trm) > 0.0f) {
          Li *= trm;
          float pdf = 0.0f;
          f3 f;
          if (is_surface) {
            f = bxdf_eval(anb, fr, mat, V, ffn, light_dir, pdf);
          } else {
            pdf = phase_hg(dot(V, light_dir), aniso);
            f = mk3(pdf);
          }
          if (pdf > 0.0f) {
            const float mis = power_heuristic(light_pdf, pdf);
            if (mis > 0.0f) Ld += mis * Li * f / light_pdf;
          }
        }
      } else {
        float pdf;
        const f3 f = bxdf_eval(anb, fr, mat, V, ffn, light_dir, pdf);
        if (pdf > 0.0f) {
          const float mis = power_heuristic(light_pdf, pdf);
          if (mis > 0.0f) {
            s_env.valid = true;
            s_env.o = scatter_pos;
            s_env.d = light_dir;
            s_env.tmax = WORLD_SIZE;
            s_env.c = weight * (mis * Li * f / light_pdf);
          }
        }
      }
    }
  }
  if (fp.num_lights > 0) {
    // rand() may return 1.0 and pick index num_lights: the zero-filled slot
    // after the last light, as in the reference's lights buffer.
    const int li = (int)(rnd(rng) * (float)fp.num_lights);
    const hpt_light& L = sv.lights[li];
    LightSample ls;
    light_sample(rng, fp.num_lights, hit_pos, L, ls);
    const f3 light_dir = ls.direction;
    const float light_pdf = ls.pdf;
    f3 Li = ls.emission;
    if (light_pdf > 0.0f && dot(light_dir, ls.normal) < 0.0f) {
      if (VMIS) {
        const f3 trm = eval_transmittance<TRANSP, COUNT>(sv, max_depth, rng, scatter_pos, light_dir, st, wc);
        if (luminance(trm) > 0.0f) {
          Li *= trm;
          float pdf = 0.0f;
          f3 f;
          if (is_surface) {
            f = bxdf_eval(anb, fr, mat, V, ffn, light_dir, pdf);
          } else {
            pdf = phase_hg(dot(V, light_dir), aniso);
            f = mk3(pdf);
          }
          if (pdf > 0.0f) {
            float mis = 1.0f;
            if (L.area > 0.0f) mis = power_heuristic(light_pdf, pdf);
            if (mis > 0.0f) Ld += mis * Li * f / light_pdf;
          }
        }
      } else {
        float pdf;
        const f3 f = bxdf_eval(anb, fr, mat, V, ffn, light_dir, pdf);
        if (pdf > 0.0f) {
          float mis = 1.0f;
          if (L.area > 0.0f) mis = power_heuristic(light_pdf, pdf);
          if (mis > 0.0f) {
            s_light.valid = true;
            s_light.o = scatter_pos;
            s_light.d = light_dir;
            s_light.tmax = ls.dist - 2.0f * EPS;
            s_light.c = weight * (mis * Li * f / light_pdf);
          }
        }
      }
    }
  }
  return Ld;
}

// One iteration of the bounce loop, pathtracer.rgen.glsl:70-256.  The
// occupancy target trades register spills for latency hiding: 3 waves per
// SIMD for the plain permutations, 5 for VOL_MIS, whose inline transmittance
// walks are latency-bound traversals (measured on the 1080p icosphere scene
// with a medium: 3 -> 577, 4 -> 626, 5 -> 633, 6 -> 618, 8 -> 556
// Msamples/s unsorted; sorted by material from bounce 1: 4 -> 722, 5 -> 727,
// 6 -> 723).
#ifndef HPT_SHADE_WAVES_PER_EU
#define HPT_SHADE_WAVES_PER_EU 3
#endif
#ifndef HPT_SHADE_VMIS_WAVES_PER_EU
#define HPT_SHADE_VMIS_WAVES_PER_EU 5
#endif
// The bounce of path p from its closest-hit record ps.hit[p]: radiance,
// AOVs and the next segment written to ps; returns whether the path lives
// on.  Without VOL_MIS the shadow rays come back in s_env / s_light (valid
// set when one was made) for the caller to trace; with VOL_MIS the
// transmittance walks run here (st: the walks' traversal stack, wc: their
// counts with COUNT).
template <bool MED, bool VMIS, bool TRANSP, bool JOBS, bool COUNT>
__device__ __forceinline__ bool shade_path(const SceneView& sv, const FrameParams& fp, const PathState& ps,
                                           uint32_t p, LaneStack& st, WalkCounts& wc, ShadowOut& s_env,
                                           ShadowOut& s_light) {
  bool cont = false;
  s_env.valid = false;
  s_light.valid = false;
  {
    const float4 o4 = ps.org[p], d4 = ps.dir[p], t4 = ps.thr[p], h4 = ps.hit[p];
    f3 org = mk3(o4.x, o4.y, o4.z), dir = mk3(d4.x, d4.y, d4.z), T = mk3(t4.x, t4.y, t4.z);
    float pdf = o4.w;
    uint32_t depth = __float_as_uint(d4.w);
    // Depth limits: the launch's, or in a wavefront of several jobs (JOBS)
    // those of the path's job.
    uint32_t max_depth = fp.max_depth, rr_depth = fp.rr_depth;
    if (JOBS) {
      const uint2 lim = fp.job_limits[ps.job[p]];
      max_depth = lim.x;
      rr_depth = lim.y;
    }
    uint32_t flags = __float_as_uint(t4.w);
    bool anb = (flags & PATH_ANB) != 0;
    bool in_medium = (flags & PATH_IN_MEDIUM) != 0;
    bool surface_scatter = (flags & PATH_SURFACE_SCATTER) != 0;
    uint4 rng = ps.rng[p];
    const float4 r4 = ps.rad[p];
    f3 radiance = mk3(r4.x, r4.y, r4.z);
    f3 med_color = mk3(0.0f);
    float med_density = 0.0f, med_aniso = 1.0f;
    uint32_t med_type = HPT_MEDIUM_NONE;
    if (MED) {
      const float4 m4 = ps.med[p];
      const float2 m2 = ps.med2[p];
      med_color = mk3(m4.x, m4.y, m4.z);
      med_density = m4.w;
      med_aniso = m2.x;
      med_type = __float_as_uint(m2.y);
    }
    const uint32_t id = __float_as_uint(h4.w);
    const float t_hit = h4.x;
    constexpr bool kScatterFlag = MED && !VMIS;

    if (id == MISS_ID) {
      const f3 env = env_eval(sv, fp, depth, pdf, dir, kScatterFlag && !surface_scatter);
      radiance += T * env;
      if (depth == 0) {
        ps.albedo[p] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
        ps.normal[p] = make_float4(-dir.x, -dir.y, -dir.z, 0.0f);
      }
    } else if (id & LIGHT_BIT) {
      const hpt_light& L = sv.lights[id & ~LIGHT_BIT];
      float light_pdf = 1.0f;
      light_isect(L, org, dir, light_pdf);
      float mis = 1.0f;
      if (depth > 0) mis = power_heuristic(pdf, light_pdf);
      if (kScatterFlag && !surface_scatter) mis = 1.0f;
      radiance += mis * ld3(L.intensity) * T;
    } else {
      HitState h = triangle_state(sv, id, t_hit, h4.y, h4.z, org, dir);
      if (h.material == INVALID) {
        radiance = mk3(0.7f, 0.0f, 0.7f);
        ps.nee_env[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ps.nee_light[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      } else {
        Mat m = load_material(sv.materials + h.material);
        Frame fr;
        fr.eta = dot(dir, h.normal) < 0.0f ? (1.0f / m.ior) : m.ior;  // before the normal map
        if ((m.base_color_map & m.metallic_roughness_map & m.emission_map & m.normal_map) != INVALID)
          apply_textures(sv, id, h4.y, h4.z, dir, m, h);
        fr.tangent = h.tangent;
        fr.bitangent = h.bitangent;
        fr.depth = depth;
        if (depth == 0) {
          ps.albedo[p] = make_float4(m.base_color.x, m.base_color.y, m.base_color.z, 0.0f);
          ps.normal[p] = make_float4(h.normal.x, h.normal.y, h.normal.z, 0.0f);
        }
        radiance += T * m.emission;
        bool alive = true;
        bool medium_sampled = false;
        if (MED) {
          if (!VMIS) surface_scatter = false;
          if (in_medium) {
            if (med_type == HPT_MEDIUM_ABSORB) {
              const f3 e = -(1.0f - med_color) * t_hit * med_density;
              T *= mk3(expf(e.x), expf(e.y), expf(e.z));
            } else if (med_type == HPT_MEDIUM_EMISSIVE) {
              radiance += med_color * t_hit * med_density * T;
            } else {
              const float scatter_dist = fminf(-logf(rnd(rng)) / med_density, t_hit);
              medium_sampled = scatter_dist < t_hit;
              if (medium_sampled) {
                T *= med_color;
                org = org + dir * scatter_dist;
                h.pos = org;
                const f3 Ld = direct_lighting<MED, VMIS, TRANSP, COUNT>(sv, fp, max_depth, rng, m, fr, h.pos, h.ffnormal, -dir,
                                                                  med_aniso, false, anb, T, s_env, s_light, st, wc);
                radiance += Ld * T;
                const float r1 = rnd(rng);
                const float r2 = rnd(rng);
                const f3 sdir = sample_hg(-dir, med_aniso, r1, r2);
                pdf = phase_hg(dot(-dir, sdir), med_aniso);
                dir = sdir;
              }
            }
          }
        }
        if (!medium_sampled) {
          bool pass = false;
          if (TRANSP) {
            if (m.opacity == 0.0f || rnd(rng) > m.opacity) {
              depth--;  // uint: wraps at depth 0 and the max_depth test ends the path
              pass = true;
            }
          }
          if (!pass) {
            if (kScatterFlag) surface_scatter = true;
            const f3 Ld = direct_lighting<MED, VMIS, TRANSP, COUNT>(sv, fp, max_depth, rng, m, fr, h.pos, h.ffnormal, -dir, med_aniso,
                                                              true, anb, T, s_env, s_light, st, wc);
            radiance += T * Ld;
            f3 L;
            float spdf;
            uint32_t sflags = 0;
            const f3 f = bxdf_sample(rng, anb, fr, m, -dir, h.ffnormal, L, spdf, sflags);
            dir = L;
            pdf = spdf;
            if (pdf > 0.0f) T *= f / pdf;
            else alive = false;
          }
          if (alive) {
            org = h.pos + dir * EPS;
            if (MED) {
              if (dot(dir, h.normal) < 0.0f && m.medium_type != HPT_MEDIUM_NONE) {
                in_medium = true;
                med_color = m.medium_color;
                med_density = m.medium_density;
                med_aniso = m.medium_anisotropy;
                med_type = m.medium_type;
              } else {
                in_medium = false;
                med_type = HPT_MEDIUM_NONE;
              }
            }
          }
        }
        if (alive && depth >= max_depth) alive = false;
        if (alive && depth >= rr_depth && rr_depth > 0) {
          const float q = fminf(fmaxf(T.x, fmaxf(T.y, T.z)) + 0.001f, 0.95f);
          if (rnd(rng) > q) alive = false;
          else T /= q;
        }
        if (alive) {
          cont = true;
          depth++;
          const uint32_t nf = (anb ? PATH_ANB : 0u) | (in_medium ? PATH_IN_MEDIUM : 0u) |
                              (surface_scatter ? PATH_SURFACE_SCATTER : 0u);
          ps.org[p] = make_float4(org.x, org.y, org.z, pdf);
          ps.dir[p] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(depth));
          ps.thr[p] = make_float4(T.x, T.y, T.z, __uint_as_float(nf));
          if (MED) {
            ps.med[p] = make_float4(med_color.x, med_color.y, med_color.z, med_density);
            ps.med2[p] = make_float2(med_aniso, __uint_as_float(med_type));
          }
        }
        ps.rng[p] = rng;
      }
    }
    ps.rad[p] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
  }
  return cont;
}

// stats: null, or the shading stage's counters (shading events as rays; with
// COUNT the volume-MIS walks' queries, node visits and triangle tests).
template <bool MED, bool VMIS, bool TRANSP, bool JOBS, bool COUNT>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(VMIS ? HPT_SHADE_VMIS_WAVES_PER_EU : HPT_SHADE_WAVES_PER_EU)))
    k_shade(SceneView sv, FrameParams fp, PathState ps, Queue qin, uint32_t n_max, Queue qout, ShadowQueue sq,
            Spill sp, unsigned long long* stats) {
  __shared__ uint2 lds_stack[VMIS ? kShortStack * kBlock : 1];
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  uint32_t p;
  const bool active = queue_entry(qin, n_max, stats, p);
  WalkCounts wc;
  bool cont = false;
  ShadowOut s_env, s_light;
  s_env.valid = false;
  s_light.valid = false;
  if (active) {
    LaneStack st{lds_stack + (VMIS ? threadIdx.x : 0), kBlock, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
    cont = shade_path<MED, VMIS, TRANSP, JOBS, COUNT>(sv, fp, ps, p, st, wc, s_env, s_light);
  }
  if (COUNT) add_walks(stats, wc);
  // Compact the survivors and the shadow rays.
  seg_append(cont, qout, p);
  if (!VMIS) {
    const bool w0 = s_env.valid, w1 = s_light.valid;
    const uint32_t i0 = seg_reserve(w0, sq);
    const uint32_t i1 = seg_reserve(w1, sq);
    if (w0) {
      sq.o[i0] = make_float4(s_env.o.x, s_env.o.y, s_env.o.z, s_env.tmax);
      sq.d[i0] = make_float4(s_env.d.x, s_env.d.y, s_env.d.z, __uint_as_float(p << 1));
      sq.c[i0] = make_float4(s_env.c.x, s_env.c.y, s_env.c.z, 0.0f);
    }
    if (w1) {
      sq.o[i1] = make_float4(s_light.o.x, s_light.o.y, s_light.o.z, s_light.tmax);
      sq.d[i1] = make_float4(s_light.d.x, s_light.d.y, s_light.d.z, __uint_as_float((p << 1) | 1u));
      sq.c[i1] = make_float4(s_light.c.x, s_light.c.y, s_light.c.z, 0.0f);
    }
  }
}

// ------------------------------------------------------------------- tail
// Pass-throughs of transparent surfaces do not count toward a path's depth
// (pathtracer.rgen.glsl:196-200), so with TRANSPARENT paths can outlive the
// max_depth + 1 bounces of the wavefront.  The tail takes the paths still in
// the queue after those bounces and runs each to its end on one lane --
// closest hit (as extend), the bounce (shade_path), the shadow rays (as
// shadow) -- with the same operations in the same order as the wavefront,
// so the images are the same; no host round trip decides how many bounces
// are left, which lets the whole batch be captured in one graph.  The tail
// can also start earlier, at any bounce: late bounces hold few paths, and a
// wavefront launch over them lasts as long as its slowest path's chain of
// dependent loads, where here each lane pays only its own path's chain.
// The grid strides over the queue: block b takes chunks b / kSeg,
// b / kSeg + G, ... (G = gridDim.x / kSeg) of segment b % kSeg.  (Handing
// out paths per lane through per-segment cursors, so a lane whose path ends
// takes the next at once, measured slower: 18.7 against 16.5 ms for two
// 2-frame 1080p batches.)  A path is cut after max_steps tail bounces; cut
// paths keep the radiance gathered so far and are counted in *cut.
// One bounce of a tail path: closest hit, shade_path, and without VOL_MIS
// the two shadow rays; returns whether the path continues.
template <bool MED, bool VMIS, bool TRANSP, bool JOBS>
__device__ __forceinline__ bool tail_bounce(const SceneView& sv, const FrameParams& fp, const PathState& ps,
                                            uint32_t p, LaneStack& st, WalkCounts& wc) {
  const float4 o4 = ps.org[p], d4 = ps.dir[p];
  const TraceResult r = trace_closest(sv, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z), 0.0f, WORLD_SIZE, st);
  ps.hit[p] = make_float4(r.t, r.u, r.v, __uint_as_float(r.id));
  wc.q++;
  ShadowOut s_env, s_light;
  const bool cont = shade_path<MED, VMIS, TRANSP, JOBS, false>(sv, fp, ps, p, st, wc, s_env, s_light);
  if (!VMIS) {
    if (s_env.valid && !trace_any(sv, s_env.o, s_env.d, 0.0f, s_env.tmax, st)) {
      float4 a = ps.nee_env[p];
      a.x += s_env.c.x;
      a.y += s_env.c.y;
      a.z += s_env.c.z;
      ps.nee_env[p] = a;
    }
    if (s_light.valid && !trace_any(sv, s_light.o, s_light.d, 0.0f, s_light.tmax, st)) {
      float4 a = ps.nee_light[p];
      a.x += s_light.c.x;
      a.y += s_light.c.y;
      a.z += s_light.c.z;
      ps.nee_light[p] = a;
    }
  }
  return cont;
}

// The tail's bounces traced (its stage's query count), one atomic per wave.
__device__ __forceinline__ void add_tail_queries(unsigned long long* stats, const WalkCounts& wc) {
  unsigned long long q = wc.q;
  for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off);
  if (lane_id() == 0) atomicAdd(stats + (3 * kSeg + blockIdx.x % kSeg) * kStatStride, q);
}

template <bool MED, bool VMIS, bool TRANSP, bool JOBS>
__global__ void __launch_bounds__(kBlock) k_tail(SceneView sv, FrameParams fp, PathState ps, Queue qin, Spill sp,
                                                 uint32_t max_steps, unsigned long long* cut,
                                                 unsigned long long* stats) {
  __shared__ uint2 lds_stack[kShortStack * kBlock];
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t seg = blockIdx.x % kSeg, n = qin.counts[seg * kSegStride];
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + seg * kStatStride, (unsigned long long)n);
  LaneStack st{lds_stack + threadIdx.x, kBlock, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
  const uint32_t stride = (gridDim.x / kSeg) * kBlock;
  WalkCounts wc;
  for (uint32_t j = (blockIdx.x / kSeg) * kBlock + threadIdx.x; j < n; j += stride) {
    const uint32_t p = qin.items[seg * qin.seg_cap + j];
    for (uint32_t step = 0;; ++step) {
      if (step == max_steps) {
        atomicAdd(cut, 1ull);
        break;
      }
      if (!tail_bounce<MED, VMIS, TRANSP, JOBS>(sv, fp, ps, p, st, wc)) break;
    }
  }
  if (stats) add_tail_queries(stats, wc);
}

// Waves per SIMD of the refilling tail kernel.  Three (168 registers, some
// spilled) measured ahead of two on the 1080p medium scene's 4-frame batches
// (tail 3.1 -> 2.6 ms) and level with it on 20-frame batches; four matched
// three.
#ifndef HPT_TAIL_WAVES_PER_EU
#define HPT_TAIL_WAVES_PER_EU 3
#endif

// The tail with lane refill (as k_extend_refill): resident blocks step
// their lanes' paths a bounce at a time, and once HPT_REFILL_IDLE lanes of a
// wave have no path those lanes take the next paths of the block's segment
// (then the other segments of its XCD), so a wave is not held by the lane
// whose paths happen to be longest.  Each path runs the same bounces as in
// k_tail.
template <bool MED, bool VMIS, bool TRANSP, bool JOBS>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HPT_TAIL_WAVES_PER_EU)))
k_tail_refill(SceneView sv, FrameParams fp, PathState ps, Queue qin,
                                                        Spill sp, uint32_t max_steps, unsigned long long* cut,
                                                        unsigned long long* stats, uint32_t* cursors) {
  __shared__ uint2 lds_stack[kShortStack * kBlock];
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  uint32_t seg = blockIdx.x % kSeg;
  uint32_t segn = qin.counts[seg * kSegStride];
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + seg * kStatStride, (unsigned long long)segn);
  LaneStack st{lds_stack + threadIdx.x, kBlock, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
  WalkCounts wc;
  const uint32_t lane = lane_id();
  uint32_t hops = 0, p = 0, step = 0;
  bool more = true, has = false;
  while (true) {
    const unsigned long long idle = __ballot(!has);
    if (__popcll(idle) >= HPT_REFILL_IDLE || idle == ~0ull) {
      unsigned long long need = idle;
      while (need != 0ull && more) {
        const uint32_t cnt = (uint32_t)__popcll(need);
        const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
        uint32_t got = 0;
        if (lane == leader) got = atomicAdd(cursors + seg * kSegStride, cnt);
        got = (uint32_t)__builtin_amdgcn_readlane((int)got, (int)leader);
        const uint32_t avail = got < segn ? min(cnt, segn - got) : 0u;
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        const bool take = ((need >> lane) & 1ull) != 0ull && rank < avail;
        if (take) {
          p = qin.items[seg * qin.seg_cap + got + rank];
          step = 0;
          has = true;
        }
        need &= ~__ballot(take);
        if (avail < cnt) {  // segment drained: the next one of this XCD
          if (++hops == kSeg / 8) {
            more = false;
          } else {
            seg = (seg + 8) % kSeg;
            segn = qin.counts[seg * kSegStride];
          }
        }
      }
      if (__ballot(has) == 0ull) break;
    }
    if (has) {
      if (step == max_steps) {
        atomicAdd(cut, 1ull);
        has = false;
      } else {
        ++step;
        has = tail_bounce<MED, VMIS, TRANSP, JOBS>(sv, fp, ps, p, st, wc);
      }
    }
  }
  if (stats) add_tail_queries(stats, wc);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(cursors + kSeg * kSegStride, 1u) == gridDim.x - 1) {
      for (uint32_t s = 0; s < kSeg; ++s) cursors[s * kSegStride] = 0;
      cursors[kSeg * kSegStride] = 0;
      __threadfence();
    }
  }
}

// ------------------------------------------ volume-MIS shading in phases
// The volume-MIS shading kernel above runs the transmittance walks of
// inc/lighting.glsl:7-61 inline: a traversal loop and the whole shading state
// are live together (registers spill), and lanes whose walks run long stall
// the rest of their wave.  The phased form cuts one bounce's shading at the
// walks, keeping the order in which a path draws from its RNG stream:
//   A  (k_vmis_a) up to the first walk: emitters and misses, material,
//      medium, transparency; at a scattering event the environment sample
//      (environment maps) or else the light pick and sample, and the walk's
//      request;
//   W  (k_walk)   the queued walks, one lane per walk: closest hits through
//      transparent or refractive surfaces with Beer-Lambert on exit faces;
//   B  (k_vmis_b, environment maps only) the environment term with its
//      walk's transmittance, then the light pick and sample and its walk's
//      request; W again;
//   C  (k_vmis_c) the light term, then phase-function or BSDF sampling, the
//      continuation and Russian roulette.
// Paths that end or pass through a surface in A continue from A.  A path's
// state between the phases lives in PathState::nee_* (below); the surface's
// hit state and material are recomputed from the hit record in each phase.
// Every value is computed by the same operations in the same order as the
// monolithic kernel, so the images are the same bit for bit.
constexpr uint32_t NEE_WALK = 1u;     // nee_li.w: a walk was queued for this sample
constexpr uint32_t NEE_AREA = 2u;     // nee_li.w: the sampled light has an area (MIS applies)
constexpr uint32_t NEE_SURFACE = 4u;  // nee_ld.w: a surface interaction (else a medium scattering event)

// Appends a walk request (origin, direction, path) to this block's segment
// of wq; want == false lanes take part in the wave's reservation only.
__device__ __forceinline__ void walk_request(bool want, const ShadowQueue& wq, f3 o, f3 d, uint32_t p) {
  const uint32_t k = seg_reserve(want, wq);
  if (want) {
    wq.o[k] = make_float4(o.x, o.y, o.z, 0.0f);
    wq.d[k] = make_float4(d.x, d.y, d.z, __uint_as_float(p));
  }
}

// Direction class of a walk: its octant and the axis of its largest
// component (24 classes; kWalkKeys - 1 marks an empty slot).
constexpr uint32_t kWalkKeys = 32;
__device__ __forceinline__ uint32_t walk_key(float4 d) {
  const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  const uint32_t major = ax >= ay ? (ax >= az ? 0u : 2u) : (ay >= az ? 1u : 2u);
  return oct * 3u + major;
}

// The light pick and sample of inc/lighting.glsl:154-175 at hit_pos into
// nee_dir / nee_li; returns whether the sample needs its walk.
__device__ __forceinline__ bool light_sample_request(const SceneView& sv, const FrameParams& fp, uint4& rng,
                                                     f3 hit_pos, const PathState& ps, uint32_t p) {
  bool want = false;
  if (fp.num_lights > 0) {
    const int li = (int)(rnd(rng) * (float)fp.num_lights);
    const hpt_light& L = sv.lights[li];
    LightSample ls;
    light_sample(rng, fp.num_lights, hit_pos, L, ls);
    want = ls.pdf > 0.0f && dot(ls.direction, ls.normal) < 0.0f;
    if (want) {
      ps.nee_dir[p] = make_float4(ls.direction.x, ls.direction.y, ls.direction.z, ls.pdf);
      ps.nee_li[p] = make_float4(ls.emission.x, ls.emission.y, ls.emission.z,
                                 __uint_as_float(NEE_WALK | (L.area > 0.0f ? NEE_AREA : 0u)));
    }
  }
  if (!want) ps.nee_li[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u));
  return want;
}

// The term of one NEE sample after its walk (inc/lighting.glsl:120-150 and
// :202-238 with VOL_MIS): Ld += mis * Li * f / light_pdf.
template <bool ENV>
__device__ __forceinline__ void nee_term(const PathState& ps, uint32_t p, bool is_surface, bool anb, const Frame& fr,
                                         const Mat& m, f3 V, f3 ffn, float aniso, f3& Ld) {
  const float4 li4 = ps.nee_li[p];
  if (!(__float_as_uint(li4.w) & NEE_WALK)) return;
  const float4 dir4 = ps.nee_dir[p], tr4 = ps.nee_tr[p];
  const f3 trm = mk3(tr4.x, tr4.y, tr4.z), light_dir = mk3(dir4.x, dir4.y, dir4.z);
  const float light_pdf = dir4.w;
  if (luminance(trm) > 0.0f) {
    f3 Li = mk3(li4.x, li4.y, li4.z);
    Li *= trm;
    float pdf = 0.0f;
    f3 f;
    if (is_surface) {
      f = bxdf_eval(anb, fr, m, V, ffn, light_dir, pdf);
    } else {
      pdf = phase_hg(dot(V, light_dir), aniso);
      f = mk3(pdf);
    }
    if (pdf > 0.0f) {
      float mis = 1.0f;
      if (ENV || (__float_as_uint(li4.w) & NEE_AREA)) mis = power_heuristic(light_pdf, pdf);
      if (mis > 0.0f) Ld += mis * Li * f / light_pdf;
    }
  }
}

// A surface interaction's state as the monolithic kernel builds it: hit
// state, material (textures applied), shading frame.
struct SurfaceState {
  HitState h;
  Mat m;
  Frame fr;
};
// s.h holds the hit state of a triangle with a valid material.
__device__ __forceinline__ void surface_state(const SceneView& sv, const float4& h4, f3 dir, uint32_t depth,
                                              SurfaceState& s) {
  const uint32_t id = __float_as_uint(h4.w);
  s.m = load_material(sv.materials + s.h.material);
  s.fr.eta = dot(dir, s.h.normal) < 0.0f ? (1.0f / s.m.ior) : s.m.ior;  // before the normal map
  if ((s.m.base_color_map & s.m.metallic_roughness_map & s.m.emission_map & s.m.normal_map) != INVALID)
    apply_textures(sv, id, h4.y, h4.z, dir, s.m, s.h);
  s.fr.tangent = s.h.tangent;
  s.fr.bitangent = s.h.bitangent;
  s.fr.depth = depth;
}

// End of a bounce for a path that lives on (pathtracer.rgen.glsl:245-256):
// depth limit, Russian roulette, the next segment's state.
__device__ __forceinline__ bool continue_path(const PathState& ps, uint32_t p, uint4& rng, bool alive, uint32_t depth,
                                              uint32_t max_depth, uint32_t rr_depth, f3 org, f3 dir, float pdf, f3 T,
                                              bool anb, bool in_medium, f3 med_color, float med_density,
                                              float med_aniso, uint32_t med_type) {
  if (alive && depth >= max_depth) alive = false;
  if (alive && depth >= rr_depth && rr_depth > 0) {
    const float q = fminf(fmaxf(T.x, fmaxf(T.y, T.z)) + 0.001f, 0.95f);
    if (rnd(rng) > q) alive = false;
    else T /= q;
  }
  if (alive) {
    depth++;
    const uint32_t nf = (anb ? PATH_ANB : 0u) | (in_medium ? PATH_IN_MEDIUM : 0u);
    ps.org[p] = make_float4(org.x, org.y, org.z, pdf);
    ps.dir[p] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(depth));
    ps.thr[p] = make_float4(T.x, T.y, T.z, __uint_as_float(nf));
    ps.med[p] = make_float4(med_color.x, med_color.y, med_color.z, med_density);
    ps.med2[p] = make_float2(med_aniso, __uint_as_float(med_type));
  }
  ps.rng[p] = rng;
  return alive;
}

#ifndef HPT_VMIS_PHASE_WAVES_PER_EU
#define HPT_VMIS_PHASE_WAVES_PER_EU 4
#endif

template <bool TRANSP, bool JOBS>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HPT_VMIS_PHASE_WAVES_PER_EU)))
    k_vmis_a(SceneView sv, FrameParams fp, PathState ps, Queue qin, uint32_t n_max, Queue qout, Queue qnee,
             ShadowQueue wq, unsigned long long* stats) {
  uint32_t p;
  const bool active = queue_entry(qin, n_max, stats, p);
  bool cont = false, nee = false, walk = false;
  f3 wo = mk3(0.0f), wd = mk3(0.0f);
  if (active) {
    const float4 o4 = ps.org[p], d4 = ps.dir[p], t4 = ps.thr[p], h4 = ps.hit[p];
    f3 org = mk3(o4.x, o4.y, o4.z), dir = mk3(d4.x, d4.y, d4.z), T = mk3(t4.x, t4.y, t4.z);
    float pdf = o4.w;
    uint32_t depth = __float_as_uint(d4.w);
    uint32_t max_depth = fp.max_depth, rr_depth = fp.rr_depth;
    if (JOBS) {
      const uint2 lim = fp.job_limits[ps.job[p]];
      max_depth = lim.x;
      rr_depth = lim.y;
    }
    const uint32_t flags = __float_as_uint(t4.w);
    const bool anb = (flags & PATH_ANB) != 0;
    bool in_medium = (flags & PATH_IN_MEDIUM) != 0;
    uint4 rng = ps.rng[p];
    const float4 r4 = ps.rad[p];
    f3 radiance = mk3(r4.x, r4.y, r4.z);
    const float4 m4 = ps.med[p];
    const float2 m2 = ps.med2[p];
    f3 med_color = mk3(m4.x, m4.y, m4.z);
    float med_density = m4.w, med_aniso = m2.x;
    uint32_t med_type = __float_as_uint(m2.y);
    const uint32_t id = __float_as_uint(h4.w);
    const float t_hit = h4.x;

    if (id == MISS_ID) {
      const f3 env = env_eval(sv, fp, depth, pdf, dir, false);
      radiance += T * env;
      if (depth == 0) {
        ps.albedo[p] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
        ps.normal[p] = make_float4(-dir.x, -dir.y, -dir.z, 0.0f);
      }
    } else if (id & LIGHT_BIT) {
      const hpt_light& L = sv.lights[id & ~LIGHT_BIT];
      float light_pdf = 1.0f;
      light_isect(L, org, dir, light_pdf);
      float mis = 1.0f;
      if (depth > 0) mis = power_heuristic(pdf, light_pdf);
      radiance += mis * ld3(L.intensity) * T;
    } else {
      SurfaceState s;
      s.h = triangle_state(sv, id, t_hit, h4.y, h4.z, org, dir);
      if (s.h.material == INVALID) {
        radiance = mk3(0.7f, 0.0f, 0.7f);
        ps.nee_env[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ps.nee_light[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ps.rad[p] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
        goto done;
      }
      surface_state(sv, h4, dir, depth, s);
      if (depth == 0) {
        ps.albedo[p] = make_float4(s.m.base_color.x, s.m.base_color.y, s.m.base_color.z, 0.0f);
        ps.normal[p] = make_float4(s.h.normal.x, s.h.normal.y, s.h.normal.z, 0.0f);
      }
      radiance += T * s.m.emission;
      bool medium_sampled = false;
      if (in_medium) {
        if (med_type == HPT_MEDIUM_ABSORB) {
          const f3 e = -(1.0f - med_color) * t_hit * med_density;
          T *= mk3(expf(e.x), expf(e.y), expf(e.z));
        } else if (med_type == HPT_MEDIUM_EMISSIVE) {
          radiance += med_color * t_hit * med_density * T;
        } else {
          const float scatter_dist = fminf(-logf(rnd(rng)) / med_density, t_hit);
          medium_sampled = scatter_dist < t_hit;
          if (medium_sampled) {
            T *= med_color;
            org = org + dir * scatter_dist;
            s.h.pos = org;
          }
        }
      }
      bool pass = false;
      if (!medium_sampled && TRANSP) {
        if (s.m.opacity == 0.0f || rnd(rng) > s.m.opacity) {
          depth--;  // uint: wraps at depth 0 and the max_depth test ends the path
          pass = true;
        }
      }
      if (pass) {
        // A pass-through: the continuation without next-event estimation.
        org = s.h.pos + dir * EPS;
        if (dot(dir, s.h.normal) < 0.0f && s.m.medium_type != HPT_MEDIUM_NONE) {
          in_medium = true;
          med_color = s.m.medium_color;
          med_density = s.m.medium_density;
          med_aniso = s.m.medium_anisotropy;
          med_type = s.m.medium_type;
        } else {
          in_medium = false;
          med_type = HPT_MEDIUM_NONE;
        }
        cont = continue_path(ps, p, rng, true, depth, max_depth, rr_depth, org, dir, pdf, T, anb, in_medium,
                             med_color, med_density, med_aniso, med_type);
      } else {
        // A scattering event (surface or medium): next-event estimation
        // starts (inc/lighting.glsl:69-75).
        nee = true;
        const f3 scatter_pos = s.h.pos + s.h.ffnormal * EPS;
        if (fp.env_type != HPT_ENV_SKY) {
          f3 Li;
          float light_pdf;
          const f3 light_dir = sample_env_map(rng, sv, fp, Li, light_pdf);
          walk = light_pdf > 0.0f;
          ps.nee_dir[p] = make_float4(light_dir.x, light_dir.y, light_dir.z, light_pdf);
          ps.nee_li[p] = make_float4(Li.x, Li.y, Li.z, __uint_as_float(walk ? NEE_WALK : 0u));
          wd = light_dir;
        } else {
          walk = light_sample_request(sv, fp, rng, s.h.pos, ps, p);
          if (walk) {
            const float4 dd = ps.nee_dir[p];
            wd = mk3(dd.x, dd.y, dd.z);
          }
        }
        wo = scatter_pos;
        ps.nee_ld[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(medium_sampled ? 0u : NEE_SURFACE));
        ps.org[p] = make_float4(org.x, org.y, org.z, pdf);
        ps.thr[p] = make_float4(T.x, T.y, T.z, __uint_as_float(flags));
        ps.rng[p] = rng;
      }
    }
    ps.rad[p] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
  }
done:
  seg_append(cont, qout, p);
  seg_append(nee, qnee, p);
  walk_request(walk, wq, wo, wd, p);
}

// The queued transmittance walks (inc/lighting.glsl:7-61), one lane per walk:
// the path's RNG stream advances for partly transparent surfaces, as inline.
template <bool TRANSP, bool JOBS, bool COUNT, uint32_t BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(HPT_WALK_WAVES_PER_EU)))
    k_walk(SceneView sv, FrameParams fp, PathState ps, ShadowQueue wq, Spill sp, unsigned long long* stats) {
  __shared__ uint2 lds_stack[kShortStack * BS];
  __shared__ uint32_t s_bin[kWalkKeys];
  __shared__ uint16_t s_order[BS];
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  SegSlot t = seg_slot<BS>(wq.counts);
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + t.seg * kStatStride, (unsigned long long)t.n);
  WalkCounts wc;
  // The block's walks grouped by direction (octant and major axis) before
  // the traversal: a block holds the walks of neighbouring paths, and the
  // light each picked is random, so unsorted a wave mixes walks toward every
  // light.  Grouped, a wave's lanes mostly share their origin region and
  // target, and with them their node fetches.  Each walk writes only its own
  // path's transmittance and RNG state, so the order changes no result.
  const uint32_t jb = t.j - threadIdx.x;  // the block's first entry
  if (jb < t.n) {
    const uint32_t key = t.j < t.n ? walk_key(wq.d[t.seg * wq.seg_cap + t.j]) : kWalkKeys - 1;
    if (threadIdx.x < kWalkKeys) s_bin[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t rank = atomicAdd(&s_bin[key], 1u);
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the bins by the first wave
      const uint32_t c = threadIdx.x < kWalkKeys ? s_bin[threadIdx.x] : 0u;
      uint32_t incl = c;
      for (uint32_t off = 1; off < kWalkKeys; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane_id() >= off) incl += v;
      }
      if (threadIdx.x < kWalkKeys) s_bin[threadIdx.x] = incl - c;
    }
    __syncthreads();
    s_order[s_bin[key] + rank] = (uint16_t)threadIdx.x;
    __syncthreads();
    t.j = jb + s_order[threadIdx.x];
  }
  if (t.j < t.n) {
    const uint32_t k = t.seg * wq.seg_cap + t.j;
    const float4 o4 = wq.o[k], d4 = wq.d[k];
    const uint32_t p = __float_as_uint(d4.w);
    const uint32_t max_depth = JOBS ? fp.job_limits[ps.job[p]].x : fp.max_depth;
    LaneStack st{lds_stack + threadIdx.x, BS, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
    uint4 rng = ps.rng[p];
    const f3 tr = eval_transmittance<TRANSP, COUNT>(sv, max_depth, rng, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z),
                                                     st, wc);
    ps.nee_tr[p] = make_float4(tr.x, tr.y, tr.z, 0.0f);
    ps.rng[p] = rng;
  }
  if (COUNT) add_walks(stats, wc);
}

// The queued walks by resident blocks that refill their lanes, as
// k_extend_refill does for extension rays: a lane traces one closest-hit
// segment of its walk at a time; once HPT_REFILL_IDLE lanes of a wave are
// idle, the lanes whose segment ended take its end together (the surface,
// its material, the RNG draw of a partly transparent one, Beer-Lambert on an
// exit face) and either start the walk's next segment or store its
// transmittance, and the lanes still idle take the next walks of the
// block's segment.  Every walk runs the operations of eval_transmittance in
// its order, its RNG draws included, so the results equal k_walk's.
// 6 waves per SIMD (80 registers) measured ahead of 7 (spills) and 5 on the
// 1080p medium scene: 10.1 against 10.8 and 10.5 ms of walks per 20-frame batch.
#ifndef HPT_WALK_REFILL_IDLE
#define HPT_WALK_REFILL_IDLE HPT_REFILL_IDLE
#endif
#ifndef HPT_WALK_REFILL_WAVES_PER_EU
#define HPT_WALK_REFILL_WAVES_PER_EU 6
#endif
template <bool TRANSP, bool JOBS, bool COUNT, uint32_t BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(HPT_WALK_REFILL_WAVES_PER_EU)))
    k_walk_refill(SceneView sv, FrameParams fp, PathState ps, ShadowQueue wq, Spill sp, unsigned long long* stats,
                  uint32_t* cursors) {
  __shared__ uint2 lds_stack[kShortStack * BS];
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  uint32_t seg = blockIdx.x % kSeg;
  uint32_t segn = wq.counts[seg * kSegStride];
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + seg * kStatStride, (unsigned long long)segn);
  LaneStack st{lds_stack + threadIdx.x, BS, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
  WalkCounts wc;
  const uint32_t lane = lane_id();
  // A walk's transmittance so far and its segment count live in its
  // path's nee_tr (xyz, w as the count) between segments: registers are
  // what limits these kernels' waves per SIMD.
  uint32_t hops = 0, k = 0;
  bool more = true, active = false, ended = false;
  Walker<false> w;
  while (true) {
    const unsigned long long idle = __ballot(!active);
    if (__popcll(idle) >= HPT_WALK_REFILL_IDLE || idle == ~0ull) {
      if (ended) {  // the end of a segment (eval_transmittance's loop body)
        ended = false;
        bool done = true;
        TraceResult r{0.0f, 0.0f, 0.0f, MISS_ID};
        if (w.found) r = TraceResult{w.best_t, w.bu, w.bv, w.best_id};
        const float4 d4 = wq.d[k];
        const uint32_t p = __float_as_uint(d4.w);
        const f3 o = w.r.org, d = mk3(d4.x, d4.y, d4.z);
        closest_lights(sv, o, d, 0.0f, WORLD_SIZE, w.found, r);
        if (COUNT) {
          wc.q++;
          wc.nv += w.nv;
          wc.nt += w.nt;
        }
        const float4 t4 = ps.nee_tr[p];
        f3 tr = mk3(t4.x, t4.y, t4.z);
        uint32_t depth = __float_as_uint(t4.w);
        if (!(r.id == MISS_ID || (r.id & LIGHT_BIT))) {
          const HitState h = triangle_state(sv, r.id, r.t, r.u, r.v, o, d);
          if (h.material == INVALID) {
            tr = mk3(0.0f);
          } else {
            const Mat m = load_material(sv.materials + h.material);
            bool is_transparent = false;
            if (TRANSP) {
              if (m.opacity == 0.0f) {
                is_transparent = true;
              } else {
                uint4 rng = ps.rng[p];
                is_transparent = rnd(rng) > m.opacity;
                ps.rng[p] = rng;
              }
            }
            const bool is_refractive = (1.0f - m.metallic) * m.specular_transmission > 0.0f;
            if (!(is_transparent || is_refractive)) {
              tr = mk3(0.0f);
            } else {
              if (dot(d, h.normal) > 0.0f && m.medium_type != HPT_MEDIUM_NONE) {
                const f3 color = m.medium_type == HPT_MEDIUM_ABSORB ? mk3(1.0f) - m.medium_color : mk3(1.0f);
                const f3 e = -color * m.medium_density * r.t;
                tr *= mk3(expf(e.x), expf(e.y), expf(e.z));
              }
              const uint32_t max_depth = JOBS ? fp.job_limits[ps.job[p]].x : fp.max_depth;
              if (++depth < max_depth) {
                w.begin(prepare_ray(h.pos + d * EPS, d), 0.0f, WORLD_SIZE, st);
                active = true;
                done = false;
                ps.nee_tr[p] = make_float4(tr.x, tr.y, tr.z, __uint_as_float(depth));
              }
            }
          }
        }
        if (done) ps.nee_tr[p] = make_float4(tr.x, tr.y, tr.z, 0.0f);
      }
      unsigned long long need = __ballot(!active);
      while (need != 0ull && more) {
        const uint32_t cnt = (uint32_t)__popcll(need);
        const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
        uint32_t got = 0;
        if (lane == leader) got = atomicAdd(cursors + seg * kSegStride, cnt);
        got = (uint32_t)__builtin_amdgcn_readlane((int)got, (int)leader);
        const uint32_t avail = got < segn ? min(cnt, segn - got) : 0u;
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        const bool take = ((need >> lane) & 1ull) != 0ull && rank < avail;
        if (take) {
          k = seg * wq.seg_cap + got + rank;
          const float4 o4 = wq.o[k], d4 = wq.d[k];
          const uint32_t p = __float_as_uint(d4.w);
          const uint32_t max_depth = JOBS ? fp.job_limits[ps.job[p]].x : fp.max_depth;
          ps.nee_tr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);  // no segment traced yet
          if (max_depth > 0) {
            w.begin(prepare_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z)), 0.0f, WORLD_SIZE, st);
            active = true;
          }
        }
        need &= ~__ballot(take);
        if (avail < cnt) {  // segment drained: the next one of this XCD
          if (++hops == kSeg / 8) {
            more = false;
          } else {
            seg = (seg + 8) % kSeg;
            segn = wq.counts[seg * kSegStride];
          }
        }
      }
      if (__ballot(active) == 0ull) break;
    }
    if (active && w.step(sv, st)) {
      active = false;
      ended = true;
    }
  }
  if (COUNT) add_walks(stats, wc);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(cursors + kSeg * kSegStride, 1u) == gridDim.x - 1) {
      for (uint32_t s = 0; s < kSeg; ++s) cursors[s * kSegStride] = 0;
      cursors[kSeg * kSegStride] = 0;
      __threadfence();
    }
  }
}

// A scattering event as A left it: the hit state of the surface that ended
// the segment (its position replaced by the path's origin, the scattering
// point, for a medium event) and, for a surface event, its material and
// frame.
struct NeeEvent {
  bool is_surface, anb;
  SurfaceState s;
  f3 V;
  float aniso;
};
__device__ __forceinline__ void nee_event(const SceneView& sv, const PathState& ps, uint32_t p, f3 org, f3 dir,
                                          uint32_t depth, uint32_t flags, NeeEvent& e) {
  e.is_surface = (__float_as_uint(ps.nee_ld[p].w) & NEE_SURFACE) != 0;
  e.anb = (flags & PATH_ANB) != 0;
  e.V = -dir;
  e.aniso = ps.med2[p].x;
  const float4 h4 = ps.hit[p];
  e.s.h = triangle_state(sv, __float_as_uint(h4.w), h4.x, h4.y, h4.z, org, dir);
  if (e.is_surface) surface_state(sv, h4, dir, depth, e.s);
  else e.s.h.pos = org;
}

// Environment maps: the environment term, then the light pick and sample.
template <bool TRANSP, bool JOBS>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HPT_VMIS_PHASE_WAVES_PER_EU)))
    k_vmis_b(SceneView sv, FrameParams fp, PathState ps, Queue qnee, ShadowQueue wq) {
  const SegSlot t = seg_slot(qnee.counts);
  bool walk = false;
  f3 wo = mk3(0.0f), wd = mk3(0.0f);
  uint32_t p = 0;
  if (t.j < t.n) {
    p = qnee.items[t.seg * qnee.seg_cap + t.j];
    const float4 o4 = ps.org[p], d4 = ps.dir[p], t4 = ps.thr[p];
    const f3 org = mk3(o4.x, o4.y, o4.z), dir = mk3(d4.x, d4.y, d4.z);
    NeeEvent e;
    nee_event(sv, ps, p, org, dir, __float_as_uint(d4.w), __float_as_uint(t4.w), e);
    f3 Ld = mk3(0.0f);
    nee_term<true>(ps, p, e.is_surface, e.anb, e.s.fr, e.s.m, e.V, e.s.h.ffnormal, e.aniso, Ld);
    uint4 rng = ps.rng[p];
    const f3 scatter_pos = e.s.h.pos + e.s.h.ffnormal * EPS;
    walk = light_sample_request(sv, fp, rng, e.s.h.pos, ps, p);
    if (walk) {
      const float4 dd = ps.nee_dir[p];
      wd = mk3(dd.x, dd.y, dd.z);
    }
    wo = scatter_pos;
    ps.nee_ld[p] = make_float4(Ld.x, Ld.y, Ld.z, ps.nee_ld[p].w);
    ps.rng[p] = rng;
  }
  walk_request(walk, wq, wo, wd, p);
}

// The light term, then the scattered direction (phase function or BSDF),
// the continuation and Russian roulette.
template <bool TRANSP, bool JOBS>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HPT_SHADE_WAVES_PER_EU)))
    k_vmis_c(SceneView sv, FrameParams fp, PathState ps, Queue qnee, Queue qout) {
  const SegSlot t = seg_slot(qnee.counts);
  bool cont = false;
  uint32_t p = 0;
  if (t.j < t.n) {
    p = qnee.items[t.seg * qnee.seg_cap + t.j];
    const float4 o4 = ps.org[p], d4 = ps.dir[p], t4 = ps.thr[p];
    f3 org = mk3(o4.x, o4.y, o4.z), dir = mk3(d4.x, d4.y, d4.z), T = mk3(t4.x, t4.y, t4.z);
    float pdf = o4.w;
    const uint32_t depth = __float_as_uint(d4.w);
    uint32_t max_depth = fp.max_depth, rr_depth = fp.rr_depth;
    if (JOBS) {
      const uint2 lim = fp.job_limits[ps.job[p]];
      max_depth = lim.x;
      rr_depth = lim.y;
    }
    const uint32_t flags = __float_as_uint(t4.w);
    bool in_medium = (flags & PATH_IN_MEDIUM) != 0;
    const float4 m4 = ps.med[p];
    const float2 m2 = ps.med2[p];
    f3 med_color = mk3(m4.x, m4.y, m4.z);
    float med_density = m4.w, med_aniso = m2.x;
    uint32_t med_type = __float_as_uint(m2.y);
    NeeEvent e;
    nee_event(sv, ps, p, org, dir, depth, flags, e);
    const float4 ld4 = ps.nee_ld[p];
    f3 Ld = mk3(ld4.x, ld4.y, ld4.z);
    nee_term<false>(ps, p, e.is_surface, e.anb, e.s.fr, e.s.m, e.V, e.s.h.ffnormal, e.aniso, Ld);
    uint4 rng = ps.rng[p];
    const float4 r4 = ps.rad[p];
    f3 radiance = mk3(r4.x, r4.y, r4.z);
    bool alive = true, anb = e.anb;
    if (!e.is_surface) {
      radiance += Ld * T;
      const float r1 = rnd(rng);
      const float r2 = rnd(rng);
      const f3 sdir = sample_hg(-dir, med_aniso, r1, r2);
      pdf = phase_hg(dot(-dir, sdir), med_aniso);
      dir = sdir;
    } else {
      radiance += T * Ld;
      f3 L;
      float spdf;
      uint32_t sflags = 0;
      const f3 f = bxdf_sample(rng, anb, e.s.fr, e.s.m, -dir, e.s.h.ffnormal, L, spdf, sflags);
      dir = L;
      pdf = spdf;
      if (pdf > 0.0f) T *= f / pdf;
      else alive = false;
      if (alive) {
        org = e.s.h.pos + dir * EPS;
        if (dot(dir, e.s.h.normal) < 0.0f && e.s.m.medium_type != HPT_MEDIUM_NONE) {
          in_medium = true;
          med_color = e.s.m.medium_color;
          med_density = e.s.m.medium_density;
          med_aniso = e.s.m.medium_anisotropy;
          med_type = e.s.m.medium_type;
        } else {
          in_medium = false;
          med_type = HPT_MEDIUM_NONE;
        }
      }
    }
    cont = continue_path(ps, p, rng, alive, depth, max_depth, rr_depth, org, dir, pdf, T, anb, in_medium, med_color,
                         med_density, med_aniso, med_type);
    ps.rad[p] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
  }
  seg_append(cont, qout, p);
}

// ------------------------------------------------------------------ shadow
// Any-hit per shadow ray; an unoccluded ray adds its contribution to the
// path's environment (slot 0) or light (slot 1) NEE sum.  Each path has at
// most one ray per slot per bounce, so the sum can be read when the ray
// starts and written when it ends.
template <bool COUNT, uint32_t BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(HPT_WALK_WAVES_PER_EU))) k_shadow(SceneView sv, PathState ps, ShadowQueue sq, Spill sp,
                                                   unsigned long long* stats) {
  __shared__ uint2 lds_stack[kShortStack * BS];
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  const LaneRange lr = lane_range<BS>(sq.counts, sq.seg_cap, false, 0, stats);
  LaneStack st{lds_stack + threadIdx.x, BS, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
  uint32_t cv = 0, ct = 0;
  // Current ray: tmax, tag, contribution, the NEE sum it adds to.
  float tmax = 0.0f;
  uint32_t tag = 0;
  float4 c4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), acc = c4;
  auto sum_of = [&](uint32_t tg) -> float4* { return ((tg & 1u) ? ps.nee_light : ps.nee_env) + (tg >> 1); };
  auto end = [&](bool occluded) {
    if (occluded) return;
    acc.x += c4.x;
    acc.y += c4.y;
    acc.z += c4.z;
    *sum_of(tag) = acc;
  };
  if (sv.node_count == 0) {  // lights only
    for (uint32_t i = 0, j = lr.j; i < kRaysPerLane && j < lr.n; ++i, j += 64) {
      const uint32_t k = lr.base + j;
      const float4 o4 = sq.o[k], d4 = sq.d[k];
      tmax = o4.w;
      tag = __float_as_uint(d4.w);
      c4 = sq.c[k];
      acc = *sum_of(tag);
      end(any_light(sv, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z), 0.0f, tmax));
    }
    return;
  }
  uint32_t i = 0, j = lr.j;
  bool active = j < lr.n;
  Walker<true> w;
  if (active) {
    const uint32_t k = lr.base + j;
    const float4 o4 = sq.o[k], d4 = sq.d[k];
    c4 = sq.c[k];
    tmax = o4.w;
    tag = __float_as_uint(d4.w);
    acc = *sum_of(tag);
    w.begin(prepare_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z)), 0.0f, tmax, st);
  }
  bool has_n = active && kRaysPerLane > 1 && j + 64 < lr.n;
  float4 on4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), dn4 = on4, cn4 = on4;
  if (has_n) {
    const uint32_t k = lr.base + j + 64;
    on4 = sq.o[k];
    dn4 = sq.d[k];
    cn4 = sq.c[k];
  }
  while (__ballot(active) != 0) {
    if (active && w.step(sv, st)) {
      end(w.found || any_light(sv, w.r.org, w.r.dir, 0.0f, tmax));
      if (COUNT) {
        cv += w.nv;
        ct += w.nt;
      }
      active = has_n;
      if (active) {
        ++i;
        j += 64;
        tmax = on4.w;
        tag = __float_as_uint(dn4.w);
        c4 = cn4;
        acc = *sum_of(tag);
        w.begin(prepare_ray(mk3(on4.x, on4.y, on4.z), mk3(dn4.x, dn4.y, dn4.z)), 0.0f, tmax, st);
        has_n = i + 1 < kRaysPerLane && j + 64 < lr.n;
        if (has_n) {
          const uint32_t k = lr.base + j + 64;
          on4 = sq.o[k];
          dn4 = sq.d[k];
          cn4 = sq.c[k];
        }
      }
    }
  }
  if (COUNT) add_visits(stats, cv, ct);
}

// Shadow rays by resident blocks that refill their lanes, as k_extend_refill
// does for extension rays.  A ray's origin, direction, tmax, tag and
// contribution are read again from its queue slot when it ends, and its NEE
// sum is read then rather than when it starts: no other ray of the launch
// adds to that sum.
template <bool COUNT, uint32_t BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(HPT_WALK_WAVES_PER_EU)))
    k_shadow_refill(SceneView sv, PathState ps, ShadowQueue sq, Spill sp, unsigned long long* stats,
                    uint32_t* cursors) {
  __shared__ uint2 lds_stack[kShortStack * BS];
  const uint32_t gid = blockIdx.x * BS + threadIdx.x;
  uint32_t seg = blockIdx.x % kSeg;
  uint32_t segn = sq.counts[seg * kSegStride];
  if (stats && threadIdx.x == 0 && blockIdx.x < kSeg) atomicAdd(stats + seg * kStatStride, (unsigned long long)segn);
  LaneStack st{lds_stack + threadIdx.x, BS, sp.buf ? sp.buf + gid : nullptr, (int)sp.stride, 0};
  const uint32_t lane = lane_id();
  uint32_t cv = 0, ct = 0, hops = 0, k = 0;
  bool more = true, active = false, ended = false;
  Walker<true> w;
  auto finish = [&] {
    const float4 o4 = sq.o[k], d4 = sq.d[k];
    if (w.found || any_light(sv, mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z), 0.0f, o4.w)) return;
    const uint32_t tag = __float_as_uint(d4.w);
    float4* sum = ((tag & 1u) ? ps.nee_light : ps.nee_env) + (tag >> 1);
    const float4 c4 = sq.c[k];
    float4 acc = *sum;
    acc.x += c4.x;
    acc.y += c4.y;
    acc.z += c4.z;
    *sum = acc;
  };
  while (true) {
    const unsigned long long idle = __ballot(!active);
    if (__popcll(idle) >= HPT_REFILL_IDLE || idle == ~0ull) {
      if (ended) {
        finish();
        if (COUNT) {
          cv += w.nv;
          ct += w.nt;
        }
        ended = false;
      }
      unsigned long long need = idle;
      while (need != 0ull && more) {
        const uint32_t cnt = (uint32_t)__popcll(need);
        const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
        uint32_t got = 0;
        if (lane == leader) got = atomicAdd(cursors + seg * kSegStride, cnt);
        got = (uint32_t)__builtin_amdgcn_readlane((int)got, (int)leader);
        const uint32_t avail = got < segn ? min(cnt, segn - got) : 0u;
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        const bool take = ((need >> lane) & 1ull) != 0ull && rank < avail;
        if (take) {
          k = seg * sq.seg_cap + got + rank;
          const float4 o4 = sq.o[k], d4 = sq.d[k];
          w.begin(prepare_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z)), 0.0f, o4.w, st);
          active = true;
        }
        need &= ~__ballot(take);
        if (avail < cnt) {  // segment drained: the next one of this XCD
          if (++hops == kSeg / 8) {
            more = false;
          } else {
            seg = (seg + 8) % kSeg;
            segn = sq.counts[seg * kSegStride];
          }
        }
      }
      if (__ballot(active) == 0ull) break;
    }
    if (active && w.step(sv, st)) {
      active = false;
      ended = true;
    }
  }
  if (COUNT) add_visits(stats, cv, ct);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(cursors + kSeg * kSegStride, 1u) == gridDim.x - 1) {
      for (uint32_t s = 0; s < kSeg; ++s) cursors[s * kSegStride] = 0;
      cursors[kSeg * kSegStride] = 0;
      __threadfence();
    }
  }
}

// -------------------------------------------------------------- accumulate
__device__ __forceinline__ float srgb(float x) {  // common/common.glsl:26-32
  if (x <= 0.0031308f) return 12.92f * x;
  return 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}
__device__ __forceinline__ unsigned char unorm8(float x) {
  const float c = fminf(fmaxf(x, 0.0f), 1.0f);
  return (unsigned char)floorf(c * 255.0f + 0.5f);
}
__device__ __forceinline__ f3 row_mul(f3 v, f3 r0, f3 r1, f3 r2) { return mk3(dot(v, r0), dot(v, r1), dot(v, r2)); }

// pathtracer.rgen.glsl:259-304 with common/tonemap.glsl.  The batch's frames
// are folded into the running means in frame order, one mix per frame with
// the weight 1 / (frame + 1) of that frame, so the images equal those of
// frame-by-frame rendering; the final image is the one after the last frame.
__global__ void __launch_bounds__(kBlock) k_accumulate(FrameParams fp, PathState ps, Images im, uint32_t npix) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= npix) return;
  f3 acc_c = mk3(0.0f), acc_a = mk3(0.0f), acc_n = mk3(0.0f);
  const uint32_t frame0 = first_frame(fp);
  if (frame0 != 0) {
    const float4 c4 = im.accum[p], a4 = im.albedo[p], n4 = im.normal[p];
    acc_c = mk3(c4.x, c4.y, c4.z);
    acc_a = mk3(a4.x, a4.y, a4.z);
    acc_n = mk3(n4.x, n4.y, n4.z);
  }
  for (uint32_t j = 0; j < fp.batch_frames; ++j) {
    const uint32_t q = j * npix + p;
    const uint32_t frame = frame0 + j;
    const float4 r = ps.rad[q], e = ps.nee_env[q], l = ps.nee_light[q];
    const f3 L = mk3(r.x, r.y, r.z) + mk3(e.x, e.y, e.z) + mk3(l.x, l.y, l.z);
    const float4 a4 = ps.albedo[q], n4 = ps.normal[q];
    const f3 A = mk3(a4.x, a4.y, a4.z), N = mk3(n4.x, n4.y, n4.z);
    if (frame == 0) {
      acc_c = L;
      acc_a = A;
      acc_n = N;
    } else {
      const float w = 1.0f / ((float)frame + 1.0f);
      acc_c = mix3(acc_c, L, w);
      acc_a = mix3(acc_a, A, w);
      acc_n = mix3(acc_n, N, w);
    }
  }
  im.accum[p] = make_float4(acc_c.x, acc_c.y, acc_c.z, 1.0f);
  im.albedo[p] = make_float4(acc_a.x, acc_a.y, acc_a.z, 1.0f);
  im.normal[p] = make_float4(acc_n.x, acc_n.y, acc_n.z, 1.0f);
  f3 c = acc_c * powf(2.0f, fp.exposure_value);
  if (fp.enable_tonemap) {
    if (fp.enable_aces) {
      if (fp.use_simple_aces) {
        const float A = 2.51f, B = 0.03f, Y = 2.43f, D = 0.59f, E = 0.14f;
        const f3 num = c * (A * c + B), den = c * (Y * c + D) + E;
        c = num / den;
        c = mk3(fminf(fmaxf(c.x, 0.f), 1.f), fminf(fmaxf(c.y, 0.f), 1.f), fminf(fmaxf(c.z, 0.f), 1.f));
      } else {
        c = row_mul(c, mk3(0.59719f, 0.35458f, 0.04823f), mk3(0.07600f, 0.90834f, 0.01566f),
                    mk3(0.02840f, 0.13383f, 0.83777f));
        const f3 aa = c * (c + 0.0245786f) - 0.000090537f;
        const f3 bb = c * (0.983729f * c + 0.4329510f) + 0.238081f;
        c = aa / bb;
        c = row_mul(c, mk3(1.60475f, -0.53108f, -0.07367f), mk3(-0.10208f, 1.10813f, -0.00605f),
                    mk3(-0.00327f, -0.07276f, 1.07602f));
        c = mk3(fminf(fmaxf(c.x, 0.f), 1.f), fminf(fmaxf(c.y, 0.f), 1.f), fminf(fmaxf(c.z, 0.f), 1.f));
      }
    } else {
      c = c * 1.0f / (1.0f + luminance(c) / 1.5f);
    }
  }
  im.final_rgba8[p] = make_uchar4(unorm8(srgb(c.x)), unorm8(srgb(c.y)), unorm8(srgb(c.z)), 255);
}

// ------------------------------------------------------ environment records
// One EnvTexel per texel from the texture and the alias table (float4 per
// texel: probability, alias bits, luminance, luminance of the alias).
__global__ void __launch_bounds__(kBlock) k_env_texels(const float4* rgb, const float4* table, uint32_t W, uint32_t H,
                                                       EnvTexel* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= W * H) return;
  const uint32_t ti = i % W, tj = i / W;
  EnvTexel& e = out[i];
  const float4 t = table[i];
  e.prob = t.x;
  e.alias = __float_as_uint(t.y);
  e.lum = t.z;
  e.lum_alias = t.w;
  for (int dy = 0; dy < 3; ++dy) {
    const int j = min(max((int)tj + dy - 1, 0), (int)H - 1);
    for (int dx = 0; dx < 3; ++dx) {
      const int c = (int)((ti + W + (uint32_t)dx - 1u) % W);
      const float4 v = rgb[(size_t)j * W + c];
      e.rgb[dy][dx][0] = v.x;
      e.rgb[dy][dx][1] = v.y;
      e.rgb[dy][dx][2] = v.z;
    }
  }
  e.pad = 0.0f;
}

// --------------------------------------------------------------- ray queries
template <bool ANY>
__global__ void __launch_bounds__(kBlock) k_trace_rays(SceneView sv, const hpt_ray* rays, uint32_t n, float4* out,
                                                       uint32_t* nvis, uint32_t* ntri, Spill sp) {
  __shared__ uint2 lds_stack[kShortStack * kBlock];
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const hpt_ray r = rays[k];
  LaneStack st{lds_stack + threadIdx.x, kBlock, sp.buf ? sp.buf + k : nullptr, (int)sp.stride, 0};
  uint32_t a = 0, b = 0;
  const f3 o = ld3(r.origin), d = ld3(r.direction);
  if (ANY) {
    const bool hit = trace_any(sv, o, d, r.tmin, r.tmax, st, &a, &b);
    out[k] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(hit ? 1u : 0u));
  } else {
    const TraceResult t = trace_closest(sv, o, d, r.tmin, r.tmax, st, &a, &b);
    out[k] = make_float4(t.t, t.u, t.v, __uint_as_float(t.id));
  }
  if (nvis) nvis[k] = a;
  if (ntri) ntri[k] = b;
}

// Closest hits written as hpt_hit records (the conversion hpt_trace_closest
// does on the host): triangles as (primitive, triangle within it), analytic
// lights as their index, misses as HPT_HIT_NONE.
__global__ void __launch_bounds__(kBlock) k_trace_hits(SceneView sv, const hpt_ray* rays, uint32_t n, hpt_hit* hits,
                                                       Spill sp) {
  __shared__ uint2 lds_stack[kShortStack * kBlock];
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  const hpt_ray r = rays[k];
  LaneStack st{lds_stack + threadIdx.x, kBlock, sp.buf ? sp.buf + k : nullptr, (int)sp.stride, 0};
  const TraceResult t = trace_closest(sv, ld3(r.origin), ld3(r.direction), r.tmin, r.tmax, st);
  hpt_hit h;
  if (t.id == MISS_ID) {
    h = hpt_hit{0.0f, 0.0f, 0.0f, HPT_HIT_NONE, HPT_INVALID_INDEX, HPT_INVALID_INDEX};
  } else if (t.id & LIGHT_BIT) {
    h = hpt_hit{t.t, 0.0f, 0.0f, HPT_HIT_LIGHT, t.id & ~LIGHT_BIT, HPT_INVALID_INDEX};
  } else {
    const uint32_t prim = __float_as_uint(reinterpret_cast<const float4*>(sv.tri_shade + t.id)[4].z);
    h = hpt_hit{t.t, t.u, t.v, HPT_HIT_TRIANGLE, prim, t.id - sv.prims[prim].local_base};
  }
  hits[k] = h;
}

// ------------------------------------------------------------------ launchers
static inline uint32_t grid_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_camera(hipStream_t s, const SceneView& sv, const FrameParams& fp, const PathState& ps, uint32_t npaths,
                   uint32_t job) {
  hipLaunchKernelGGL(k_camera, dim3(grid_for(npaths)), dim3(kBlock), 0, s, sv, fp, ps, npaths, job);
}

// Blocks of a launch over qin: the identity queue's grid, or kSeg blocks per
// chunk of the segment capacity.
static inline uint32_t queue_grid(const Queue& qin, uint32_t n_max) {
  return qin.items ? kSeg * seg_geom(n_max).chunks : grid_for(n_max);
}

// Blocks of a traversal launch of bs-thread blocks (kRaysPerLane entries per lane).
static inline uint32_t walk_blocks(uint32_t entries, uint32_t bs) {
  const uint32_t span = bs * kRaysPerLane;
  return (entries + span - 1) / span;
}

// The block sizes the traversal kernels are built for (kTraversalBlocks).
#define HPT_BY_BS(bs, CALL) \
  switch (bs) {             \
    case 64: CALL(64); break;   \
    case 256: CALL(256); break; \
    default: CALL(128); break;  \
  }

// Blocks of bs threads per segment that fit on the device at once at the
// traversal kernels' waves per SIMD (4 SIMDs per CU).
static uint32_t resident_blocks_per_segment(uint32_t bs, uint32_t waves_per_simd = HPT_WALK_WAVES_PER_EU) {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  const uint32_t waves = (uint32_t)cus * 4u * waves_per_simd;
  return std::max<uint32_t>(1u, waves / (bs / 64u) / kSeg);
}

void launch_extend(hipStream_t s, const SceneView& sv, const PathState& ps, const Queue& qin, uint32_t n_max,
                   const Queue& qnext, const ShadowQueue& sq, Spill sp, StageCounters sc, uint8_t* keys,
                   uint32_t* phase_counts, uint32_t bs, uint32_t* cursors) {
  const bool count = sc.count && sc.stats;
  if (cursors && sv.node_count > 0) {
    const dim3 grid(kSeg * std::min(walk_blocks(qin.seg_cap, bs), resident_blocks_per_segment(bs)));
#define HPT_EXTEND_R(C, S, BS_)                                                                                   \
  hipLaunchKernelGGL((k_extend_refill<C, S, BS_>), grid, dim3(BS_), 0, s, sv, ps, qin, n_max, qnext, sq, sp, sc.stats, \
                     keys, phase_counts, cursors)
#define HPT_R_TT(BS_) HPT_EXTEND_R(true, true, BS_)
#define HPT_R_FT(BS_) HPT_EXTEND_R(false, true, BS_)
#define HPT_R_TF(BS_) HPT_EXTEND_R(true, false, BS_)
#define HPT_R_FF(BS_) HPT_EXTEND_R(false, false, BS_)
    if (keys) {
      if (count) HPT_BY_BS(bs, HPT_R_TT)
      else HPT_BY_BS(bs, HPT_R_FT)
    } else {
      if (count) HPT_BY_BS(bs, HPT_R_TF)
      else HPT_BY_BS(bs, HPT_R_FF)
    }
#undef HPT_R_TT
#undef HPT_R_FT
#undef HPT_R_TF
#undef HPT_R_FF
#undef HPT_EXTEND_R
    return;
  }
  const dim3 grid(qin.items ? kSeg * walk_blocks(qin.seg_cap, bs) : walk_blocks(n_max, bs));
#define HPT_EXTEND(C, S, BS_)                                                                                   \
  hipLaunchKernelGGL((k_extend<C, S, BS_>), grid, dim3(BS_), 0, s, sv, ps, qin, n_max, qnext, sq, sp, sc.stats, keys, \
                     phase_counts)
#define HPT_E_TT(BS_) HPT_EXTEND(true, true, BS_)
#define HPT_E_FT(BS_) HPT_EXTEND(false, true, BS_)
#define HPT_E_TF(BS_) HPT_EXTEND(true, false, BS_)
#define HPT_E_FF(BS_) HPT_EXTEND(false, false, BS_)
  if (keys) {
    if (count) HPT_BY_BS(bs, HPT_E_TT)
    else HPT_BY_BS(bs, HPT_E_FT)
  } else {
    if (count) HPT_BY_BS(bs, HPT_E_TF)
    else HPT_BY_BS(bs, HPT_E_FF)
  }
#undef HPT_E_TT
#undef HPT_E_FT
#undef HPT_E_TF
#undef HPT_E_FF
#undef HPT_EXTEND
}

void launch_shade_sort(hipStream_t s, const Queue& qin, uint32_t n_max, const ShadeSort& ss) {
  // kSeg blocks per chunk of kBlock slots of the segment capacity: the
  // geometry of a shading launch over a compacted queue.
  const dim3 grid(kSeg * ((qin.seg_cap + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_sort_count, grid, dim3(kBlock), 0, s, qin, n_max, ss);
  hipLaunchKernelGGL(k_sort_scan, dim3(kSeg), dim3(kSortKeys), 0, s, ss, qin.seg_cap);
  hipLaunchKernelGGL(k_sort_scatter, grid, dim3(kBlock), 0, s, qin, n_max, ss);
}

// The walk-counting variant exists for the volume-MIS permutations of a
// single-job launch (the only shading kernels that walk).
template <bool MED, bool VMIS, bool TRANSP>
static void shade_variant(hipStream_t s, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                          const Queue& qin, uint32_t n_max, const Queue& qout, const ShadowQueue& sq, Spill sp,
                          StageCounters sc) {
  const dim3 grid(queue_grid(qin, n_max));
  if (fp.job_limits)
    hipLaunchKernelGGL((k_shade<MED, VMIS, TRANSP, true, false>), grid, dim3(kBlock), 0, s, sv, fp, ps, qin, n_max,
                       qout, sq, sp, sc.stats);
  else if (VMIS && sc.count && sc.stats)
    hipLaunchKernelGGL((k_shade<MED, VMIS, TRANSP, false, VMIS>), grid, dim3(kBlock), 0, s, sv, fp, ps, qin, n_max,
                       qout, sq, sp, sc.stats);
  else
    hipLaunchKernelGGL((k_shade<MED, VMIS, TRANSP, false, false>), grid, dim3(kBlock), 0, s, sv, fp, ps, qin, n_max,
                       qout, sq, sp, sc.stats);
}

void launch_shade(hipStream_t s, uint32_t features, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                  const Queue& qin, uint32_t n_max, const Queue& qout, const ShadowQueue& sq, Spill sp,
                  StageCounters sc) {
  const bool med = features & HPT_FEATURE_MEDIUM;
  const bool vmis = med && (features & HPT_FEATURE_VOL_MIS);
  const bool tr = features & HPT_FEATURE_TRANSPARENT;
#define HPT_SHADE(M, V, T) shade_variant<M, V, T>(s, sv, fp, ps, qin, n_max, qout, sq, sp, sc)
  if (!med) {
    if (tr) HPT_SHADE(false, false, true);
    else HPT_SHADE(false, false, false);
  } else if (!vmis) {
    if (tr) HPT_SHADE(true, false, true);
    else HPT_SHADE(true, false, false);
  } else {
    if (tr) HPT_SHADE(true, true, true);
    else HPT_SHADE(true, true, false);
  }
#undef HPT_SHADE
}

void launch_tail(hipStream_t s, uint32_t features, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                 const Queue& qin, Spill sp, uint32_t max_steps, unsigned long long* cut, StageCounters sc,
                 uint32_t* cursors) {
  const uint32_t chunks = (qin.seg_cap + kBlock - 1) / kBlock;
  const bool refill = cursors != nullptr;
  const uint32_t per_seg = refill ? std::min(chunks, resident_blocks_per_segment(kBlock, HPT_TAIL_WAVES_PER_EU))
                                  : std::min(chunks, kTailChunks);
  const dim3 grid(kSeg * per_seg);
  const bool med = features & HPT_FEATURE_MEDIUM;
  const bool vmis = med && (features & HPT_FEATURE_VOL_MIS);
  const bool tr = features & HPT_FEATURE_TRANSPARENT;
#define HPT_T(M, V, T, J)                                                                                      \
  if (refill)                                                                                                  \
    hipLaunchKernelGGL((k_tail_refill<M, V, T, J>), grid, dim3(kBlock), 0, s, sv, fp, ps, qin, sp, max_steps, cut, \
                       sc.stats, cursors);                                                                     \
  else                                                                                                         \
    hipLaunchKernelGGL((k_tail<M, V, T, J>), grid, dim3(kBlock), 0, s, sv, fp, ps, qin, sp, max_steps, cut, sc.stats)
#define HPT_TJ(M, V, T)            \
  if (fp.job_limits) HPT_T(M, V, T, true); \
  else HPT_T(M, V, T, false)
  if (!med) {
    if (tr) HPT_TJ(false, false, true);
    else HPT_TJ(false, false, false);
  } else if (!vmis) {
    if (tr) HPT_TJ(true, false, true);
    else HPT_TJ(true, false, false);
  } else {
    if (tr) HPT_TJ(true, true, true);
    else HPT_TJ(true, true, false);
  }
#undef HPT_TJ
#undef HPT_T
}

void launch_vmis_a(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   const Queue& qin, uint32_t n_max, const Queue& qout, const Queue& qnee, const ShadowQueue& wq,
                   StageCounters sc) {
  const dim3 grid(queue_grid(qin, n_max));
#define HPT_A(T, J) hipLaunchKernelGGL((k_vmis_a<T, J>), grid, dim3(kBlock), 0, s, sv, fp, ps, qin, n_max, qout, qnee, wq, \
                                       sc.stats)
  if (fp.job_limits) {
    if (transp) HPT_A(true, true);
    else HPT_A(false, true);
  } else {
    if (transp) HPT_A(true, false);
    else HPT_A(false, false);
  }
#undef HPT_A
}

void launch_walk(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                 const ShadowQueue& wq, uint32_t seg_cap, Spill sp, StageCounters sc, uint32_t bs, uint32_t* cursors) {
  const bool refill = cursors && sv.node_count > 0;
  const uint32_t per_seg = (seg_cap + bs - 1) / bs;
  const dim3 grid(kSeg * (refill ? std::min(per_seg, resident_blocks_per_segment(bs, HPT_WALK_REFILL_WAVES_PER_EU))
                                  : per_seg));
  const bool count = sc.count && sc.stats;
#define HPT_W(T, J, C, BS_)                                                                                   \
  if (refill)                                                                                                 \
    hipLaunchKernelGGL((k_walk_refill<T, J, C, BS_>), grid, dim3(BS_), 0, s, sv, fp, ps, wq, sp, sc.stats, cursors); \
  else                                                                                                        \
    hipLaunchKernelGGL((k_walk<T, J, C, BS_>), grid, dim3(BS_), 0, s, sv, fp, ps, wq, sp, sc.stats)
#define HPT_W_TTF(BS_) HPT_W(true, true, false, BS_)
#define HPT_W_FTF(BS_) HPT_W(false, true, false, BS_)
#define HPT_W_TFT(BS_) HPT_W(true, false, true, BS_)
#define HPT_W_FFT(BS_) HPT_W(false, false, true, BS_)
#define HPT_W_TFF(BS_) HPT_W(true, false, false, BS_)
#define HPT_W_FFF(BS_) HPT_W(false, false, false, BS_)
  if (fp.job_limits) {
    if (transp) HPT_BY_BS(bs, HPT_W_TTF)
    else HPT_BY_BS(bs, HPT_W_FTF)
  } else if (count) {
    if (transp) HPT_BY_BS(bs, HPT_W_TFT)
    else HPT_BY_BS(bs, HPT_W_FFT)
  } else {
    if (transp) HPT_BY_BS(bs, HPT_W_TFF)
    else HPT_BY_BS(bs, HPT_W_FFF)
  }
#undef HPT_W_TTF
#undef HPT_W_FTF
#undef HPT_W_TFT
#undef HPT_W_FFT
#undef HPT_W_TFF
#undef HPT_W_FFF
#undef HPT_W
}

void launch_vmis_b(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   const Queue& qnee, const ShadowQueue& wq) {
  const dim3 grid(kSeg * ((qnee.seg_cap + kBlock - 1) / kBlock));
#define HPT_B(T, J) hipLaunchKernelGGL((k_vmis_b<T, J>), grid, dim3(kBlock), 0, s, sv, fp, ps, qnee, wq)
  if (fp.job_limits) {
    if (transp) HPT_B(true, true);
    else HPT_B(false, true);
  } else {
    if (transp) HPT_B(true, false);
    else HPT_B(false, false);
  }
#undef HPT_B
}

void launch_vmis_c(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   const Queue& qnee, const Queue& qout) {
  const dim3 grid(kSeg * ((qnee.seg_cap + kBlock - 1) / kBlock));
#define HPT_C(T, J) hipLaunchKernelGGL((k_vmis_c<T, J>), grid, dim3(kBlock), 0, s, sv, fp, ps, qnee, qout)
  if (fp.job_limits) {
    if (transp) HPT_C(true, true);
    else HPT_C(false, true);
  } else {
    if (transp) HPT_C(true, false);
    else HPT_C(false, false);
  }
#undef HPT_C
}

void launch_shadow(hipStream_t s, const SceneView& sv, const PathState& ps, const ShadowQueue& sq, uint32_t n_max,
                   Spill sp, StageCounters sc, uint32_t bs, uint32_t* cursors) {
  if (cursors && sv.node_count > 0) {
    const dim3 grid(kSeg * std::min(walk_blocks(sq.seg_cap, bs), resident_blocks_per_segment(bs)));
#define HPT_S_T(BS_) \
  hipLaunchKernelGGL((k_shadow_refill<true, BS_>), grid, dim3(BS_), 0, s, sv, ps, sq, sp, sc.stats, cursors)
#define HPT_S_F(BS_) \
  hipLaunchKernelGGL((k_shadow_refill<false, BS_>), grid, dim3(BS_), 0, s, sv, ps, sq, sp, sc.stats, cursors)
    if (sc.count && sc.stats) HPT_BY_BS(bs, HPT_S_T)
    else HPT_BY_BS(bs, HPT_S_F)
#undef HPT_S_T
#undef HPT_S_F
    return;
  }
  const dim3 grid(kSeg * walk_blocks(sq.seg_cap, bs));
#define HPT_S_T(BS_) hipLaunchKernelGGL((k_shadow<true, BS_>), grid, dim3(BS_), 0, s, sv, ps, sq, sp, sc.stats)
#define HPT_S_F(BS_) hipLaunchKernelGGL((k_shadow<false, BS_>), grid, dim3(BS_), 0, s, sv, ps, sq, sp, sc.stats)
  if (sc.count && sc.stats) HPT_BY_BS(bs, HPT_S_T)
  else HPT_BY_BS(bs, HPT_S_F)
#undef HPT_S_T
#undef HPT_S_F
}

void launch_env_texels(hipStream_t s, const float4* rgb, const float4* table, uint32_t w, uint32_t h, EnvTexel* out) {
  hipLaunchKernelGGL(k_env_texels, dim3(grid_for(w * h)), dim3(kBlock), 0, s, rgb, table, w, h, out);
}

void launch_accumulate(hipStream_t s, const FrameParams& fp, const PathState& ps, const Images& im, uint32_t npaths) {
  hipLaunchKernelGGL(k_accumulate, dim3(grid_for(npaths)), dim3(kBlock), 0, s, fp, ps, im, npaths);
}

__global__ void k_job_limits(uint2* dst, JobLimits lim, uint32_t count) {
  if (threadIdx.x < count) dst[threadIdx.x] = lim.v[threadIdx.x];
}

void launch_job_limits(hipStream_t s, uint2* dst, const JobLimits& lim, uint32_t count) {
  hipLaunchKernelGGL(k_job_limits, dim3(1), dim3(kMaxWaveJobs), 0, s, dst, lim, count);
}

// The device frame counter of graph replays: set from the host's value, or
// advanced by the batch size at the end of a captured batch.
__global__ void k_frame_counter(uint32_t* counter, uint32_t value, int add) {
  if (threadIdx.x == 0) counter[0] = add ? counter[0] + value : value;
}

void launch_frame_counter(hipStream_t s, uint32_t* counter, uint32_t value, bool add) {
  hipLaunchKernelGGL(k_frame_counter, dim3(1), dim3(64), 0, s, counter, value, add ? 1 : 0);
}

void launch_trace_hits(hipStream_t s, const SceneView& sv, const hpt_ray* rays, uint32_t n, hpt_hit* hits, Spill sp) {
  hipLaunchKernelGGL(k_trace_hits, dim3(grid_for(n)), dim3(kBlock), 0, s, sv, rays, n, hits, sp);
}

void launch_trace_rays(hipStream_t s, const SceneView& sv, const hpt_ray* rays, uint32_t n, int any_hit, float4* out,
                       uint32_t* node_visits, uint32_t* tri_tests, Spill sp) {
  if (any_hit)
    hipLaunchKernelGGL(k_trace_rays<true>, dim3(grid_for(n)), dim3(kBlock), 0, s, sv, rays, n, out, node_visits,
                       tri_tests, sp);
  else
    hipLaunchKernelGGL(k_trace_rays<false>, dim3(grid_for(n)), dim3(kBlock), 0, s, sv, rays, n, out, node_visits,
                       tri_tests, sp);
}

}  // namespace hpt
