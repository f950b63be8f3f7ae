// kernels.h — launchers of the gfx950 kernels, called by the C ABI layer.
#pragma once

#include "common.h"

namespace hpt {

// Structure-of-arrays path state, one slot per pixel of the frame
// (pathtracer.rgen.glsl keeps the same values in registers across its loop).
struct PathState {
  uint4* rng;         // s0 of RNGState (common/wltcRS.glsl:4-7)
  float4* org;        // ray origin, .w = state.pdf
  float4* dir;        // ray direction, .w = state.depth (uint bits)
  float4* thr;        // throughput, .w = flag bits (PATH_*)
  float4* rad;        // radiance gathered on the path itself
  float4* albedo;     // first-hit albedo
  float4* normal;     // first-hit normal
  float4* med;        // state.medium.color, .w = density
  float2* med2;       // (anisotropy, type bits)
  float4* nee_env;    // sum of unoccluded environment NEE terms
  float4* nee_light;  // sum of unoccluded light NEE terms
  float4* hit;        // closest hit (t, u, v, id bits)
  uint16_t* job;      // job of each path in a wavefront of several jobs, else null
  // Between the phases of the volume-MIS shading (null when it is not phased):
  float4* nee_dir;    // NEE sample direction, .w = its pdf
  float4* nee_li;     // NEE sample radiance, .w = flag bits (walk queued, light with an area)
  float4* nee_ld;     // the event's NEE sum so far, .w = flag bits (surface event)
  float4* nee_tr;     // transmittance of the last walk
};

// Path state arrays from path `base` on: the slice of one job in a wavefront.
inline PathState path_slice(const PathState& ps, uint32_t base) {
  PathState s = ps;
  s.rng += base;
  s.org += base;
  s.dir += base;
  s.thr += base;
  s.rad += base;
  s.albedo += base;
  s.normal += base;
  s.med += base;
  s.med2 += base;
  s.nee_env += base;
  s.nee_light += base;
  s.hit += base;
  if (s.job) s.job += base;
  if (s.nee_dir) {
    s.nee_dir += base;
    s.nee_li += base;
    s.nee_ld += base;
    s.nee_tr += base;
  }
  return s;
}
constexpr uint32_t PATH_ANB = 1u;             // any_non_specular_bounce
constexpr uint32_t PATH_IN_MEDIUM = 2u;       // is_in_medium
constexpr uint32_t PATH_SURFACE_SCATTER = 4u; // is_surface_scatter (MEDIUM without VOL_MIS)

constexpr int kBlock = 256;
// Block sizes the traversal kernels (extend, walk, shadow) are built for:
// hpt_set_traversal_block picks one; queue segments are multiples of
// kBlock entries, so each divides them.  128 measured fastest on both 1080p
// benchmark scenes (+3-4 % over 256).
constexpr uint32_t kTraversalBlocks[] = {64, 128, 256};
constexpr uint32_t kDefaultTraversalBlock = 128;

// Compacted queues are split into kSeg segments, each with its own append
// counter kSegStride words (256 B) from the next.  Block b of a launch reads
// segment b % kSeg and appends to the same segment, so appends spread over
// kSeg counters -- a single counter serialises one atomic per wave at about
// 11 ns on MI355X -- and, with blocks dealt round-robin to the 8 XCDs, a
// path keeps to one XCD from bounce to bounce.
constexpr uint32_t kSeg = 64;
// Queue entries one lane of the extend and shadow kernels traces.
#ifndef HPT_RAYS_PER_LANE
#define HPT_RAYS_PER_LANE 1
#endif
constexpr uint32_t kRaysPerLane = HPT_RAYS_PER_LANE;
constexpr uint32_t kSegStride = 64;

// Segment geometry of a launch over at most n_max paths: the first bounce
// runs grid_for(n_max) blocks over the identity queue; later launches run
// kSeg * chunks blocks, block b taking entries (b / kSeg) * kBlock + tid of
// segment b % kSeg.  Segments hold chunks * kBlock paths (twice that for
// shadow rays), which bounds what one segment's producer blocks can append.
struct SegGeom {
  uint32_t chunks;   // blocks per segment
  uint32_t seg_cap;  // path entries per segment = chunks * kBlock
};
inline SegGeom seg_geom(uint32_t n_max) {
  const uint32_t blocks = (n_max + kBlock - 1) / kBlock;
  const uint32_t chunks = (blocks + kSeg - 1) / kSeg;
  return SegGeom{chunks, chunks * (uint32_t)kBlock};
}

struct Queue {
  uint32_t* items;   // path ids, segment s at items + s * seg_cap; null = identity over n_max
  uint32_t* counts;  // segment s count at counts[s * kSegStride]
  uint32_t seg_cap;
};

struct ShadowQueue {
  float4* o;  // origin, .w = tmax
  float4* d;  // direction, .w = (path << 1 | slot) bits
  float4* c;  // throughput-weighted contribution
  uint32_t* counts;  // segment s holds entries [s * seg_cap, s * seg_cap + counts[s * kSegStride])
  uint32_t seg_cap;  // 2 * the path segment capacity
};

// Material-sorted shading: extend writes one key per queue slot -- the
// material bucket of the hit triangle (materials past kSortKeys - 3 share the
// last bucket), kSortKeys - 2 for an analytic light, kSortKeys - 1 for a miss
// -- and a counting sort groups each segment's entries by key, so the lanes
// of a shading wave mostly share one material.  Shading a path does not
// depend on its queue position: sorted and unsorted launches give the same
// images bit for bit.
constexpr uint32_t kSortKeys = 64;
// Words of a handle's segment-counter buffer: path queues A and B, the shadow
// queue, the sorted shading queue, the sort's key counts and cursors, the
// phased shading's two queues, then the refill cursors (kCursorOffset).
constexpr size_t kCursorOffset = 6 * kSeg * kSegStride + 2 * kSeg * kSortKeys;
constexpr size_t kCountWords = kCursorOffset + 2 * kSeg * kSegStride;
struct ShadeSort {
  uint8_t* keys;     // one per queue slot, segment s at keys + s * seg_cap
  uint32_t* items;   // sorted path ids, same segment layout
  uint32_t* counts;  // entries of sorted segment s at counts[s * kSegStride]
  uint32_t* bucket;  // kSeg * kSortKeys key counts (zero between sorts), then kSeg * kSortKeys cursors
};

struct Spill {
  uint2* buf;      // traversal stack entries beyond the LDS part, may be null
  uint32_t stride; // threads of the launch
};

// Image planes the frame accumulates into (inc/global.glsl:1-4).
struct Images {
  float4* accum;
  float4* albedo;
  float4* normal;
  uchar4* final_rgba8;
};

// job: the index stored in ps.job (when non-null) for each path launched.
void launch_camera(hipStream_t s, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   uint32_t npaths, uint32_t job = 0);
// Device counters of one stage, 64-bit, spread like the queue counters:
// rays at stats[seg * kStatStride], node visits at stats[(kSeg + seg) *
// kStatStride], triangle tests at stats[(2 * kSeg + seg) * kStatStride],
// closest-hit queries at stats[(3 * kSeg + seg) * kStatStride].  Rays (for
// shading: shading events) are counted whenever stats is non-null; visits,
// tests and queries only by the counting kernel variants (count == true):
// the traversal kernels' own walks, and the volume-MIS transmittance walks
// of the shading kernel.
constexpr uint32_t kStatStride = 32;
constexpr uint32_t kStatCounters = 4;
constexpr uint32_t kStatsPerStage = kStatCounters * kSeg * kStatStride;
struct StageCounters {
  unsigned long long* stats;
  bool count;
};

// qin.items == nullptr: the first bounce, the identity queue over n_max
// paths.  Extend zeroes the counters of qnext and sq, which the shade launch
// after it fills.
// keys: null, or the sort keys of qin's slots to write (see ShadeSort).
// phase_counts: null, or two counter blocks (kSeg * kSegStride words each)
// of the phased volume-MIS shading's queues, zeroed with qnext and sq.
// bs: threads per block of the traversal kernels (extend, walk, shadow):
// one of kTraversalBlocks.
// cursors: null, or kSeg + 1 refill counters (kSegStride words apart, zero
// between launches: the launch's last block clears them).  With cursors a
// launch over a compacted queue runs resident blocks whose lanes take new
// entries when enough of a wave's rays have ended (k_extend_refill).
void launch_extend(hipStream_t s, const SceneView& sv, const PathState& ps, const Queue& qin, uint32_t n_max,
                   const Queue& qnext, const ShadowQueue& sq, Spill sp, StageCounters sc, uint8_t* keys,
                   uint32_t* phase_counts, uint32_t bs, uint32_t* cursors = nullptr);
// Groups qin's entries by the keys extend wrote: ss.items and ss.counts form
// the segmented queue shading then reads (qin may be the identity queue).
void launch_shade_sort(hipStream_t s, const Queue& qin, uint32_t n_max, const ShadeSort& ss);
void launch_shade(hipStream_t s, uint32_t features, const SceneView& sv, const FrameParams& fp,
                  const PathState& ps, const Queue& qin, uint32_t n_max, const Queue& qout, const ShadowQueue& sq,
                  Spill sp, StageCounters sc);
// The paths left in qin after the wavefront's bounces, each traced to its
// end by one lane (k_tail): at most kTailChunks blocks per segment, a path cut
// after max_steps bounces and counted in *cut.
constexpr uint32_t kTailChunks = 16;
// cursors (see launch_extend): non-null runs k_tail_refill, lanes that take
// new paths as others end.
void launch_tail(hipStream_t s, uint32_t features, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                 const Queue& qin, Spill sp, uint32_t max_steps, unsigned long long* cut, StageCounters sc,
                 uint32_t* cursors = nullptr);
void launch_shadow(hipStream_t s, const SceneView& sv, const PathState& ps, const ShadowQueue& sq, uint32_t n_max,
                   Spill sp, StageCounters sc, uint32_t bs, uint32_t* cursors = nullptr);
// Volume-MIS shading in phases (MEDIUM|VOL_MIS): A over qin (finished and
// passing paths to qout, scattering events to qnee, their first walk to wq),
// the walks of a walk queue, B over qnee (environment maps: the second walk
// to wq), C over qnee (survivors to qout).  Same images as launch_shade.
void launch_vmis_a(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   const Queue& qin, uint32_t n_max, const Queue& qout, const Queue& qnee, const ShadowQueue& wq,
                   StageCounters sc);
void launch_walk(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                 const ShadowQueue& wq, uint32_t seg_cap, Spill sp, StageCounters sc, uint32_t bs,
                 uint32_t* cursors = nullptr);
void launch_vmis_b(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   const Queue& qnee, const ShadowQueue& wq);
void launch_vmis_c(hipStream_t s, bool transp, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   const Queue& qnee, const Queue& qout);
// Depth limits of the jobs of one wavefront (FrameParams::job_limits), passed
// by value so that a captured launch carries them.
constexpr uint32_t kMaxWaveJobs = 256;
struct JobLimits {
  uint2 v[kMaxWaveJobs];  // (max_depth, rr_depth)
};
void launch_job_limits(hipStream_t s, uint2* dst, const JobLimits& lim, uint32_t count);
// counter[0] = value, or counter[0] += value when `add` (one thread).
void launch_frame_counter(hipStream_t s, uint32_t* counter, uint32_t value, bool add);
void launch_accumulate(hipStream_t s, const FrameParams& fp, const PathState& ps, const Images& im,
                       uint32_t npaths);

// Builds the EnvTexel records of a w x h environment map.
void launch_env_texels(hipStream_t s, const float4* rgb, const float4* table, uint32_t w, uint32_t h,
                       EnvTexel* out);

// Closest hits of device-resident rays as hpt_hit records.
void launch_trace_hits(hipStream_t s, const SceneView& sv, const hpt_ray* rays, uint32_t n, hpt_hit* hits, Spill sp);

// Ray-query entry points: out = (t, u, v, id bits) per ray; counters optional.
void launch_trace_rays(hipStream_t s, const SceneView& sv, const hpt_ray* rays, uint32_t n, int any_hit,
                       float4* out, uint32_t* node_visits, uint32_t* tri_tests, Spill sp);

}  // namespace hpt
