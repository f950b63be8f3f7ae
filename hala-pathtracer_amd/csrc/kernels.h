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
};
constexpr uint32_t PATH_ANB = 1u;             // any_non_specular_bounce
constexpr uint32_t PATH_IN_MEDIUM = 2u;       // is_in_medium
constexpr uint32_t PATH_SURFACE_SCATTER = 4u; // is_surface_scatter (MEDIUM without VOL_MIS)

struct ShadowQueue {
  float4* o;  // origin, .w = tmax
  float4* d;  // direction, .w = (path << 1 | slot) bits
  float4* c;  // throughput-weighted contribution
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

constexpr int kBlock = 256;

void launch_camera(hipStream_t s, const SceneView& sv, const FrameParams& fp, const PathState& ps,
                   uint32_t npaths);
// Device counters of one stage: stats[stage*3 + {0,1,2}] = rays, node
// visits, triangle tests (64-bit).  Rays are counted whenever stats is
// non-null; visits only by the counting kernel variant (count == true).
struct StageCounters {
  unsigned long long* stats;
  bool count;
};

// q == nullptr: the queue is the identity over n_max paths.
void launch_extend(hipStream_t s, const SceneView& sv, const PathState& ps, const uint32_t* q,
                   const uint32_t* qcount, uint32_t n_max, uint32_t* zero0, uint32_t* zero1, Spill sp,
                   StageCounters sc);
void launch_shade(hipStream_t s, uint32_t features, const SceneView& sv, const FrameParams& fp,
                  const PathState& ps, const uint32_t* qin, const uint32_t* qin_count, uint32_t n_max,
                  uint32_t* qout, uint32_t* qout_count, const ShadowQueue& sq, uint32_t* sq_count, Spill sp);
void launch_shadow(hipStream_t s, const SceneView& sv, const PathState& ps, const ShadowQueue& sq,
                   const uint32_t* sq_count, uint32_t n_max, Spill sp, StageCounters sc);
void launch_accumulate(hipStream_t s, const FrameParams& fp, const PathState& ps, const Images& im,
                       uint32_t npaths);

// Ray-query entry points: out = (t, u, v, id bits) per ray; counters optional.
void launch_trace_rays(hipStream_t s, const SceneView& sv, const hpt_ray* rays, uint32_t n, int any_hit,
                       float4* out, uint32_t* node_visits, uint32_t* tri_tests, Spill sp);

}  // namespace hpt
