// api.cpp — the C ABI of include/hpt.h over the gfx950 kernels.
//
// Host-side work per call:
//  set_scene   validate and pack the std140 arrays and meshes (world-space
//              triangle vertices, shading vertices, inverse-transpose 3x3s);
//  commit      build the 8-wide BVH, upload everything, size the path state;
//  render      per frame: camera, (extend, shade, shadow) x bounces,
//              accumulate, all on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "bvh_gpu.h"
#include "common.h"
#include "exr.h"
#include "kernels.h"

namespace {

using namespace hpt;

thread_local std::string g_last_error;

struct Fail : std::runtime_error {
  hpt_status code;
  Fail(hpt_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Fail(e == hipErrorOutOfMemory ? HPT_ERROR_OUT_OF_MEMORY : HPT_ERROR_DEVICE,
               std::string(what) + ": " + hipGetErrorString(e));
  }
}

// Owning device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    if (n == bytes && p) return;
    release();
    if (n == 0) return;
    // A request the device cannot hold fails here with the sizes; a failure
    // inside the runtime's own host bookkeeping is reported as such.
    size_t free_b = 0, total_b = 0;
    check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (n > free_b)
      throw Fail(HPT_ERROR_OUT_OF_MEMORY, "device memory: " + std::to_string(n) + " B requested, " +
                                              std::to_string(free_b) + " B free of " + std::to_string(total_b));
    try {
      check(hipMalloc(&p, n), "hipMalloc");
    } catch (const std::bad_alloc&) {
      p = nullptr;
      throw Fail(HPT_ERROR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(n) + " B: host allocation inside the "
                                          "HIP runtime failed (" + std::to_string(free_b) + " B of device memory free)");
    }
    bytes = n;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
  // Takes ownership of device memory from hipMalloc.
  void adopt(void* q, size_t n) {
    release();
    p = q;
    bytes = n;
  }
};

template <typename T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

// The committed scene on the device: BVH, shading arrays, environment,
// textures, blue noise.
struct SceneDevice {
  uint32_t node_count = 0, leaf_tri_count = 0, bvh_depth = 0;
  DevBuf d_nodes, d_ltris, d_tri_shade, d_tri_key, d_prims, d_materials, d_lights, d_cameras, d_env_rgb,
      d_env_table, d_env_texels, d_tex_texels, d_tex_info, d_srgb_lut, d_blue_noise;
};

struct StageTimer {
  hipEvent_t a, b;
  int stage;
};

}  // namespace

// Each handle struct starts with a tag of its kind: an entry point given a
// handle of another kind (or of a destroyed one) fails instead of writing
// into memory that is not the struct it expects.
constexpr uint32_t kRendererTag = 0x52545048u;   // "HPTR"
constexpr uint32_t kGroupTag = 0x47545048u;      // "HPTG"
constexpr uint32_t kPresenterTag = 0x50545048u;  // "HPTP"

struct hpt_renderer {
  uint32_t tag = kRendererTag;
  hpt_create_info info{};
  hipStream_t stream = nullptr;
  // Scene, host side.
  std::vector<hpt_camera> cameras;
  uint32_t camera_count = 0;
  std::vector<hpt_light> lights;  // HPT_MAX_LIGHTS + 1 slots
  uint32_t num_lights = 0;
  std::vector<hpt_material> materials;
  std::vector<float> tri_pos;       // 9 floats per global triangle
  std::vector<TriShade> tri_shade;
  std::vector<uint32_t> tri_prim;  // primitive of each global triangle
  std::vector<PrimInfo> prims;
  bool scene_set = false;
  bool committed = false;
  // Environment.
  uint32_t env_type = HPT_ENV_SKY;
  float ground[3] = {0.5f, 0.5f, 0.5f}, sky[3] = {0.5f, 0.5f, 0.5f};
  float env_intensity = 1.0f, env_rotation = 0.0f, env_total = 0.0f;
  uint32_t env_w = 0, env_h = 0;
  std::vector<float4> env_rgb, env_table;
  std::vector<uchar4> blue_noise;
  // Material textures: one RGBA8 pool, per-texture offsets and sizes.
  std::vector<uchar4> tex_texels;
  std::vector<TexInfo> tex_info;
  // Device copy of the committed scene; shared by handles that render the
  // same scene (hpt_share_scene), freed with the last of them.
  std::shared_ptr<SceneDevice> sd;
  DevBuf d_rng, d_org, d_dir, d_thr, d_rad, d_alb, d_nrm, d_med, d_med2, d_nee_e, d_nee_l, d_hit;
  DevBuf d_qa, d_qb, d_sh_o, d_sh_d, d_sh_c, d_counts, d_spill, d_stats;
  DevBuf d_tail_cut;  // paths cut at kMaxBounces since hpt_create (one device word)
  // Phased volume-MIS shading (MEDIUM|VOL_MIS): the per-path NEE state
  // between its kernels and the queue of its scattering events; off = the
  // single shading kernel with inline walks.
  uint32_t vmis_phases = 1;
  DevBuf d_nee_dir, d_nee_li, d_nee_ld, d_nee_tr, d_qc;
  // Material-sorted shading from bounce shade_sort_from on (HPT_SHADE_SORT_OFF:
  // never): one key byte and one sorted entry per queue slot.
  uint32_t shade_sort_from = HPT_SHADE_SORT_OFF;
  DevBuf d_sort_keys, d_sort_items;
  // Bounces from tail_from on run in the tail kernel (HPT_TAIL_FROM_NONE:
  // only the paths that outlive max_depth + 1 bounces with TRANSPARENT).
  uint32_t tail_from = HPT_TAIL_FROM_NONE;
  // Threads per block of the traversal kernels (extend, walk, shadow).
  uint32_t trav_block = kDefaultTraversalBlock;
  bool refill = true;  // lane refill in the traversal kernels (hpt_set_traversal_refill)
  DevBuf d_accum, d_albedo, d_normal, d_final;
  uint32_t npix = 0;          // width * height
  uint32_t batch_cap = 0;     // frames the path buffers hold
  uint32_t frames_in_flight = 0;  // requested batch size, 0 = sized from kPathBudget
  uint64_t frame_index = 0;
  // Graph replay of whole batches (HPT_LAUNCH_GRAPH): the instantiated batch,
  // the launch parameters it was captured with, and the device frame counter
  // it reads and advances.
  uint32_t launch_mode = HPT_LAUNCH_GRAPH;
  // BVH construction at commit: the requested builder, the one that ran and its time.
  uint32_t bvh_builder = HPT_BVH_GPU, bvh_builder_used = HPT_BVH_HOST;
  double bvh_build_ms = 0.0;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> gkey;
  DevBuf d_frame;
  bool frame_synced = false;  // d_frame holds frame_index
  uint64_t graph_launches = 0, graph_captures = 0;
  // Marks the end of this handle's queued work for a group step to wait on.
  hipEvent_t ev = nullptr;
  // Profiling.
  int profiling = 0;
  std::vector<StageTimer> timers;
  size_t timers_used = 0;
  hpt_stage_stats stats[HPT_STAGE_COUNT] = {};
};

// A wavefront shared by several jobs of one scene (hpt_group_render): its own
// path state, queues and counters, sized for the largest step so far.
struct hpt_group {
  uint32_t tag = kGroupTag;
  int32_t device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // End of each of the last two steps: a step waits for the one two before it,
  // so at most two steps are queued (the host can prepare the next step while
  // the device runs the current one, and jobs admitted meanwhile join soon).
  hipEvent_t step_end[2] = {nullptr, nullptr};
  uint32_t launch_mode = HPT_LAUNCH_GRAPH;
  uint32_t shade_sort_from = HPT_SHADE_SORT_OFF;
  uint32_t tail_from = HPT_TAIL_FROM_NONE;
  uint32_t trav_block = kDefaultTraversalBlock;
  bool refill = true;  // lane refill in the traversal kernels (hpt_set_traversal_refill)
  uint32_t vmis_phases = 1;  // volume-MIS shading in phases (hpt_group_set_vmis_phases)
  size_t cap = 0;  // paths the path buffers hold
  DevBuf d_rng, d_org, d_dir, d_thr, d_rad, d_alb, d_nrm, d_med, d_med2, d_nee_e, d_nee_l, d_hit, d_job;
  DevBuf d_qa, d_qb, d_sh_o, d_sh_d, d_sh_c, d_counts, d_spill, d_sort_keys, d_sort_items, d_limits;
  DevBuf d_nee_dir, d_nee_li, d_nee_ld, d_nee_tr, d_qc;  // phased volume-MIS shading, sized like the paths
  DevBuf d_tail_cut;  // paths cut at kMaxBounces since hpt_group_create (one device word)
  // The last captured step and the key of the step before this one: a step
  // repeated with the same jobs, frames and buffers is captured and replayed.
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> gkey, last_key;
  uint64_t steps = 0, graph_launches = 0, graph_captures = 0;
};

namespace {

hpt_status fail_with(const Fail& f) {
  g_last_error = f.what();
  return f.code;
}

template <typename F>
hpt_status guarded(F&& fn) {
  try {
    fn();
    g_last_error.clear();
    return HPT_OK;
  } catch (const Fail& f) {
    return fail_with(f);
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return HPT_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return HPT_ERROR_DEVICE;
  } catch (...) {
    g_last_error = "unknown error";
    return HPT_ERROR_DEVICE;
  }
}

void need(bool c, const std::string& m, hpt_status code = HPT_ERROR_INVALID_ARGUMENT) {
  if (!c) throw Fail(code, m);
}

bool ok_renderer(const hpt_renderer* r) { return r && r->tag == kRendererTag; }
bool ok_group(const hpt_group* g) { return g && g->tag == kGroupTag; }

void set_device(hpt_renderer* r) { check(hipSetDevice(r->info.device), "hipSetDevice"); }

float luminance(float r, float g, float b) { return 0.212671f * r + 0.715160f * g + 0.072169f * b; }

// Vose's alias method in double precision with both work lists filled in
// increasing index order; returns the luminance total.
double build_alias(const std::vector<float>& lum, std::vector<float>& prob, std::vector<uint32_t>& alias) {
  const size_t n = lum.size();
  prob.assign(n, 1.0f);
  alias.resize(n);
  double total = 0.0;
  for (size_t i = 0; i < n; ++i) total += (double)lum[i];
  for (size_t i = 0; i < n; ++i) alias[i] = (uint32_t)i;
  if (!(total > 0.0)) return total;
  std::vector<double> scaled(n);
  std::vector<uint32_t> small_, large_;
  small_.reserve(n);
  large_.reserve(n);
  for (size_t i = 0; i < n; ++i) {
    scaled[i] = (double)lum[i] * (double)n / total;
    if (scaled[i] < 1.0) small_.push_back((uint32_t)i);
    else large_.push_back((uint32_t)i);
  }
  while (!small_.empty() && !large_.empty()) {
    const uint32_t s = small_.back();
    small_.pop_back();
    const uint32_t l = large_.back();
    large_.pop_back();
    prob[s] = (float)scaled[s];
    alias[s] = l;
    scaled[l] = (scaled[l] + scaled[s]) - 1.0;
    if (scaled[l] < 1.0) small_.push_back(l);
    else large_.push_back(l);
  }
  for (uint32_t i : large_) { prob[i] = 1.0f; alias[i] = i; }
  for (uint32_t i : small_) { prob[i] = 1.0f; alias[i] = i; }
  return total;
}

// 8-bit sRGB code -> linear, the IEC 61966-2-1 decode in double rounded to float.
float srgb_to_linear(int code) {
  const double c = code / 255.0;
  return (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
}

void default_blue_noise(std::vector<uchar4>& bn) {
  // 1024x1024 stand-in for assets/textures/blue_noise.png (not in the
  // reference checkout): top bytes of pcg4d(x, y, 0x6A09E667, 0xBB67AE85).
  bn.resize(1024 * 1024);
  for (uint32_t y = 0; y < 1024; ++y)
    for (uint32_t x = 0; x < 1024; ++x) {
      uint32_t v[4] = {x, y, 0x6A09E667u, 0xBB67AE85u};
      for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
      v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
      for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16u;
      v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
      bn[(size_t)y * 1024 + x] = make_uchar4((unsigned char)(v[0] >> 24), (unsigned char)(v[1] >> 24),
                                             (unsigned char)(v[2] >> 24), (unsigned char)(v[3] >> 24));
    }
}

SceneView scene_view(const hpt_renderer* r) {
  SceneView sv{};  // zeroed padding: the bytes are compared as a graph key
  sv.nodes = r->sd->d_nodes.as<Node8>();
  sv.ltris = r->sd->d_ltris.as<LeafTri>();
  sv.tri_shade = r->sd->d_tri_shade.as<TriShade>();
  sv.tri_key = r->sd->d_tri_key.as<uint8_t>();
  sv.prims = r->sd->d_prims.as<PrimInfo>();
  sv.materials = r->sd->d_materials.as<hpt_material>();
  sv.lights = r->sd->d_lights.as<hpt_light>();
  sv.cameras = r->sd->d_cameras.as<hpt_camera>();
  sv.env_rgb = r->sd->d_env_rgb.as<float4>();
  sv.env_texels = r->sd->d_env_texels.as<EnvTexel>();
  sv.blue_noise = r->sd->d_blue_noise.as<uchar4>();
  sv.tex_texels = r->sd->d_tex_texels.as<uchar4>();
  sv.tex_info = r->sd->d_tex_info.as<TexInfo>();
  sv.srgb_lut = r->sd->d_srgb_lut.as<float>();
  sv.node_count = r->sd->node_count;
  sv.tri_count = r->sd->leaf_tri_count;
  sv.num_lights = r->num_lights;
  sv.material_count = (uint32_t)r->materials.size();
  return sv;
}

PathState path_state(const hpt_renderer* r) {
  PathState ps{};
  ps.rng = r->d_rng.as<uint4>();
  ps.org = r->d_org.as<float4>();
  ps.dir = r->d_dir.as<float4>();
  ps.thr = r->d_thr.as<float4>();
  ps.rad = r->d_rad.as<float4>();
  ps.albedo = r->d_alb.as<float4>();
  ps.normal = r->d_nrm.as<float4>();
  ps.med = r->d_med.as<float4>();
  ps.med2 = r->d_med2.as<float2>();
  ps.nee_env = r->d_nee_e.as<float4>();
  ps.nee_light = r->d_nee_l.as<float4>();
  ps.hit = r->d_hit.as<float4>();
  ps.nee_dir = r->d_nee_dir.as<float4>();
  ps.nee_li = r->d_nee_li.as<float4>();
  ps.nee_ld = r->d_nee_ld.as<float4>();
  ps.nee_tr = r->d_nee_tr.as<float4>();
  return ps;
}

// Traversal stack entries beyond the LDS part: needed when the tree has
// more levels than the short stack holds.
Spill ensure_spill(hpt_renderer* r, uint32_t threads) {
  Spill sp{nullptr, 0};
  if (r->sd->bvh_depth <= (uint32_t)kShortStack) return sp;
  const size_t entries = (size_t)(r->sd->bvh_depth - kShortStack) * threads;
  if (r->d_spill.bytes < entries * sizeof(uint2)) r->d_spill.alloc(entries * sizeof(uint2));
  sp.buf = r->d_spill.as<uint2>();
  sp.stride = threads;
  return sp;
}

uint32_t round_threads(uint32_t n) { return ((n + kBlock - 1) / kBlock) * kBlock; }

void stage_begin(hpt_renderer* r, int stage, StageTimer*& t) {
  t = nullptr;
  if (!(r->profiling & 1)) return;
  if (r->timers_used == r->timers.size()) {
    StageTimer st;
    check(hipEventCreate(&st.a), "hipEventCreate");
    check(hipEventCreate(&st.b), "hipEventCreate");
    r->timers.push_back(st);
  }
  t = &r->timers[r->timers_used++];
  t->stage = stage;
  check(hipEventRecord(t->a, r->stream), "hipEventRecord");
}

void stage_end(hpt_renderer* r, StageTimer* t) {
  if (!t) return;
  check(hipEventRecord(t->b, r->stream), "hipEventRecord");
  r->stats[t->stage].launches++;
}

void collect_timers(hpt_renderer* r) {
  if (!r->timers_used) return;
  check(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
  for (size_t i = 0; i < r->timers_used; ++i) {
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, r->timers[i].a, r->timers[i].b), "hipEventElapsedTime");
    r->stats[r->timers[i].stage].total_ms += ms;
  }
  r->timers_used = 0;
}

FrameParams frame_params(const hpt_renderer* r, uint32_t camera_index) {
  FrameParams fp{};
  for (int i = 0; i < 3; ++i) {
    fp.ground[i] = r->ground[i];
    fp.sky[i] = r->sky[i];
  }
  fp.width = r->info.width;
  fp.height = r->info.height;
  fp.res_x = (float)r->info.width;
  fp.res_y = (float)r->info.height;
  fp.max_depth = r->info.max_depth;
  fp.rr_depth = r->info.rr_depth;
  fp.frame_index = (uint32_t)r->frame_index;
  fp.batch_frames = 1;
  fp.pixels = r->npix;
  fp.camera_index = camera_index;
  fp.env_type = r->env_type;
  fp.env_w = r->env_w;
  fp.env_h = r->env_h;
  fp.env_total_sum = r->env_total;
  fp.env_rotation = r->env_rotation;
  fp.env_intensity = r->env_intensity;
  fp.exposure_value = r->info.exposure_value;
  fp.enable_tonemap = r->info.enable_tonemap;
  fp.enable_aces = r->info.enable_aces;
  fp.use_simple_aces = r->info.use_simple_aces;
  fp.num_lights = r->num_lights;
  fp.tan_half_fov = std::tan(r->cameras[camera_index].yfov * 0.5f);
  return fp;
}

// One job's part of a wavefront: its frame parameters (camera, frames,
// accumulation), its images and its slice [base, base + npaths) of the
// path state.
struct WaveJob {
  FrameParams fp;
  Images im;
  uint32_t base, npaths;
};

// The device buffers one wavefront runs in.
struct WaveBuffers {
  PathState ps;
  uint32_t *qa, *qb, *counts;
  uint32_t* qc;       // the phased volume-MIS shading's event queue (null: not phased)
  float4 *sh_o, *sh_d, *sh_c;
  Spill sp;
  uint8_t* sort_keys;
  uint32_t* sort_items;
  unsigned long long* tail_cut;  // paths the tail cut at kMaxBounces (one device word)
  bool refill = true;  // traversal lanes refill from their queues (k_extend_refill, k_walk_refill)
};

// Bounces (closest-hit queries) a path may take; the tail cuts a path that
// passes through transparent surfaces beyond this and counts it.
constexpr uint32_t kMaxBounces = 4096;

// Camera paths of every job, the bounce loop over all n paths at once
// (extend, shade with shade_fp, shadow), then every job's accumulation.
// `timing`: the handle whose stage timers and counters bracket the
// launches, or null.
void trace_wavefront(hipStream_t s, const SceneView& sv, uint32_t features, const FrameParams& shade_fp,
                     const WaveBuffers& wb, const std::vector<WaveJob>& jobs, uint32_t n, uint32_t fixed_bounces,
                     uint32_t sort_from, uint32_t tail_from, uint32_t trav_block, hpt_renderer* timing) {
  const SegGeom g = seg_geom(n);
  uint32_t* counts = wb.counts;
  // Segment counters: path queues A and B (ping-pong), the shadow queue, the
  // sorted shading queue, then the sort's key counts and cursors.
  const Queue qbuf[2] = {{wb.qa, counts, g.seg_cap}, {wb.qb, counts + kSeg * kSegStride, g.seg_cap}};
  const ShadowQueue sq{wb.sh_o, wb.sh_d, wb.sh_c, counts + 2 * kSeg * kSegStride, 2 * g.seg_cap};
  const ShadeSort ss{wb.sort_keys, wb.sort_items, counts + 3 * kSeg * kSegStride, counts + 4 * kSeg * kSegStride};
  const PathState& ps = wb.ps;
  const Spill sp = wb.sp;
  StageTimer* t = nullptr;
  auto begin = [&](int stage) {
    if (timing) stage_begin(timing, stage, t);
  };
  auto end = [&] {
    if (timing) stage_end(timing, t);
  };
  auto* stats = timing && timing->profiling ? timing->d_stats.as<unsigned long long>() : nullptr;
  const bool count = timing && (timing->profiling & HPT_PROFILE_COUNT_VISITS) != 0;
  const StageCounters sc_extend{stats ? stats + kStatsPerStage * HPT_STAGE_EXTEND : nullptr, count};
  const StageCounters sc_shadow{stats ? stats + kStatsPerStage * HPT_STAGE_SHADOW : nullptr, count};
  const StageCounters sc_shade{stats ? stats + kStatsPerStage * HPT_STAGE_SHADE : nullptr, count};
  const StageCounters sc_walk{stats ? stats + kStatsPerStage * HPT_STAGE_WALK : nullptr, count};
  const StageCounters sc_tail{stats ? stats + kStatsPerStage * HPT_STAGE_TAIL : nullptr, count};

  begin(HPT_STAGE_CAMERA);
  for (size_t j = 0; j < jobs.size(); ++j)
    launch_camera(s, sv, jobs[j].fp, path_slice(ps, jobs[j].base), jobs[j].npaths, (uint32_t)j);
  end();
  check(hipGetLastError(), "camera launch");

  Queue qin{nullptr, nullptr, g.seg_cap};  // the first bounce walks the camera paths in order
  int out = 0;
  const bool transparent = (features & HPT_FEATURE_TRANSPARENT) != 0;
  const bool vmis = (features & HPT_FEATURE_MEDIUM) && (features & HPT_FEATURE_VOL_MIS);
  // Phased volume-MIS shading: scattering events queue (qnee), the first
  // walk's queue in the shadow queue's first half of each segment (sq's
  // counters), the second walk's in the other half.
  const bool phased = vmis && wb.qc && ps.nee_dir;
  uint32_t* phase_counts = phased ? counts + 4 * kSeg * kSegStride + 2 * kSeg * kSortKeys : nullptr;
  const Queue qnee{wb.qc, phase_counts, g.seg_cap};
  const ShadowQueue wq1{wb.sh_o + g.seg_cap, wb.sh_d + g.seg_cap, wb.sh_c + g.seg_cap,
                        phased ? phase_counts + kSeg * kSegStride : nullptr, 2 * g.seg_cap};
  // Bounces [0, wave_bounces) run as wavefront launches; the paths left then
  // go to the tail kernel.
  const uint32_t wave_bounces = std::max<uint32_t>(1u, std::min(fixed_bounces, tail_from));
  for (uint32_t b = 0; b < wave_bounces; ++b) {
    const Queue& qout = qbuf[out];
    const bool sort = b >= sort_from && ss.keys && ss.items;
    begin(HPT_STAGE_EXTEND);
    launch_extend(s, sv, ps, qin, n, qout, sq, sp, sc_extend, sort ? ss.keys : nullptr, phase_counts, trav_block,
                  wb.refill ? counts + kCursorOffset : nullptr);
    end();
    Queue qshade = qin;
    if (sort) {
      begin(HPT_STAGE_SORT);
      launch_shade_sort(s, qin, n, ss);
      end();
      qshade = Queue{ss.items, ss.counts, g.seg_cap};
    }
    if (phased) {
      begin(HPT_STAGE_SHADE);
      launch_vmis_a(s, transparent, sv, shade_fp, ps, qshade, n, qout, qnee, sq, sc_shade);
      end();
      begin(HPT_STAGE_WALK);
      launch_walk(s, transparent, sv, shade_fp, ps, sq, g.seg_cap, sp, sc_walk, trav_block,
                  wb.refill ? counts + kCursorOffset : nullptr);
      end();
      if (shade_fp.env_type != HPT_ENV_SKY) {
        begin(HPT_STAGE_SHADE);
        launch_vmis_b(s, transparent, sv, shade_fp, ps, qnee, wq1);
        end();
        begin(HPT_STAGE_WALK);
        launch_walk(s, transparent, sv, shade_fp, ps, wq1, g.seg_cap, sp, sc_walk, trav_block,
                    wb.refill ? counts + kCursorOffset : nullptr);
        end();
      }
      begin(HPT_STAGE_SHADE);
      launch_vmis_c(s, transparent, sv, shade_fp, ps, qnee, qout);
      end();
    } else {
      begin(HPT_STAGE_SHADE);
      launch_shade(s, features, sv, shade_fp, ps, qshade, n, qout, sq, sp, sc_shade);
      end();
    }
    if (!vmis) {
      begin(HPT_STAGE_SHADOW);
      launch_shadow(s, sv, ps, sq, n, sp, sc_shadow, trav_block, wb.refill ? counts + kCursorOffset : nullptr);
      end();
    }
    check(hipGetLastError(), "bounce launch");
    qin = qout;
    out ^= 1;
  }
  if (transparent || wave_bounces < fixed_bounces) {
    // The paths still queued run to their end in the tail kernel, on the
    // device, with no host round trip: those past a tail_from bounce, and
    // with TRANSPARENT those that outlive max_depth + 1 bounces
    // (pass-throughs do not count toward depth, pathtracer.rgen.glsl:196-200).
    begin(HPT_STAGE_TAIL);
    launch_tail(s, features, sv, shade_fp, ps, qin, sp, wave_bounces < kMaxBounces ? kMaxBounces - wave_bounces : 0u,
                wb.tail_cut, sc_tail, wb.refill ? counts + kCursorOffset : nullptr);
    end();
    check(hipGetLastError(), "tail launch");
  }
  begin(HPT_STAGE_ACCUMULATE);
  for (const WaveJob& j : jobs) launch_accumulate(s, j.fp, path_slice(ps, j.base), j.im, j.fp.pixels);
  end();
  check(hipGetLastError(), "accumulate launch");
}

Images images_of(const hpt_renderer* r) {
  return Images{r->d_accum.as<float4>(), r->d_albedo.as<float4>(), r->d_normal.as<float4>(), r->d_final.as<uchar4>()};
}

// Traces `frames` consecutive frames (starting at r->frame_index, or at the
// device counter `frame_ptr` when given) as one wavefront of
// frames * width * height paths.
void render_batch(hpt_renderer* r, uint32_t camera_index, uint32_t frames, const uint32_t* frame_ptr = nullptr) {
  FrameParams fp = frame_params(r, camera_index);
  fp.batch_frames = frames;
  if (frame_ptr) {
    fp.frame_index = 0;
    fp.frame_ptr = frame_ptr;
  }
  const uint32_t n = frames * r->npix;
  WaveBuffers wb;
  wb.ps = path_state(r);
  wb.qa = r->d_qa.as<uint32_t>();
  wb.qb = r->d_qb.as<uint32_t>();
  wb.counts = r->d_counts.as<uint32_t>();
  wb.sh_o = r->d_sh_o.as<float4>();
  wb.sh_d = r->d_sh_d.as<float4>();
  wb.sh_c = r->d_sh_c.as<float4>();
  wb.sp = ensure_spill(r, 2 * kSeg * seg_geom(n).seg_cap);  // threads of the widest launch (shadow)
  wb.sort_keys = r->d_sort_keys.as<uint8_t>();
  wb.sort_items = r->d_sort_items.as<uint32_t>();
  wb.tail_cut = r->d_tail_cut.as<unsigned long long>();
  wb.qc = r->vmis_phases ? r->d_qc.as<uint32_t>() : nullptr;
  wb.refill = r->refill;
  const std::vector<WaveJob> jobs{WaveJob{fp, images_of(r), 0, n}};
  trace_wavefront(r->stream, scene_view(r), r->info.features, fp, wb, jobs, n, r->info.max_depth + 1,
                  r->shade_sort_from, r->tail_from, r->trav_block, r);
}

template <typename T>
void key_append(std::vector<uint8_t>& k, const T& v) {
  const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

// Everything a captured batch bakes into its kernel arguments: the scene,
// frame parameters (but the first frame, which the device counter supplies),
// path state and queue buffers, batch size and permutation.
std::vector<uint8_t> graph_key(hpt_renderer* r, uint32_t camera_index, uint32_t frames) {
  std::vector<uint8_t> k;
  key_append(k, scene_view(r));
  FrameParams fp = frame_params(r, camera_index);
  fp.frame_index = 0;
  key_append(k, fp);
  key_append(k, path_state(r));
  for (const DevBuf* b : {&r->d_qa, &r->d_qb, &r->d_sh_o, &r->d_sh_d, &r->d_sh_c, &r->d_counts, &r->d_spill,
                          &r->d_accum, &r->d_albedo, &r->d_normal, &r->d_final, &r->d_frame, &r->d_sort_keys,
                          &r->d_sort_items, &r->d_qc, &r->d_tail_cut})
    key_append(k, b->p);
  key_append(k, r->vmis_phases);
  key_append(k, frames);
  key_append(k, r->info.features);
  key_append(k, r->shade_sort_from);
  key_append(k, r->tail_from);
  key_append(k, r->trav_block);
  key_append(k, r->refill);
  return k;
}

// One full batch as a graph replay: captured on first use (and again when
// any captured parameter changed), then launched with one call.  The graph
// ends by advancing the device frame counter, so replays need no updates.
void render_batch_graph(hpt_renderer* r, uint32_t camera_index, uint32_t frames) {
  hipStream_t s = r->stream;
  if (!r->d_frame.p) r->d_frame.alloc(256);
  const uint32_t n = frames * r->npix;
  (void)ensure_spill(r, 2 * kSeg * seg_geom(n).seg_cap);  // any allocation happens before capture
  std::vector<uint8_t> key = graph_key(r, camera_index, frames);
  if (!r->gexec || key != r->gkey) {
    if (r->gexec) {
      check(hipStreamSynchronize(s), "graph");
      check(hipGraphExecDestroy(r->gexec), "hipGraphExecDestroy");
      r->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      render_batch(r, camera_index, frames, r->d_frame.as<uint32_t>());
      launch_frame_counter(s, r->d_frame.as<uint32_t>(), frames, true);
      check(hipGetLastError(), "frame counter launch");
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(&r->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    check(e, "hipGraphInstantiate");
    r->gkey.swap(key);
    r->graph_captures++;
  }
  if (!r->frame_synced) {
    launch_frame_counter(s, r->d_frame.as<uint32_t>(), (uint32_t)r->frame_index, false);
    check(hipGetLastError(), "frame counter launch");
    r->frame_synced = true;
  }
  check(hipGraphLaunch(r->gexec, s), "hipGraphLaunch");
  r->graph_launches++;
}

// Device memory a batch may take for path state and queues; 1080p fits 32
// frames in about 19 GB of the 288 GB.
constexpr size_t kPathBudget = size_t(24) << 30;
constexpr uint32_t kMaxBatchFrames = 64;
constexpr uint32_t kAutoBatchFrames = 32;
// Bytes per path: 11 float4 path fields, the float2 medium field, two queue
// slots and two shadow-queue entries of 3 float4.
constexpr size_t kPathBytes = 11 * 16 + 8 + 2 * 4 + 2 * 48;  // (+ segment slack)

// Phased volume-MIS shading: four float4 of NEE state and an event-queue slot.
constexpr size_t kPhaseBytes = 4 * 16 + 4;

bool vmis_features(uint32_t features) {
  return (features & HPT_FEATURE_MEDIUM) && (features & HPT_FEATURE_VOL_MIS);
}

uint32_t batch_frames_for(const hpt_renderer* r) {
  if (r->frames_in_flight) return r->frames_in_flight;
  const size_t per_frame = (size_t)r->npix * (kPathBytes + (vmis_features(r->info.features) ? kPhaseBytes : 0));
  size_t k = kPathBudget / (per_frame ? per_frame : 1);
  if (k > kAutoBatchFrames) k = kAutoBatchFrames;
  if (k < 1) k = 1;
  while (k > 1 && (uint64_t)k * r->npix * 2 > 0xFFFFFFFFull) --k;  // queue indices are 32-bit
  return (uint32_t)k;
}

// Sort keys and sorted entries for `frames` frames per batch while shading
// sort is on; released when it is off.
void alloc_sort(hpt_renderer* r, uint32_t frames) {
  if (r->shade_sort_from == HPT_SHADE_SORT_OFF) {
    r->d_sort_keys.release();
    r->d_sort_items.release();
    return;
  }
  const size_t qn = (size_t)kSeg * seg_geom(frames * r->npix).seg_cap;
  r->d_sort_keys.alloc(qn);
  r->d_sort_items.alloc(qn * 4);
}

// Sizes the path state and the queues for `frames` frames per batch.
void alloc_paths(hpt_renderer* r, uint32_t frames) {
  const size_t n = (size_t)frames * r->npix;
  const size_t f4 = n * 16;
  for (DevBuf* b : {&r->d_rng, &r->d_org, &r->d_dir, &r->d_thr, &r->d_rad, &r->d_alb, &r->d_nrm, &r->d_med,
                    &r->d_nee_e, &r->d_nee_l, &r->d_hit})
    b->alloc(f4);
  r->d_med2.alloc(n * 8);
  const size_t qn = (size_t)kSeg * seg_geom((uint32_t)n).seg_cap;  // segment capacity, >= n
  r->d_qa.alloc(qn * 4);
  r->d_qb.alloc(qn * 4);
  r->d_sh_o.alloc(2 * qn * 16);
  r->d_sh_d.alloc(2 * qn * 16);
  r->d_sh_c.alloc(2 * qn * 16);
  if (vmis_features(r->info.features)) {
    for (DevBuf* b : {&r->d_nee_dir, &r->d_nee_li, &r->d_nee_ld, &r->d_nee_tr}) b->alloc(f4);
    r->d_qc.alloc(qn * 4);
  } else {
    for (DevBuf* b : {&r->d_nee_dir, &r->d_nee_li, &r->d_nee_ld, &r->d_nee_tr, &r->d_qc}) b->release();
  }
  alloc_sort(r, frames);
  if (!r->d_tail_cut.p) {
    r->d_tail_cut.alloc(256);
    check(hipMemsetAsync(r->d_tail_cut.p, 0, r->d_tail_cut.bytes, r->stream), "hipMemsetAsync");
  }
  r->batch_cap = frames;
}

// Uploads the environment texture and builds its EnvTexel records (the sky
// keeps a 1x1 black map so the pointers stay valid).
void upload_env(hpt_renderer* r, hipStream_t s) {
  if (r->env_rgb.empty()) {
    r->env_rgb.assign(1, make_float4(0, 0, 0, 0));
    r->env_table.assign(1, make_float4(1, 0, 0, 0));
  }
  const uint32_t w = r->env_type == HPT_ENV_MAP ? r->env_w : 1, h = r->env_type == HPT_ENV_MAP ? r->env_h : 1;
  upload(r->sd->d_env_rgb, r->env_rgb, s);
  upload(r->sd->d_env_table, r->env_table, s);
  r->sd->d_env_texels.alloc((size_t)w * h * sizeof(EnvTexel));
  launch_env_texels(s, r->sd->d_env_rgb.as<float4>(), r->sd->d_env_table.as<float4>(), w, h, r->sd->d_env_texels.as<EnvTexel>());
  check(hipGetLastError(), "environment records");
}

// Path state, queues, counters and images of a committed handle, zeroed.
void alloc_frame_state(hpt_renderer* r) {
  hipStream_t s = r->stream;
  const uint32_t n = r->info.width * r->info.height;
  r->npix = n;
  alloc_paths(r, batch_frames_for(r));
  const size_t f4 = (size_t)n * 16;
  r->d_counts.alloc(kCountWords * sizeof(uint32_t));
  check(hipMemsetAsync(r->d_counts.p, 0, r->d_counts.bytes, s), "hipMemsetAsync");
  r->d_accum.alloc(f4);
  r->d_albedo.alloc(f4);
  r->d_normal.alloc(f4);
  r->d_final.alloc((size_t)n * 4);
  check(hipMemsetAsync(r->d_accum.p, 0, f4, s), "hipMemsetAsync");
  check(hipMemsetAsync(r->d_albedo.p, 0, f4, s), "hipMemsetAsync");
  check(hipMemsetAsync(r->d_normal.p, 0, f4, s), "hipMemsetAsync");
  check(hipMemsetAsync(r->d_final.p, 0, (size_t)n * 4, s), "hipMemsetAsync");
}

void require_committed(const hpt_renderer* r) {
  need(r->committed, "hpt_commit has not been called since the last scene change", HPT_ERROR_STATE);
}

}  // namespace

namespace {

ExrImage decode_exr_or_throw(const void* data, size_t size, bool pixels) {
  ExrImage img;
  std::string err;
  const ExrResult res = decode_exr(static_cast<const uint8_t*>(data), size, pixels, img, err);
  if (res != kExrOk)
    throw Fail(res == kExrUnsupported ? HPT_ERROR_UNSUPPORTED : HPT_ERROR_INVALID_ARGUMENT, "OpenEXR: " + err);
  return img;
}

}  // namespace

namespace {

// Sizes a group's buffers for n paths (grow only, with a quarter of slack so
// that steps of slowly growing size do not reallocate each time).
void group_alloc(hpt_group* g, uint32_t n, uint32_t bvh_depth, bool vmis) {
  // Buffers are freed or replaced only once the queued steps that use them end.
  const bool sort = g->shade_sort_from != HPT_SHADE_SORT_OFF;
  const size_t spill_need = bvh_depth > (uint32_t)kShortStack
                                ? (size_t)(bvh_depth - kShortStack) * 2 * kSeg *
                                      seg_geom((uint32_t)std::max<size_t>(n, g->cap)).seg_cap * sizeof(uint2)
                                : 0;
  if (n > g->cap || sort != (g->d_sort_keys.p != nullptr) || g->d_spill.bytes < spill_need ||
      vmis != (g->d_qc.p != nullptr))
    check(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
  if (n > g->cap) {
    const size_t cap = std::max<size_t>(n, std::min<size_t>(g->cap + g->cap / 4, 0x7FFFFFFFu));
    const size_t f4 = cap * 16;
    for (DevBuf* b : {&g->d_rng, &g->d_org, &g->d_dir, &g->d_thr, &g->d_rad, &g->d_alb, &g->d_nrm, &g->d_med,
                      &g->d_nee_e, &g->d_nee_l, &g->d_hit})
      b->alloc(f4);
    g->d_med2.alloc(cap * 8);
    g->d_job.alloc(cap * 2);
    const size_t qn = (size_t)kSeg * seg_geom((uint32_t)cap).seg_cap;
    g->d_qa.alloc(qn * 4);
    g->d_qb.alloc(qn * 4);
    g->d_sh_o.alloc(2 * qn * 16);
    g->d_sh_d.alloc(2 * qn * 16);
    g->d_sh_c.alloc(2 * qn * 16);
    g->d_sort_keys.release();
    g->d_sort_items.release();
    for (DevBuf* b : {&g->d_nee_dir, &g->d_nee_li, &g->d_nee_ld, &g->d_nee_tr, &g->d_qc}) b->release();
    g->cap = cap;
  }
  if (vmis && !g->d_qc.p) {
    for (DevBuf* b : {&g->d_nee_dir, &g->d_nee_li, &g->d_nee_ld, &g->d_nee_tr}) b->alloc(g->cap * 16);
    g->d_qc.alloc((size_t)kSeg * seg_geom((uint32_t)g->cap).seg_cap * 4);
  } else if (!vmis && g->d_qc.p) {
    for (DevBuf* b : {&g->d_nee_dir, &g->d_nee_li, &g->d_nee_ld, &g->d_nee_tr, &g->d_qc}) b->release();
  }
  const size_t qn = (size_t)kSeg * seg_geom((uint32_t)g->cap).seg_cap;
  if (g->shade_sort_from != HPT_SHADE_SORT_OFF) {
    if (!g->d_sort_keys.p) g->d_sort_keys.alloc(qn);
    if (!g->d_sort_items.p) g->d_sort_items.alloc(qn * 4);
  } else {
    g->d_sort_keys.release();
    g->d_sort_items.release();
  }
  if (bvh_depth > (uint32_t)kShortStack) {
    const size_t entries = (size_t)(bvh_depth - kShortStack) * 2 * qn;  // threads of a shadow launch
    if (g->d_spill.bytes < entries * sizeof(uint2)) g->d_spill.alloc(entries * sizeof(uint2));
  }
  if (!g->d_counts.p) {
    g->d_counts.alloc(kCountWords * sizeof(uint32_t));
    check(hipMemsetAsync(g->d_counts.p, 0, g->d_counts.bytes, g->stream), "hipMemsetAsync");
  }
  if (!g->d_limits.p) g->d_limits.alloc(kMaxWaveJobs * sizeof(uint2));
  if (!g->d_tail_cut.p) {
    g->d_tail_cut.alloc(256);
    check(hipMemsetAsync(g->d_tail_cut.p, 0, g->d_tail_cut.bytes, g->stream), "hipMemsetAsync");
  }
}

// Environment and light settings shading reads once per launch: equal for
// every job of a group step.
bool same_environment(const hpt_renderer* a, const hpt_renderer* b) {
  return a->env_type == b->env_type && a->env_w == b->env_w && a->env_h == b->env_h &&
         a->env_total == b->env_total && a->env_rotation == b->env_rotation &&
         a->env_intensity == b->env_intensity && a->num_lights == b->num_lights &&
         std::memcmp(a->ground, b->ground, sizeof a->ground) == 0 && std::memcmp(a->sky, b->sky, sizeof a->sky) == 0;
}

void ensure_event(hipEvent_t& e) {
  if (!e) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
}

}  // namespace

extern "C" {

hpt_status hpt_create(const hpt_create_info* info, hpt_renderer** out) {
  return guarded([&] {
    need(info && out, "null argument");
    *out = nullptr;
    need(info->width > 0 && info->height > 0, "width and height must be non-zero");
    need((uint64_t)info->width * info->height <= (1u << 30), "image too large");
    need(info->rr_depth <= info->max_depth, "the Russian Roulette depth is greater than the max depth");
    need((info->features & ~7u) == 0, "unknown feature bits");
    int ndev = 0;
    check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    need(info->device >= 0 && info->device < ndev, "no such HIP device", HPT_ERROR_DEVICE);
    auto* r = new hpt_renderer();
    r->info = *info;
    try {
      set_device(r);
      check(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking), "hipStreamCreate");
    } catch (...) {
      delete r;
      throw;
    }
    *out = r;
  });
}

hpt_status hpt_set_scene(hpt_renderer* r, const hpt_scene_desc* d) {
  return guarded([&] {
    need(ok_renderer(r) && d, "null argument");
    need(d->camera_count >= 1 && d->camera_count <= HPT_MAX_CAMERAS && d->cameras, "1..8 cameras required");
    need(d->light_count <= HPT_MAX_LIGHTS, "at most 16 lights");
    need(d->light_count == 0 || d->lights, "null lights");
    need(d->material_count == 0 || d->materials, "null materials");
    need(d->primitive_count == 0 || d->primitives, "null primitives");
    for (uint32_t i = 0; i < d->camera_count; ++i)
      need(d->cameras[i].type == HPT_CAMERA_PERSPECTIVE || d->cameras[i].type == HPT_CAMERA_ORTHOGRAPHIC,
           "unknown camera type");
    for (uint32_t i = 0; i < d->light_count; ++i)
      need(d->lights[i].type >= HPT_LIGHT_POINT && d->lights[i].type <= HPT_LIGHT_SPHERE, "unknown light type");
    for (uint32_t i = 0; i < d->material_count; ++i) {
      const hpt_material& m = d->materials[i];
      need(m.type == HPT_MATERIAL_DIFFUSE || m.type == HPT_MATERIAL_DISNEY, "unknown material type");
      need(m.medium.type <= HPT_MEDIUM_EMISSIVE, "unknown medium type");
    }
    std::vector<float> tri_pos;
    std::vector<TriShade> tri_shade;
    std::vector<uint32_t> tri_prim;
    std::vector<PrimInfo> prims;
    uint64_t ntri = 0, nvtx = 0;
    for (uint32_t p = 0; p < d->primitive_count; ++p) {
      const hpt_primitive_desc& pd = d->primitives[p];
      need(pd.index_count % 3 == 0, "index_count must be a multiple of 3");
      need(pd.vertex_count == 0 || pd.vertices, "null vertices");
      need(pd.index_count == 0 || pd.indices, "null indices");
      need(pd.material_index == HPT_INVALID_INDEX || pd.material_index < d->material_count,
           "material index out of range");
      ntri += pd.index_count / 3;
      nvtx += pd.vertex_count;
    }
    need(ntri < (1u << 31) && nvtx < (1u << 31), "too many triangles or vertices");
    tri_pos.reserve(ntri * 9);
    tri_shade.reserve(ntri);
    tri_prim.reserve(ntri);
    uint32_t tri_base = 0;
    for (uint32_t p = 0; p < d->primitive_count; ++p) {
      const hpt_primitive_desc& pd = d->primitives[p];
      const float* m = pd.transform;
      // transpose(inverse(mat3(transform))) (triangles.rchit.glsl:21): for
      // columns a, b, c its columns are cross(b,c), cross(c,a), cross(a,b) / det.
      const f3 a = mk3(m[0], m[1], m[2]), b = mk3(m[4], m[5], m[6]), c = mk3(m[8], m[9], m[10]);
      const float det = dot(a, cross(b, c));
      const f3 i0 = cross(b, c) / det, i1 = cross(c, a) / det, i2 = cross(a, b) / det;
      PrimInfo pi{};
      const float it[9] = {i0.x, i0.y, i0.z, i1.x, i1.y, i1.z, i2.x, i2.y, i2.z};
      std::memcpy(pi.it, it, sizeof it);
      pi.material_index = pd.material_index;
      pi.local_base = tri_base;
      prims.push_back(pi);
      std::vector<float> wp((size_t)pd.vertex_count * 3);
      for (uint32_t k = 0; k < pd.vertex_count; ++k) {
        const hpt_vertex& v = pd.vertices[k];
        const float x = v.position[0], y = v.position[1], z = v.position[2];
        // world = transform * vec4(p, 1), summed left to right.
        wp[k * 3 + 0] = ((m[0] * x + m[4] * y) + m[8] * z) + m[12];
        wp[k * 3 + 1] = ((m[1] * x + m[5] * y) + m[9] * z) + m[13];
        wp[k * 3 + 2] = ((m[2] * x + m[6] * y) + m[10] * z) + m[14];
      }
      for (uint32_t t = 0; t < pd.index_count / 3; ++t) {
        uint32_t ix[3];
        for (int k = 0; k < 3; ++k) {
          ix[k] = pd.indices[t * 3 + k];
          need(ix[k] < pd.vertex_count, "vertex index out of range");
          for (int a2 = 0; a2 < 3; ++a2) tri_pos.push_back(wp[(size_t)ix[k] * 3 + a2]);
        }
        TriShade ts{};
        for (int k = 0; k < 3; ++k) {
          const hpt_vertex& v = pd.vertices[ix[k]];
          std::memcpy(ts.n[k], v.normal, 12);
          std::memcpy(ts.t[k], v.tangent, 12);
          std::memcpy(ts.uv[k], v.tex_coord, 8);
        }
        ts.prim = p;
        ts.material = pd.material_index;
        tri_shade.push_back(ts);
        tri_prim.push_back(p);
      }
      tri_base += pd.index_count / 3;
    }
    r->cameras.assign(d->cameras, d->cameras + d->camera_count);
    r->cameras.resize(HPT_MAX_CAMERAS, hpt_camera{});  // unused slots zero-filled
    r->lights.assign(HPT_MAX_LIGHTS + 1, hpt_light{});
    for (uint32_t i = 0; i < d->light_count; ++i) r->lights[i] = d->lights[i];
    r->num_lights = d->light_count;
    r->materials.assign(d->materials, d->materials + d->material_count);
    if (r->materials.empty()) r->materials.push_back(hpt_material{});
    r->tri_pos.swap(tri_pos);
    r->tri_shade.swap(tri_shade);
    r->tri_prim.swap(tri_prim);
    r->prims.swap(prims);
    r->camera_count = d->camera_count;
    r->scene_set = true;
    r->committed = false;
  });
}

hpt_status hpt_exr_info(const void* data, size_t size, uint32_t* width, uint32_t* height) {
  return guarded([&] {
    need(data && width && height, "null argument");
    const ExrImage img = decode_exr_or_throw(data, size, false);
    *width = img.width;
    *height = img.height;
  });
}

hpt_status hpt_exr_decode_rgb(const void* data, size_t size, float* rgb, uint32_t width, uint32_t height) {
  return guarded([&] {
    need(data && rgb, "null argument");
    const ExrImage img = decode_exr_or_throw(data, size, true);
    need(img.width == width && img.height == height,
         "OpenEXR image is " + std::to_string(img.width) + "x" + std::to_string(img.height) + ", not " +
             std::to_string(width) + "x" + std::to_string(height));
    std::memcpy(rgb, img.rgb.data(), img.rgb.size() * sizeof(float));
  });
}

hpt_status hpt_set_env_map(hpt_renderer* r, const float* rgb, uint32_t w, uint32_t h, float rotation) {
  return guarded([&] {
    need(ok_renderer(r) && rgb && w > 0 && h > 0, "invalid environment map");
    need((uint64_t)w * h < (1ull << 31), "environment map too large");
    const size_t n = (size_t)w * h;
    std::vector<float4> texels(n);
    std::vector<float> lum(n);
    for (size_t i = 0; i < n; ++i) {
      const float a = rgb[i * 3], b = rgb[i * 3 + 1], c = rgb[i * 3 + 2];
      need(std::isfinite(a) && std::isfinite(b) && std::isfinite(c) && a >= 0 && b >= 0 && c >= 0,
           "environment texels must be finite and non-negative");
      texels[i] = make_float4(a, b, c, 0.0f);
      lum[i] = luminance(a, b, c);
    }
    std::vector<float> prob;
    std::vector<uint32_t> alias;
    const double total = build_alias(lum, prob, alias);
    std::vector<float4> table(n);
    for (size_t i = 0; i < n; ++i) {
      uint32_t al = alias[i];
      float alf;
      std::memcpy(&alf, &al, 4);
      table[i] = make_float4(prob[i], alf, lum[i], lum[al]);
    }
    r->env_rgb.swap(texels);
    r->env_table.swap(table);
    r->env_total = (float)total;
    r->env_w = w;
    r->env_h = h;
    need(!(r->committed && r->sd.use_count() > 1),
         "the environment map belongs to a scene shared with other handles; set it before sharing",
         HPT_ERROR_STATE);
    r->env_rotation = rotation;
    r->env_type = HPT_ENV_MAP;
    if (r->committed) {
      set_device(r);
      upload_env(r, r->stream);
    }
  });
}

hpt_status hpt_set_ground_color(hpt_renderer* r, const float rgba[4]) {
  return guarded([&] {
    need(ok_renderer(r) && rgba, "null argument");
    for (int i = 0; i < 3; ++i) r->ground[i] = rgba[i];
    r->env_type = HPT_ENV_SKY;
  });
}

hpt_status hpt_set_sky_color(hpt_renderer* r, const float rgba[4]) {
  return guarded([&] {
    need(ok_renderer(r) && rgba, "null argument");
    for (int i = 0; i < 3; ++i) r->sky[i] = rgba[i];
    r->env_type = HPT_ENV_SKY;
  });
}

hpt_status hpt_set_env_intensity(hpt_renderer* r, float v) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    r->env_intensity = v;
  });
}

hpt_status hpt_set_exposure_value(hpt_renderer* r, float ev) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    r->info.exposure_value = ev;
  });
}

hpt_status hpt_set_blue_noise(hpt_renderer* r, const uint8_t* rgba8, uint32_t w, uint32_t h) {
  return guarded([&] {
    need(ok_renderer(r) && rgba8, "null argument");
    need(w == 1024 && h == 1024, "the blue-noise texture must be 1024x1024 (rand4blue wraps at 1024)");
    need(!(r->committed && r->sd.use_count() > 1),
         "the blue-noise texture belongs to a scene shared with other handles; set it before sharing",
         HPT_ERROR_STATE);
    r->blue_noise.resize(1024 * 1024);
    std::memcpy(r->blue_noise.data(), rgba8, 1024 * 1024 * 4);
    if (r->committed) {
      set_device(r);
      upload(r->sd->d_blue_noise, r->blue_noise, r->stream);
    }
  });
}

hpt_status hpt_set_textures(hpt_renderer* r, const hpt_texture_desc* textures, uint32_t count) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(count == 0 || textures, "null textures");
    std::vector<uchar4> texels;
    std::vector<TexInfo> info;
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; ++i) {
      const hpt_texture_desc& t = textures[i];
      need(t.width > 0 && t.height > 0 && t.rgba8, "texture " + std::to_string(i) + " is empty");
      total += (uint64_t)t.width * t.height;
      need(total < (1ull << 32), "textures too large");
    }
    texels.resize(total);
    uint64_t off = 0;
    for (uint32_t i = 0; i < count; ++i) {
      const hpt_texture_desc& t = textures[i];
      const size_t n = (size_t)t.width * t.height;
      std::memcpy(texels.data() + off, t.rgba8, n * 4);
      info.push_back(TexInfo{(uint32_t)off, t.width, t.height, 0});
      off += n;
    }
    r->tex_texels.swap(texels);
    r->tex_info.swap(info);
    r->committed = false;
  });
}

hpt_status hpt_commit(hpt_renderer* r) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(r->scene_set, "hpt_set_scene has not been called", HPT_ERROR_STATE);
    set_device(r);
    hipStream_t s = r->stream;
    const uint32_t ntri = (uint32_t)(r->tri_pos.size() / 9);
    auto sd = std::make_shared<SceneDevice>();
    const auto t_build = std::chrono::steady_clock::now();
    bool on_gpu = false;
    if (r->bvh_builder == HPT_BVH_GPU && ntri > 0) {
      // The same tree as the host build, made on the device; a large node whose
      // centroids all coincide falls back to the host.
      DevBuf d_pos;
      upload(d_pos, r->tri_pos, s);
      GpuBvh g;
      std::string err;
      if (build_bvh8_gpu(d_pos.as<float>(), ntri, s, g, err)) {
        sd->d_nodes.adopt(g.nodes, (size_t)g.node_count * sizeof(Node8));
        sd->d_ltris.adopt(g.tris, (size_t)ntri * sizeof(LeafTri));
        sd->node_count = g.node_count;
        sd->leaf_tri_count = ntri;
        sd->bvh_depth = g.depth;
        on_gpu = true;
      } else {
        need(err.find("not finite") == std::string::npos, err);
      }
    }
    if (!on_gpu) {
      BvhBuildResult bvh;
      std::string err;
      need(build_bvh8(r->tri_pos.data(), ntri, bvh, err), err);
      sd->node_count = (uint32_t)bvh.nodes.size();
      sd->leaf_tri_count = (uint32_t)bvh.tris.size();
      sd->bvh_depth = bvh.depth;
      upload(sd->d_nodes, bvh.nodes, s);
      upload(sd->d_ltris, bvh.tris, s);
      check(hipStreamSynchronize(s), "upload");  // bvh is a local
    }
    r->bvh_builder_used = on_gpu ? HPT_BVH_GPU : HPT_BVH_HOST;
    r->bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_build).count();
    r->sd = sd;  // handles sharing the previous scene keep it
    upload(r->sd->d_tri_shade, r->tri_shade, s);
    std::vector<uint8_t> tri_key(r->tri_prim.size());
    for (size_t i = 0; i < tri_key.size(); ++i) {
      const uint32_t m = r->prims[r->tri_prim[i]].material_index;
      tri_key[i] = (uint8_t)std::min<uint32_t>(m, kSortKeys - 3);  // HPT_INVALID_INDEX too
    }
    if (tri_key.empty()) tri_key.push_back(0);
    upload(r->sd->d_tri_key, tri_key, s);
    upload(r->sd->d_prims, r->prims, s);
    upload(r->sd->d_materials, r->materials, s);
    upload(r->sd->d_lights, r->lights, s);
    upload(r->sd->d_cameras, r->cameras, s);
    upload_env(r, s);
    const uint32_t ntex = (uint32_t)r->tex_info.size();
    for (const hpt_material& m : r->materials)
      for (uint32_t t : {m.base_color_map_index, m.normal_map_index, m.metallic_roughness_map_index,
                         m.emission_map_index})
        need(t == HPT_INVALID_INDEX || t < ntex, "material texture index out of range of hpt_set_textures");
    // Uploads are asynchronous: every host array stays alive until the
    // stream synchronisation at the end of this call.
    const std::vector<uchar4> no_texels(1);
    const std::vector<TexInfo> no_info(1);
    upload(r->sd->d_tex_texels, r->tex_texels.empty() ? no_texels : r->tex_texels, s);
    upload(r->sd->d_tex_info, r->tex_info.empty() ? no_info : r->tex_info, s);
    std::vector<float> lut(256);
    for (int i = 0; i < 256; ++i) lut[i] = srgb_to_linear(i);
    upload(r->sd->d_srgb_lut, lut, s);
    // The procedural stand-in is made here, when first needed: handles that
    // render a shared scene never upload one.
    if (r->blue_noise.empty()) default_blue_noise(r->blue_noise);
    upload(r->sd->d_blue_noise, r->blue_noise, s);
    alloc_frame_state(r);
    check(hipStreamSynchronize(s), "commit");
    r->frame_index = 0;
    r->committed = true;
  });
}

hpt_status hpt_share_scene(hpt_renderer* dst, const hpt_renderer* src) {
  return guarded([&] {
    need(ok_renderer(dst) && ok_renderer(src), "null or invalid renderer handle");
    need(src->committed, "the source renderer has no committed scene", HPT_ERROR_STATE);
    need(dst != src, "a renderer cannot share its own scene");
    need(dst->info.device == src->info.device, "renderers on different devices cannot share a scene");
    set_device(dst);
    // Host-side copies the frame parameters and the ray-query decoding read.
    try {
      dst->cameras = src->cameras;
      dst->lights = src->lights;
      dst->materials = src->materials;
      dst->tri_prim = src->tri_prim;
      dst->prims = src->prims;
    } catch (const std::bad_alloc&) {
      throw Fail(HPT_ERROR_OUT_OF_MEMORY,
                 "share scene: host copy of the scene arrays failed (cameras " + std::to_string(src->cameras.size()) +
                     ", lights " + std::to_string(src->lights.size()) + ", materials " +
                     std::to_string(src->materials.size()) + ", triangles " + std::to_string(src->tri_prim.size()) +
                     ", primitives " + std::to_string(src->prims.size()) + ")");
    }
    dst->camera_count = src->camera_count;
    dst->num_lights = src->num_lights;
    dst->env_type = src->env_type;
    std::memcpy(dst->ground, src->ground, sizeof dst->ground);
    std::memcpy(dst->sky, src->sky, sizeof dst->sky);
    dst->env_intensity = src->env_intensity;
    dst->env_rotation = src->env_rotation;
    dst->env_total = src->env_total;
    dst->env_w = src->env_w;
    dst->env_h = src->env_h;
    dst->tri_pos.clear();
    dst->tri_shade.clear();
    dst->sd = src->sd;
    dst->scene_set = false;  // a later hpt_commit needs hpt_set_scene first
    alloc_frame_state(dst);
    check(hipStreamSynchronize(dst->stream), "share scene");
    dst->frame_index = 0;
    dst->committed = true;
  });
}

hpt_status hpt_query_idle(hpt_renderer* r, int* idle) {
  return guarded([&] {
    need(ok_renderer(r) && idle, "null argument");
    set_device(r);
    const hipError_t e = hipStreamQuery(r->stream);
    if (e == hipErrorNotReady) {
      *idle = 0;
      return;
    }
    check(e, "hipStreamQuery");
    *idle = 1;
  });
}

hpt_status hpt_get_memory_usage(hpt_renderer* r, uint64_t* frame_bytes, uint64_t* scene_bytes) {
  return guarded([&] {
    need(ok_renderer(r) && frame_bytes && scene_bytes, "null argument");
    uint64_t f = 0, sc = 0;
    for (const DevBuf* b : {&r->d_rng, &r->d_org, &r->d_dir, &r->d_thr, &r->d_rad, &r->d_alb, &r->d_nrm, &r->d_med,
                            &r->d_med2, &r->d_nee_e, &r->d_nee_l, &r->d_hit, &r->d_qa, &r->d_qb, &r->d_sh_o,
                            &r->d_sh_d, &r->d_sh_c, &r->d_counts, &r->d_spill, &r->d_stats, &r->d_sort_keys,
                            &r->d_sort_items, &r->d_accum,
                            &r->d_albedo, &r->d_normal, &r->d_final})
      f += b->bytes;
    if (r->sd)
      for (const DevBuf* b : {&r->sd->d_nodes, &r->sd->d_ltris, &r->sd->d_tri_shade, &r->sd->d_tri_key,
                              &r->sd->d_prims, &r->sd->d_materials, &r->sd->d_lights, &r->sd->d_cameras, &r->sd->d_env_rgb,
                              &r->sd->d_env_table, &r->sd->d_env_texels, &r->sd->d_tex_texels, &r->sd->d_tex_info,
                              &r->sd->d_srgb_lut, &r->sd->d_blue_noise})
        sc += b->bytes;
    *frame_bytes = f;
    *scene_bytes = sc;
  });
}

hpt_status hpt_render(hpt_renderer* r, uint32_t frames, uint32_t camera_index) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    require_committed(r);
    need(camera_index < r->camera_count, "camera index out of range");
    set_device(r);
    const uint32_t cap = batch_frames_for(r);
    if (cap != r->batch_cap) alloc_paths(r, cap);
    // Queue counters start from zero on every call, whatever an earlier
    // failed launch left behind.
    check(hipMemsetAsync(r->d_counts.p, 0, r->d_counts.bytes, r->stream), "hipMemsetAsync");
    uint64_t left = frames;
    if (r->info.max_samples)
      left = r->frame_index >= r->info.max_samples ? 0 : std::min<uint64_t>(left, r->info.max_samples - r->frame_index);
    left = std::min<uint64_t>(left, 0xFFFFFFFFull - r->frame_index);  // frame indices stay 32-bit
    const bool graph = r->launch_mode == HPT_LAUNCH_GRAPH && !r->profiling;
    while (left) {
      const uint32_t k = (uint32_t)std::min<uint64_t>(left, cap);
      if (graph && k == cap) {
        render_batch_graph(r, camera_index, k);
      } else {
        render_batch(r, camera_index, k);
        r->frame_synced = false;
      }
      r->frame_index += k;
      left -= k;
    }
    if (r->profiling & 1) collect_timers(r);
  });
}

hpt_status hpt_get_truncated_paths(hpt_renderer* r, uint64_t* paths) {
  return guarded([&] {
    need(ok_renderer(r) && paths, "null argument");
    *paths = 0;
    if (!r->d_tail_cut.p) return;
    set_device(r);
    unsigned long long v = 0;
    check(hipMemcpyAsync(&v, r->d_tail_cut.p, sizeof(v), hipMemcpyDeviceToHost, r->stream), "hipMemcpyAsync");
    check(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
    *paths = v;
  });
}

hpt_status hpt_set_bvh_builder(hpt_renderer* r, uint32_t builder) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(builder == HPT_BVH_HOST || builder == HPT_BVH_GPU, "unknown BVH builder");
    r->bvh_builder = builder;
  });
}

hpt_status hpt_get_bvh_build_stats(hpt_renderer* r, uint32_t* builder_used, double* build_ms) {
  return guarded([&] {
    need(ok_renderer(r) && builder_used && build_ms, "null argument");
    require_committed(r);
    *builder_used = r->bvh_builder_used;
    *build_ms = r->bvh_build_ms;
  });
}

hpt_status hpt_set_shade_sort(hpt_renderer* r, uint32_t first_bounce) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    r->shade_sort_from = first_bounce;
    if (r->committed && r->batch_cap) {
      set_device(r);
      check(hipStreamSynchronize(r->stream), "shade sort");  // buffers may be in use
      alloc_sort(r, r->batch_cap);
    }
  });
}

hpt_status hpt_set_tail_from(hpt_renderer* r, uint32_t first_bounce) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(first_bounce >= 1, "the first bounce always runs as a wavefront");
    r->tail_from = first_bounce;
  });
}

hpt_status hpt_group_set_tail_from(hpt_group* g, uint32_t first_bounce) {
  return guarded([&] {
    need(ok_group(g), "null or invalid group handle");
    need(first_bounce >= 1, "the first bounce always runs as a wavefront");
    g->tail_from = first_bounce;
  });
}

static bool traversal_block_ok(uint32_t threads) {
  for (uint32_t b : kTraversalBlocks)
    if (b == threads) return true;
  return false;
}

hpt_status hpt_set_traversal_block(hpt_renderer* r, uint32_t threads) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(traversal_block_ok(threads), "the traversal kernels are built for 64, 128 or 256 threads per block");
    r->trav_block = threads;
  });
}

hpt_status hpt_group_set_traversal_block(hpt_group* g, uint32_t threads) {
  return guarded([&] {
    need(ok_group(g), "null or invalid group handle");
    need(traversal_block_ok(threads), "the traversal kernels are built for 64, 128 or 256 threads per block");
    g->trav_block = threads;
  });
}

hpt_status hpt_set_traversal_refill(hpt_renderer* r, uint32_t on) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    r->refill = on != 0;
  });
}

hpt_status hpt_group_set_traversal_refill(hpt_group* g, uint32_t on) {
  return guarded([&] {
    need(ok_group(g), "null or invalid group handle");
    g->refill = on != 0;
  });
}

hpt_status hpt_set_vmis_phases(hpt_renderer* r, uint32_t on) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(on <= 1, "on must be 0 or 1");
    r->vmis_phases = on;
  });
}

hpt_status hpt_set_launch_mode(hpt_renderer* r, uint32_t mode) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(mode == HPT_LAUNCH_EAGER || mode == HPT_LAUNCH_GRAPH, "unknown launch mode");
    r->launch_mode = mode;
  });
}

hpt_status hpt_get_launch_stats(hpt_renderer* r, uint64_t* graph_launches, uint64_t* graph_captures) {
  return guarded([&] {
    need(ok_renderer(r) && graph_launches && graph_captures, "null argument");
    *graph_launches = r->graph_launches;
    *graph_captures = r->graph_captures;
  });
}

hpt_status hpt_reset_accumulation(hpt_renderer* r) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    r->frame_index = 0;
    r->frame_synced = false;
    if (r->committed) {
      set_device(r);
      check(hipMemsetAsync(r->d_accum.p, 0, r->d_accum.bytes, r->stream), "hipMemsetAsync");
      check(hipMemsetAsync(r->d_albedo.p, 0, r->d_albedo.bytes, r->stream), "hipMemsetAsync");
      check(hipMemsetAsync(r->d_normal.p, 0, r->d_normal.bytes, r->stream), "hipMemsetAsync");
      check(hipMemsetAsync(r->d_final.p, 0, r->d_final.bytes, r->stream), "hipMemsetAsync");
    }
  });
}

hpt_status hpt_set_frame_index(hpt_renderer* r, uint64_t frame_index) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(frame_index <= 0xFFFFFFFFull, "frame index must fit the 32-bit UBO field");
    r->frame_index = frame_index;
    r->frame_synced = false;
  });
}

hpt_status hpt_get_frame_index(hpt_renderer* r, uint64_t* fi) {
  return guarded([&] {
    need(ok_renderer(r) && fi, "null argument");
    *fi = r->frame_index;
  });
}

hpt_status hpt_read_image(hpt_renderer* r, hpt_image which, void* dst, size_t bytes) {
  return guarded([&] {
    need(ok_renderer(r) && dst, "null argument");
    require_committed(r);
    set_device(r);
    const DevBuf* b = nullptr;
    switch (which) {
      case HPT_IMAGE_FINAL_RGBA8: b = &r->d_final; break;
      case HPT_IMAGE_ACCUM_RGBA32F: b = &r->d_accum; break;
      case HPT_IMAGE_ALBEDO_RGBA32F: b = &r->d_albedo; break;
      case HPT_IMAGE_NORMAL_RGBA32F: b = &r->d_normal; break;
      default: need(false, "unknown image");
    }
    need(bytes == b->bytes, "destination size does not match the image size");
    check(hipMemcpyAsync(dst, b->p, bytes, hipMemcpyDeviceToHost, r->stream), "read image");
    check(hipStreamSynchronize(r->stream), "read image");
  });
}

hpt_status hpt_get_image_pointer(hpt_renderer* r, hpt_image which, void** ptr, size_t* bytes) {
  return guarded([&] {
    need(ok_renderer(r) && ptr && bytes, "null argument");
    require_committed(r);
    const DevBuf* b = nullptr;
    switch (which) {
      case HPT_IMAGE_FINAL_RGBA8: b = &r->d_final; break;
      case HPT_IMAGE_ACCUM_RGBA32F: b = &r->d_accum; break;
      case HPT_IMAGE_ALBEDO_RGBA32F: b = &r->d_albedo; break;
      case HPT_IMAGE_NORMAL_RGBA32F: b = &r->d_normal; break;
      default: need(false, "unknown image");
    }
    *ptr = b->p;
    *bytes = b->bytes;
  });
}

static void trace_api(hpt_renderer* r, const hpt_ray* rays, uint32_t n, int any, std::vector<float4>& out,
                      uint32_t* nvis = nullptr, uint32_t* ntri = nullptr) {
  set_device(r);
  require_committed(r);
  out.resize(n);
  if (n == 0) return;
  DevBuf dr, dout, dcnt;
  dr.alloc((size_t)n * sizeof(hpt_ray));
  dout.alloc((size_t)n * sizeof(float4));
  uint32_t* dv = nullptr;
  uint32_t* dt = nullptr;
  if (nvis && ntri) {
    dcnt.alloc((size_t)n * 8);
    dv = dcnt.as<uint32_t>();
    dt = dv + n;
  }
  check(hipMemcpyAsync(dr.p, rays, (size_t)n * sizeof(hpt_ray), hipMemcpyHostToDevice, r->stream), "rays");
  const Spill sp = ensure_spill(r, round_threads(n));
  launch_trace_rays(r->stream, scene_view(r), dr.as<hpt_ray>(), n, any, dout.as<float4>(), dv, dt, sp);
  check(hipGetLastError(), "trace launch");
  check(hipMemcpyAsync(out.data(), dout.p, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, r->stream), "hits");
  if (dv) {
    check(hipMemcpyAsync(nvis, dv, (size_t)n * 4, hipMemcpyDeviceToHost, r->stream), "counts");
    check(hipMemcpyAsync(ntri, dt, (size_t)n * 4, hipMemcpyDeviceToHost, r->stream), "counts");
  }
  check(hipStreamSynchronize(r->stream), "trace");
}

hpt_status hpt_trace_counts(hpt_renderer* r, const hpt_ray* rays, uint32_t n, int any_hit, uint32_t* node_visits,
                            uint32_t* tri_tests) {
  return guarded([&] {
    need(ok_renderer(r) && (n == 0 || (rays && node_visits && tri_tests)), "null argument");
    std::vector<float4> out;
    trace_api(r, rays, n, any_hit ? 1 : 0, out, node_visits, tri_tests);
  });
}

hpt_status hpt_trace_closest(hpt_renderer* r, const hpt_ray* rays, uint32_t n, hpt_hit* hits) {
  return guarded([&] {
    need(ok_renderer(r) && (n == 0 || (rays && hits)), "null argument");
    std::vector<float4> out;
    trace_api(r, rays, n, 0, out);
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t id;
      std::memcpy(&id, &out[i].w, 4);
      hpt_hit& h = hits[i];
      if (id == MISS_ID) {
        h = hpt_hit{0.0f, 0.0f, 0.0f, HPT_HIT_NONE, HPT_INVALID_INDEX, HPT_INVALID_INDEX};
      } else if (id & LIGHT_BIT) {
        h = hpt_hit{out[i].x, 0.0f, 0.0f, HPT_HIT_LIGHT, id & ~LIGHT_BIT, HPT_INVALID_INDEX};
      } else {
        const uint32_t prim = r->tri_prim[id];
        h = hpt_hit{out[i].x, out[i].y, out[i].z, HPT_HIT_TRIANGLE, prim, id - r->prims[prim].local_base};
      }
    }
  });
}

hpt_status hpt_trace_any(hpt_renderer* r, const hpt_ray* rays, uint32_t n, uint32_t* occluded) {
  return guarded([&] {
    need(ok_renderer(r) && (n == 0 || (rays && occluded)), "null argument");
    std::vector<float4> out;
    trace_api(r, rays, n, 1, out);
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t v;
      std::memcpy(&v, &out[i].w, 4);
      occluded[i] = v;
    }
  });
}

// ------------------------------------------------ allocation-free ray queries
// hpt_trace_closest_device keeps the traversal stack entries that do not fit
// a lane's LDS short stack (trees deeper than kShortStack levels) in this
// array instead of a buffer of the handle.  Each GPU holds its own copy until
// the process ends; host code uses the address hipGetSymbolAddress reports
// for the current device, never the array's name.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchEntries = kScratchBytes / sizeof(uint2);
static __device__ __align__(256) uint2 scratch[kScratchEntries];

namespace {

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mutex;
static uint2* g_scratch[kScratchDevices] = {};

// The scratch array's address on the current device `device`, looked up
// once per device.
static uint2* scratch_on(int device) {
  need(device >= 0 && device < kScratchDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch[device] = static_cast<uint2*>(ptr);
  }
  return g_scratch[device];
}

}  // namespace

hpt_status hpt_trace_closest_device(hpt_renderer* r, const hpt_ray* rays, uint32_t n, hpt_hit* hits) {
  return guarded([&] {
    need(ok_renderer(r) && (n == 0 || (rays && hits)), "null argument");
    set_device(r);
    require_committed(r);
    if (n == 0) return;
    Spill sp{nullptr, 0};
    if (r->sd->bvh_depth > (uint32_t)kShortStack) {
      const size_t threads = round_threads(n);
      const size_t entries = (size_t)(r->sd->bvh_depth - kShortStack) * threads;
      need(entries <= kScratchEntries,
           "hpt_trace_closest_device: " + std::to_string(n) + " rays over a " + std::to_string(r->sd->bvh_depth) +
               "-level tree need " + std::to_string(entries) + " stack entries beyond the LDS stack, above the " +
               std::to_string(kScratchEntries) + " the device scratch array holds; split the call");
      sp.buf = scratch_on(r->info.device);
      sp.stride = (uint32_t)threads;
    }
    launch_trace_hits(r->stream, scene_view(r), rays, n, hits, sp);
    check(hipGetLastError(), "trace launch");
  });
}

hpt_status hpt_wait_idle(hpt_renderer* r) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    set_device(r);
    check(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
  });
}

void hpt_destroy(hpt_renderer* r) {
  if (!ok_renderer(r)) return;  // null, another kind of handle, or destroyed
  r->tag = 0;
  (void)hipSetDevice(r->info.device);
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  if (r->gexec) (void)hipGraphExecDestroy(r->gexec);
  if (r->ev) (void)hipEventDestroy(r->ev);
  for (auto& t : r->timers) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

// ---------------------------------------------------------------- presentation
// A ring of pinned host images that the renderer's final rgba8 image is
// copied into behind its frames: the window loop takes the newest finished
// one without waiting for the tracer (the role of the swapchain hand-off in
// the reference's update()/render() loop, hala-pathtracer.rs:236-289).
struct hpt_presenter {
  uint32_t tag = kPresenterTag;
  hpt_renderer* r = nullptr;
  int32_t device = 0;
  hipStream_t copy_stream = nullptr;
  size_t bytes = 0;
  DevBuf staging;  // one device copy of the final image per slot
  std::vector<uint8_t*> host;
  std::vector<hipEvent_t> staged, ready;
  std::vector<uint64_t> frame;
  // Slot states: free, copy in flight, ready, held by the caller.
  enum { kFree = 0, kInFlight = 1, kReady = 2, kHeld = 3 };

  std::vector<int> state;
  uint32_t next = 0;
  uint64_t presented = 0, dropped = 0;
};

static bool ok_presenter(const hpt_presenter* p) { return p && p->tag == kPresenterTag; }

namespace {

static void present_poll(hpt_presenter* p) {
  for (size_t i = 0; i < p->state.size(); ++i) {
    if (p->state[i] != hpt_presenter::kInFlight) continue;
    const hipError_t e = hipEventQuery(p->ready[i]);
    if (e == hipSuccess) p->state[i] = hpt_presenter::kReady;
    else if (e != hipErrorNotReady) check(e, "presentation copy");
  }
}

static void present_free(hpt_presenter* p) {
  (void)hipSetDevice(p->device);
  if (p->copy_stream) (void)hipStreamSynchronize(p->copy_stream);
  for (uint8_t* h : p->host)
    if (h) (void)hipHostFree(h);
  for (hipEvent_t e : p->staged)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->ready)
    if (e) (void)hipEventDestroy(e);
  if (p->copy_stream) (void)hipStreamDestroy(p->copy_stream);
  delete p;
}

}  // namespace

hpt_status hpt_present_create(hpt_renderer* r, uint32_t slots, hpt_presenter** out) {
  hpt_presenter* p = nullptr;
  const hpt_status st = guarded([&] {
    need(ok_renderer(r) && out, "null argument");
    *out = nullptr;
    need(slots >= 2 && slots <= 16, "slots must be 2..16");
    require_committed(r);
    set_device(r);
    p = new hpt_presenter;
    p->r = r;
    p->device = r->info.device;
    p->bytes = r->d_final.bytes;
    check(hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
    p->staging.alloc(p->bytes * slots);
    p->host.assign(slots, nullptr);
    p->staged.assign(slots, nullptr);
    p->ready.assign(slots, nullptr);
    p->frame.assign(slots, 0);
    p->state.assign(slots, hpt_presenter::kFree);
    for (uint32_t i = 0; i < slots; ++i) {
      check(hipHostMalloc(reinterpret_cast<void**>(&p->host[i]), p->bytes, hipHostMallocDefault), "hipHostMalloc");
      check(hipEventCreateWithFlags(&p->staged[i], hipEventDisableTiming), "hipEventCreate");
      check(hipEventCreateWithFlags(&p->ready[i], hipEventDisableTiming), "hipEventCreate");
    }
    *out = p;
  });
  if (st != HPT_OK && p) present_free(p);
  return st;
}

hpt_status hpt_present_frame(hpt_presenter* p) {
  return guarded([&] {
    need(ok_presenter(p), "null or invalid presenter handle");
    hpt_renderer* r = p->r;
    set_device(r);
    need(r->d_final.bytes == p->bytes, "the renderer's image size changed");
    present_poll(p);
    // Next slot in ring order that is not in flight or held; a ready frame
    // that was never taken is overwritten (newer frames replace it).
    const uint32_t n = (uint32_t)p->state.size();
    int slot = -1;
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t i = (p->next + k) % n;
      if (p->state[i] == hpt_presenter::kFree || p->state[i] == hpt_presenter::kReady) {
        slot = (int)i;
        break;
      }
    }
    if (slot < 0) {  // every slot busy: drop this frame rather than wait
      p->dropped++;
      return;
    }
    uint8_t* dev = p->staging.as<uint8_t>() + (size_t)slot * p->bytes;
    // The device copy is ordered on the renderer's stream, before any later
    // frame rewrites the final image; the host copy then runs on a stream of
    // its own and the tracer never waits for it.
    check(hipMemcpyAsync(dev, r->d_final.p, p->bytes, hipMemcpyDeviceToDevice, r->stream), "present stage");
    check(hipEventRecord(p->staged[slot], r->stream), "hipEventRecord");
    check(hipStreamWaitEvent(p->copy_stream, p->staged[slot], 0), "hipStreamWaitEvent");
    check(hipMemcpyAsync(p->host[slot], dev, p->bytes, hipMemcpyDeviceToHost, p->copy_stream), "present copy");
    check(hipEventRecord(p->ready[slot], p->copy_stream), "hipEventRecord");
    p->frame[slot] = r->frame_index;
    p->state[slot] = hpt_presenter::kInFlight;
    p->next = (uint32_t)(slot + 1) % n;
    p->presented++;
  });
}

hpt_status hpt_present_acquire(hpt_presenter* p, const void** pixels, uint64_t* frame_index) {
  return guarded([&] {
    need(ok_presenter(p) && pixels && frame_index, "null argument");
    *pixels = nullptr;
    *frame_index = 0;
    set_device(p->r);
    present_poll(p);
    int best = -1;
    for (size_t i = 0; i < p->state.size(); ++i) {
      const int s = p->state[i];
      if ((s == hpt_presenter::kReady || s == hpt_presenter::kHeld) && (best < 0 || p->frame[i] > p->frame[best]))
        best = (int)i;
    }
    if (best < 0) return;  // nothing finished yet
    for (size_t i = 0; i < p->state.size(); ++i) {
      const int s = p->state[i];
      // frames older than the one handed out are stale: their slots free up
      if ((int)i != best && (s == hpt_presenter::kHeld || (s == hpt_presenter::kReady && p->frame[i] < p->frame[best])))
        p->state[i] = hpt_presenter::kFree;
    }
    p->state[best] = hpt_presenter::kHeld;
    *pixels = p->host[best];
    *frame_index = p->frame[best];
  });
}

hpt_status hpt_present_get_stats(hpt_presenter* p, uint64_t* presented, uint64_t* dropped) {
  return guarded([&] {
    need(ok_presenter(p) && presented && dropped, "null argument");
    *presented = p->presented;
    *dropped = p->dropped;
  });
}

void hpt_present_destroy(hpt_presenter* p) {
  if (!ok_presenter(p)) return;
  p->tag = 0;
  present_free(p);
}

const char* hpt_last_error(void) { return g_last_error.c_str(); }

hpt_status hpt_get_image_size(hpt_renderer* r, uint32_t* width, uint32_t* height) {
  return guarded([&] {
    need(ok_renderer(r) && width && height, "null argument");
    *width = r->info.width;
    *height = r->info.height;
  });
}

hpt_status hpt_get_bvh_info(hpt_renderer* r, hpt_bvh_info* info) {
  return guarded([&] {
    need(ok_renderer(r) && info, "null argument");
    require_committed(r);
    info->node_count = r->sd->node_count;
    info->triangle_count = r->sd->leaf_tri_count;
    info->max_depth = r->sd->bvh_depth;
    info->_pad = 0;
  });
}

hpt_status hpt_export_bvh(hpt_renderer* r, void* nodes, size_t node_bytes, void* tris, size_t tri_bytes) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    require_committed(r);
    need(node_bytes == (size_t)r->sd->node_count * sizeof(Node8), "node buffer size mismatch");
    need(tri_bytes == (size_t)r->sd->leaf_tri_count * sizeof(LeafTri), "triangle buffer size mismatch");
    set_device(r);
    if (node_bytes) check(hipMemcpy(nodes, r->sd->d_nodes.p, node_bytes, hipMemcpyDeviceToHost), "export nodes");
    if (tri_bytes) check(hipMemcpy(tris, r->sd->d_ltris.p, tri_bytes, hipMemcpyDeviceToHost), "export tris");
  });
}

hpt_status hpt_set_frames_in_flight(hpt_renderer* r, uint32_t frames) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need(frames <= kMaxBatchFrames, "at most 64 frames in flight");
    const uint64_t px = (uint64_t)r->info.width * r->info.height;
    need((uint64_t)frames * px <= (1ull << 31), "frames * width * height must not exceed 2^31");
    r->frames_in_flight = frames;
  });
}

hpt_status hpt_get_frames_in_flight(hpt_renderer* r, uint32_t* frames) {
  return guarded([&] {
    need(ok_renderer(r) && frames, "null argument");
    *frames = r->npix ? batch_frames_for(r) : r->frames_in_flight;
  });
}

hpt_status hpt_set_profiling(hpt_renderer* r, int enable) {
  return guarded([&] {
    need(ok_renderer(r), "null or invalid renderer handle");
    need((enable & ~3) == 0, "unknown profiling bits");
    set_device(r);
    if (r->timers_used) collect_timers(r);
    r->profiling = enable;
    for (auto& s : r->stats) s = hpt_stage_stats{};
    r->timers_used = 0;
    if (enable) {
      r->d_stats.alloc((size_t)kStatsPerStage * HPT_STAGE_COUNT * sizeof(unsigned long long));
      check(hipMemsetAsync(r->d_stats.p, 0, r->d_stats.bytes, r->stream), "hipMemsetAsync");
      check(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
    }
  });
}

hpt_status hpt_get_stage_stats(hpt_renderer* r, int stage, hpt_stage_stats* out) {
  return guarded([&] {
    need(ok_renderer(r) && out, "null argument");
    need(stage >= 0 && stage < HPT_STAGE_COUNT, "unknown stage");
    set_device(r);
    if (r->timers_used) collect_timers(r);
    *out = r->stats[stage];
    if (r->profiling && r->d_stats.p) {
      std::vector<unsigned long long> v(kStatsPerStage);
      check(hipStreamSynchronize(r->stream), "hipStreamSynchronize");
      check(hipMemcpy(v.data(), r->d_stats.as<unsigned long long>() + (size_t)kStatsPerStage * stage,
                      v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost),
            "stats");
      uint64_t sum[kStatCounters] = {};
      for (uint32_t j = 0; j < kStatCounters; ++j)
        for (uint32_t sg = 0; sg < kSeg; ++sg) sum[j] += v[(j * kSeg + sg) * kStatStride];
      out->rays = sum[0];
      out->node_visits = sum[1];
      out->tri_tests = sum[2];
      out->queries = sum[3];
    }
  });
}

hpt_status hpt_group_create(int32_t device, hpt_group** out) {
  return guarded([&] {
    need(out != nullptr, "null argument");
    *out = nullptr;
    int n = 0;
    check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    need(device >= 0 && device < n, "device ordinal out of range");
    std::unique_ptr<hpt_group> g(new hpt_group());
    g->device = device;
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
    ensure_event(g->done);
    ensure_event(g->step_end[0]);
    ensure_event(g->step_end[1]);
    *out = g.release();
  });
}

hpt_status hpt_group_set_launch_mode(hpt_group* g, uint32_t mode) {
  return guarded([&] {
    need(ok_group(g), "null or invalid group handle");
    need(mode == HPT_LAUNCH_EAGER || mode == HPT_LAUNCH_GRAPH, "unknown launch mode");
    g->launch_mode = mode;
  });
}

hpt_status hpt_group_set_vmis_phases(hpt_group* g, uint32_t on) {
  return guarded([&] {
    need(ok_group(g), "null or invalid group handle");
    need(on <= 1, "on must be 0 or 1");
    g->vmis_phases = on;
  });
}

hpt_status hpt_group_set_shade_sort(hpt_group* g, uint32_t first_bounce) {
  return guarded([&] {
    need(ok_group(g), "null or invalid group handle");
    g->shade_sort_from = first_bounce;
  });
}

hpt_status hpt_group_render(hpt_group* g, hpt_renderer* const* jobs, const uint32_t* frames,
                            const uint32_t* camera_indices, uint32_t count) {
  return guarded([&] {
    need(ok_group(g) && jobs && frames, "null argument");
    need(count >= 1 && count <= kMaxWaveJobs, "a group step takes 1 to 256 jobs");
    hpt_renderer* r0 = jobs[0];
    need(ok_renderer(r0), "null or invalid job handle");
    for (uint32_t i = 0; i < count; ++i) {
      hpt_renderer* r = jobs[i];
      need(ok_renderer(r), "null or invalid job handle");
      require_committed(r);
      need(r->info.device == g->device, "job on another device than the group");
      need(r->sd == r0->sd, "the jobs of a group step must render one committed scene (hpt_share_scene)");
      need(r->info.features == r0->info.features, "the jobs of a group step must use one shader permutation");
      need(same_environment(r, r0), "the jobs of a group step must share environment and light settings");
      need(!camera_indices || camera_indices[i] < r->camera_count, "camera index out of range");
      for (uint32_t j = 0; j < i; ++j) need(jobs[j] != r, "a job appears twice in a group step");
    }
    check(hipSetDevice(g->device), "hipSetDevice");
    // Frames per job, capped like hpt_render (max_samples, 32-bit frame indices).
    std::vector<uint32_t> k(count);
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; ++i) {
      const hpt_renderer* r = jobs[i];
      uint64_t left = frames[i];
      if (r->info.max_samples)
        left = r->frame_index >= r->info.max_samples ? 0 : std::min<uint64_t>(left, r->info.max_samples - r->frame_index);
      left = std::min<uint64_t>(left, 0xFFFFFFFFull - r->frame_index);
      k[i] = (uint32_t)left;
      total += left * r->npix;
    }
    need(2 * total < 0xFFFFFFFFull, "group step too large: queue indices are 32-bit");
    if (total == 0) return;
    const uint32_t n = (uint32_t)total;
    check(hipEventSynchronize(g->step_end[g->steps & 1]), "hipEventSynchronize");  // the step two before
    group_alloc(g, n, r0->sd->bvh_depth, vmis_features(r0->info.features));
    // The step runs after each job's earlier work on its own stream.
    for (uint32_t i = 0; i < count; ++i) {
      ensure_event(jobs[i]->ev);
      check(hipEventRecord(jobs[i]->ev, jobs[i]->stream), "hipEventRecord");
      check(hipStreamWaitEvent(g->stream, jobs[i]->ev, 0), "hipStreamWaitEvent");
    }
    check(hipMemsetAsync(g->d_counts.p, 0, g->d_counts.bytes, g->stream), "hipMemsetAsync");

    const bool graph = g->launch_mode == HPT_LAUNCH_GRAPH;
    std::vector<WaveJob> wj;
    std::vector<hpt_renderer*> live;
    JobLimits lim{};
    uint32_t base = 0, max_depth = 0;
    std::vector<uint8_t> key;
    for (uint32_t i = 0; i < count; ++i) {
      hpt_renderer* r = jobs[i];
      if (!k[i]) continue;
      if (graph && !r->d_frame.p) r->d_frame.alloc(256);
      FrameParams fp = frame_params(r, camera_indices ? camera_indices[i] : 0);
      fp.batch_frames = k[i];
      FrameParams kfp = fp;
      kfp.frame_index = 0;
      kfp.frame_ptr = r->d_frame.as<uint32_t>();
      key_append(key, kfp);
      key_append(key, images_of(r));
      lim.v[wj.size()] = make_uint2(r->info.max_depth, r->info.rr_depth);
      max_depth = std::max(max_depth, r->info.max_depth);
      wj.push_back(WaveJob{fp, images_of(r), base, k[i] * r->npix});
      live.push_back(r);
      base += k[i] * r->npix;
    }
    const SceneView sv = scene_view(r0);
    FrameParams shade_fp = frame_params(r0, 0);
    shade_fp.frame_index = 0;  // shading reads no frame index (the path RNG carries it)
    shade_fp.max_depth = max_depth;
    shade_fp.job_limits = g->d_limits.as<uint2>();
    WaveBuffers wb{};
    wb.ps = PathState{g->d_rng.as<uint4>(),   g->d_org.as<float4>(),   g->d_dir.as<float4>(),
                      g->d_thr.as<float4>(),  g->d_rad.as<float4>(),   g->d_alb.as<float4>(),
                      g->d_nrm.as<float4>(),  g->d_med.as<float4>(),   g->d_med2.as<float2>(),
                      g->d_nee_e.as<float4>(), g->d_nee_l.as<float4>(), g->d_hit.as<float4>(),
                      g->d_job.as<uint16_t>(), g->d_nee_dir.as<float4>(), g->d_nee_li.as<float4>(),
                      g->d_nee_ld.as<float4>(), g->d_nee_tr.as<float4>()};
    wb.qa = g->d_qa.as<uint32_t>();
    wb.qb = g->d_qb.as<uint32_t>();
    wb.counts = g->d_counts.as<uint32_t>();
    wb.sh_o = g->d_sh_o.as<float4>();
    wb.sh_d = g->d_sh_d.as<float4>();
    wb.sh_c = g->d_sh_c.as<float4>();
    wb.sp = Spill{g->d_spill.as<uint2>(), 2 * kSeg * seg_geom(n).seg_cap};
    if (!g->d_spill.p) wb.sp = Spill{nullptr, 0};
    wb.sort_keys = g->d_sort_keys.as<uint8_t>();
    wb.sort_items = g->d_sort_items.as<uint32_t>();
    wb.tail_cut = g->d_tail_cut.as<unsigned long long>();
    wb.qc = g->vmis_phases ? g->d_qc.as<uint32_t>() : nullptr;
    wb.refill = g->refill;
    key_append(key, sv);
    key_append(key, shade_fp);
    key_append(key, wb.ps);
    for (const void* q : {(const void*)wb.qa, (const void*)wb.qb, (const void*)wb.counts, (const void*)wb.sh_o,
                          (const void*)wb.sh_d, (const void*)wb.sh_c, (const void*)wb.sp.buf,
                          (const void*)wb.sort_keys, (const void*)wb.sort_items, (const void*)wb.qc,
                          (const void*)wb.tail_cut})
      key_append(key, q);
    key_append(key, wb.sp.stride);
    key_append(key, n);
    key_append(key, r0->info.features);
    key_append(key, g->shade_sort_from);
    key_append(key, g->tail_from);
    key_append(key, g->trav_block);
    key_append(key, g->refill);
    key_append(key, g->vmis_phases);
    for (const WaveJob& j : wj) key_append(key, j.base);
    const uint32_t njobs = (uint32_t)wj.size();
    const uint32_t fixed_bounces = max_depth + 1;
    const bool replay = graph && key == g->last_key;
    g->last_key = key;
    if (replay) {
      for (hpt_renderer* r : live)
        if (!r->frame_synced) {
          launch_frame_counter(g->stream, r->d_frame.as<uint32_t>(), (uint32_t)r->frame_index, false);
          r->frame_synced = true;
        }
      check(hipGetLastError(), "frame counter launch");
      if (!g->gexec || key != g->gkey) {
        if (g->gexec) {
          check(hipStreamSynchronize(g->stream), "graph");
          check(hipGraphExecDestroy(g->gexec), "hipGraphExecDestroy");
          g->gexec = nullptr;
        }
        std::vector<WaveJob> cj = wj;
        for (size_t i = 0; i < cj.size(); ++i) {
          cj[i].fp.frame_index = 0;
          cj[i].fp.frame_ptr = live[i]->d_frame.as<uint32_t>();
        }
        hipGraph_t gr = nullptr;
        check(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
          launch_job_limits(g->stream, g->d_limits.as<uint2>(), lim, njobs);
          trace_wavefront(g->stream, sv, r0->info.features, shade_fp, wb, cj, n, fixed_bounces, g->shade_sort_from, g->tail_from, g->trav_block,
                          nullptr);
          for (size_t i = 0; i < cj.size(); ++i)
            launch_frame_counter(g->stream, live[i]->d_frame.as<uint32_t>(), cj[i].fp.batch_frames, true);
          check(hipGetLastError(), "group launch");
        } catch (...) {
          (void)hipStreamEndCapture(g->stream, &gr);
          if (gr) (void)hipGraphDestroy(gr);
          throw;
        }
        check(hipStreamEndCapture(g->stream, &gr), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&g->gexec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        check(e, "hipGraphInstantiate");
        g->gkey = key;
        g->graph_captures++;
      }
      check(hipGraphLaunch(g->gexec, g->stream), "hipGraphLaunch");
      g->graph_launches++;
    } else {
      launch_job_limits(g->stream, g->d_limits.as<uint2>(), lim, njobs);
      check(hipGetLastError(), "job limits launch");
      trace_wavefront(g->stream, sv, r0->info.features, shade_fp, wb, wj, n, fixed_bounces, g->shade_sort_from, g->tail_from, g->trav_block,
                      nullptr);
      for (hpt_renderer* r : live) r->frame_synced = false;
    }
    for (size_t i = 0; i < live.size(); ++i) live[i]->frame_index += wj[i].fp.batch_frames;
    // Later work queued on each job's stream (reads, renders) waits for this step.
    check(hipEventRecord(g->done, g->stream), "hipEventRecord");
    for (hpt_renderer* r : live) check(hipStreamWaitEvent(r->stream, g->done, 0), "hipStreamWaitEvent");
    check(hipEventRecord(g->step_end[g->steps & 1], g->stream), "hipEventRecord");
    g->steps++;
  });
}

hpt_status hpt_group_get_stats(hpt_group* g, uint64_t* steps, uint64_t* graph_launches, uint64_t* device_bytes) {
  return guarded([&] {
    need(ok_group(g) && steps && graph_launches && device_bytes, "null argument");
    uint64_t b = 0;
    for (const DevBuf* d : {&g->d_rng, &g->d_org, &g->d_dir, &g->d_thr, &g->d_rad, &g->d_alb, &g->d_nrm, &g->d_med,
                            &g->d_med2, &g->d_nee_e, &g->d_nee_l, &g->d_hit, &g->d_job, &g->d_qa, &g->d_qb,
                            &g->d_sh_o, &g->d_sh_d, &g->d_sh_c, &g->d_counts, &g->d_spill, &g->d_sort_keys,
                            &g->d_sort_items, &g->d_limits, &g->d_nee_dir, &g->d_nee_li, &g->d_nee_ld,
                            &g->d_nee_tr, &g->d_qc, &g->d_tail_cut})
      b += d->bytes;
    *steps = g->steps;
    *graph_launches = g->graph_launches;
    *device_bytes = b;
  });
}

hpt_status hpt_group_get_truncated_paths(hpt_group* g, uint64_t* paths) {
  return guarded([&] {
    need(ok_group(g) && paths, "null argument");
    *paths = 0;
    if (!g->d_tail_cut.p) return;
    check(hipSetDevice(g->device), "hipSetDevice");
    unsigned long long v = 0;
    check(hipMemcpyAsync(&v, g->d_tail_cut.p, sizeof(v), hipMemcpyDeviceToHost, g->stream), "hipMemcpyAsync");
    check(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
    *paths = v;
  });
}

hpt_status hpt_group_wait_idle(hpt_group* g) {
  return guarded([&] {
    need(ok_group(g), "null or invalid group handle");
    check(hipSetDevice(g->device), "hipSetDevice");
    check(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
  });
}

void hpt_group_destroy(hpt_group* g) {
  if (!ok_group(g)) return;
  g->tag = 0;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->done) (void)hipEventDestroy(g->done);
  for (hipEvent_t e : g->step_end)
    if (e) (void)hipEventDestroy(e);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

}  // extern "C"

namespace hpt {
// The calling thread's message for hpt_last_error(), for entry points
// implemented in other files of the library (the denoiser's).
void set_last_error(const std::string& message) { g_last_error = message; }
}  // namespace hpt
