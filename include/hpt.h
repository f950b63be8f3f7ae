/*
 * hpt.h — C ABI of the MI355X (gfx950) path tracer.
 *
 * This library replaces the GPU half of the renderer that the reference host
 * drives (`hala_renderer::rt_renderer::HalaRenderer`, called from
 * src/bin/hala-pathtracer.rs:128-285) together with the GLSL programs under
 * shaders/src/hala-pathtracer/.  The host keeps its scene loader, its YAML
 * config and its camera/light/material structs; it hands those structs to
 * this library in their std140 layout (shaders/src/hala-pathtracer/inc/types.glsl)
 * and calls the entry points below instead of the Vulkan renderer.
 *
 * Conventions
 *  - Every entry point returns hpt_status (0 = HPT_OK).  On failure the
 *    calling thread's message is available from hpt_last_error().
 *    C++ exceptions never cross this boundary.
 *  - Inputs are borrowed for the duration of the call only and are copied to
 *    device memory.  The handle owns every device allocation it makes.
 *  - One handle is externally synchronised (the reference calls its renderer
 *    through `&mut self` from one thread).  Distinct handles are independent.
 *  - Work is queued on the handle's own HIP stream; hpt_read_image and
 *    hpt_wait_idle synchronise with it.
 */
#ifndef HPT_H
#define HPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPT_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
typedef int32_t hpt_status;
enum {
  HPT_OK = 0,
  HPT_ERROR_INVALID_ARGUMENT = 1, /* bad pointer, size or enum value       */
  HPT_ERROR_OUT_OF_MEMORY = 2,    /* device or host allocation failed      */
  HPT_ERROR_DEVICE = 3,           /* a HIP call failed                     */
  HPT_ERROR_STATE = 4,            /* call made in the wrong order          */
  HPT_ERROR_UNSUPPORTED = 5       /* feature or file variant not supported */
};

/* ------------------------------------------------------------------------ */
/* Constants shared with the GLSL ABI (inc/constants.glsl:1-44)              */
/* ------------------------------------------------------------------------ */
#define HPT_MAX_CAMERAS 8
#define HPT_MAX_LIGHTS 16
#define HPT_INVALID_INDEX 0xFFFFFFFFu
#define HPT_WORLD_SIZE 10000.0f

enum { HPT_CAMERA_PERSPECTIVE = 0, HPT_CAMERA_ORTHOGRAPHIC = 1 };
enum {
  HPT_LIGHT_POINT = 0,
  HPT_LIGHT_DIRECTIONAL = 1,
  HPT_LIGHT_SPOT = 2,
  HPT_LIGHT_QUAD = 3,
  HPT_LIGHT_SPHERE = 4
};
enum { HPT_ENV_SKY = 0, HPT_ENV_MAP = 1 };
enum { HPT_MEDIUM_NONE = 0, HPT_MEDIUM_ABSORB = 1, HPT_MEDIUM_SCATTER = 2, HPT_MEDIUM_EMISSIVE = 3 };
enum { HPT_MATERIAL_DIFFUSE = 0, HPT_MATERIAL_DISNEY = 1 };

/* Shader permutation bits (shaders/src/make_shaders.yaml:7-22).  The
 * reference host selects MEDIUM and VOL_MIS together when the scene has a
 * medium and TRANSPARENT when it has a transparent material
 * (src/bin/hala-pathtracer.rs:101-114).  Every one of the six combinations
 * the YAML lists is accepted here.  ROUGHNESS_REGULARIZE_EVAL is always on
 * (make_shaders.yaml:4). */
enum {
  HPT_FEATURE_MEDIUM = 1u << 0,
  HPT_FEATURE_VOL_MIS = 1u << 1,
  HPT_FEATURE_TRANSPARENT = 1u << 2
};

/* ------------------------------------------------------------------------ */
/* std140 structs (inc/types.glsl).  Byte offsets are asserted at the end.   */
/* ------------------------------------------------------------------------ */

/* Camera, inc/types.glsl:2-11 (80-byte std140 array stride). */
typedef struct hpt_camera {
  float position[3];
  float _pad0;
  float right[3];
  float _pad1;
  float up[3];
  float _pad2;
  float forward[3];
  float yfov;                    /* radians */
  float focal_distance_or_xmag;  /* perspective: focal distance; ortho: xmag */
  float aperture_or_ymag;        /* perspective: aperture; ortho: ymag */
  uint32_t type;                 /* HPT_CAMERA_* */
  float _pad3;
} hpt_camera;

/* Light, inc/types.glsl:14-39 (80-byte std140 array stride). */
typedef struct hpt_light {
  float intensity[3];
  float _pad0;
  float position[3];
  float _pad1;
  float u[3];
  float _pad2;
  float v[3];
  float radius;
  float area;
  int32_t type; /* HPT_LIGHT_* */
  float _pad3[2];
} hpt_light;

/* Medium, inc/types.glsl:42-49. */
typedef struct hpt_medium {
  float color[3];
  float density;
  float anisotropy;
  uint32_t type; /* HPT_MEDIUM_* */
  float _padding0;
  float _padding1;
} hpt_medium;

/* Material, inc/types.glsl:52-83 (144 bytes). */
typedef struct hpt_material {
  hpt_medium medium;
  float base_color[3];
  float opacity;
  float emission[3];
  float anisotropic;
  float metallic;
  float roughness;
  float subsurface;
  float specular_tint;
  float sheen;
  float sheen_tint;
  float clearcoat;
  float clearcoat_roughness;
  float clearcoat_tint[3];
  float specular_transmission;
  float ior;
  float ax;
  float ay;
  uint32_t base_color_map_index;
  uint32_t normal_map_index;
  uint32_t metallic_roughness_map_index;
  uint32_t emission_map_index;
  uint32_t type; /* HPT_MATERIAL_* */
} hpt_material;

/* Vertex, inc/types.glsl:86-98 (std430, 44 bytes, 4-byte aligned). */
typedef struct hpt_vertex {
  float position[3];
  float normal[3];
  float tangent[3];
  float tex_coord[2];
} hpt_vertex;

/* One mesh primitive: the host-side view of Primitive (inc/types.glsl:109-114)
 * plus the vertex/index buffers that the GLSL reaches through
 * buffer_reference.  `transform` is a column-major mat4 (GLSL / glam order):
 * world = transform * vec4(position, 1). */
typedef struct hpt_primitive_desc {
  float transform[16];
  uint32_t material_index;
  uint32_t vertex_count;
  uint32_t index_count; /* multiple of 3 */
  uint32_t _pad;
  const hpt_vertex* vertices;
  const uint32_t* indices;
} hpt_primitive_desc;

typedef struct hpt_scene_desc {
  const hpt_camera* cameras;
  uint32_t camera_count; /* 1..HPT_MAX_CAMERAS */
  uint32_t light_count;  /* 0..HPT_MAX_LIGHTS */
  const hpt_light* lights;
  const hpt_material* materials;
  uint32_t material_count;
  uint32_t primitive_count;
  const hpt_primitive_desc* primitives;
} hpt_scene_desc;

/* Arguments of HalaRenderer::new (src/bin/hala-pathtracer.rs:128-138) minus
 * the window, plus the feature bits that choose the shader permutation
 * (hala-pathtracer.rs:101-114) and the exposure set at :211. */
typedef struct hpt_create_info {
  uint32_t width;
  uint32_t height;
  uint32_t max_depth;
  uint32_t rr_depth;
  uint32_t enable_tonemap;
  uint32_t enable_aces;
  uint32_t use_simple_aces;
  uint32_t features;     /* HPT_FEATURE_* */
  uint64_t max_samples;  /* 0 = unlimited */
  float exposure_value;
  int32_t device;        /* HIP device ordinal */
} hpt_create_info;

/* Images the renderer keeps resident (inc/global.glsl:2-4). */
typedef enum hpt_image {
  HPT_IMAGE_FINAL_RGBA8 = 0,  /* width*height*4 bytes, tonemapped sRGB */
  HPT_IMAGE_ACCUM_RGBA32F = 1, /* width*height*16 bytes, running mean radiance */
  HPT_IMAGE_ALBEDO_RGBA32F = 2,
  HPT_IMAGE_NORMAL_RGBA32F = 3
} hpt_image;

/* Ray for the trace entry points: t is searched on [tmin, tmax] inclusive. */
typedef struct hpt_ray {
  float origin[3];
  float tmin;
  float direction[3];
  float tmax;
} hpt_ray;

enum { HPT_HIT_NONE = 0, HPT_HIT_TRIANGLE = 1, HPT_HIT_LIGHT = 2 };

/* Closest hit.  For triangles `instance` is the primitive index (the GLSL's
 * gl_InstanceCustomIndexEXT, triangles.rchit.glsl:17) and `primitive` the
 * triangle index inside it (gl_PrimitiveID, :24-26); u,v are the barycentric
 * weights of the triangle's second and third vertex (hit attribute `attribs`,
 * :58).  For analytic lights `instance` is the light index (lights.rint.glsl:21)
 * and `primitive` is HPT_INVALID_INDEX. */
typedef struct hpt_hit {
  float t;
  float u;
  float v;
  uint32_t kind; /* HPT_HIT_* */
  uint32_t instance;
  uint32_t primitive;
} hpt_hit;

typedef struct hpt_renderer hpt_renderer;

/* ------------------------------------------------------------------------ */
/* Entry points                                                              */
/* ------------------------------------------------------------------------ */

/* HalaRenderer::new (hala-pathtracer.rs:128-138). */
HPT_API hpt_status hpt_create(const hpt_create_info* info, hpt_renderer** out);

/* HalaRenderer::set_scene (hala-pathtracer.rs:198).  Copies cameras, lights,
 * materials and meshes; the BVH is built by hpt_commit. */
HPT_API hpt_status hpt_set_scene(hpt_renderer* r, const hpt_scene_desc* scene);

/* HalaRenderer::set_envmap (hala-pathtracer.rs:203).  `rgb` is width*height
 * linear RGB float triples, row 0 = theta 0 (+Z).  `rotation` is the u offset
 * of inc/env_map.glsl:18 / env_map_eval.rcall.glsl:18 in turns.  Builds the
 * luminance alias table that replaces the reference's marginal/conditional
 * distribution textures (inc/env_map.glsl:6-37).  Selects ENV_TYPE_MAP. */
HPT_API hpt_status hpt_set_env_map(hpt_renderer* r, const float* rgb, uint32_t width,
                                   uint32_t height, float rotation);

/* OpenEXR decoding for hpt_set_env_map: the reference's configurations name
 * OpenEXR HDRIs (`hdri_file` in conf/dragon.yaml and its siblings), which
 * HalaRenderer::set_envmap decodes from the path it is given
 * (hala-pathtracer.rs:202-203).  These take the file's bytes.  Read: single-part scanline files, NONE, RLE, ZIPS,
 * ZIP or PIZ compression, UINT, HALF or FLOAT channels without subsampling;
 * RGB from the channels R, G and B (a missing one reads 0), else Y copied to
 * all three.  Other files fail with HPT_ERROR_UNSUPPORTED, malformed ones
 * with HPT_ERROR_INVALID_ARGUMENT.  Host code only: no device is used. */
HPT_API hpt_status hpt_exr_info(const void* data, size_t size, uint32_t* width, uint32_t* height);
/* Writes height*width*3 floats, row 0 = the top line of the data window
 * (theta 0 of an equirectangular map). */
HPT_API hpt_status hpt_exr_decode_rgb(const void* data, size_t size, float* rgb, uint32_t width,
                                      uint32_t height);

/* HalaRenderer::set_ground_color / set_sky_color (hala-pathtracer.rs:205-206).
 * Either selects ENV_TYPE_SKY (env_sky_eval.rcall.glsl:12-15). */
HPT_API hpt_status hpt_set_ground_color(hpt_renderer* r, const float rgba[4]);
HPT_API hpt_status hpt_set_sky_color(hpt_renderer* r, const float rgba[4]);

/* HalaRenderer::set_env_intensity (hala-pathtracer.rs:208). */
HPT_API hpt_status hpt_set_env_intensity(hpt_renderer* r, float intensity);

/* HalaRenderer::set_exposure_value (hala-pathtracer.rs:211). */
HPT_API hpt_status hpt_set_exposure_value(hpt_renderer* r, float ev);

/* HalaRenderer::load_blue_noise_texture (hala-pathtracer.rs:195): RGBA8 texels
 * read by rand4blue (common/wltcRS.glsl:46-53).  Without this call the
 * renderer uses its built-in 1024x1024 procedural substitute. */
HPT_API hpt_status hpt_set_blue_noise(hpt_renderer* r, const uint8_t* rgba8, uint32_t width,
                                      uint32_t height);

/* Material textures: the g_textures[] table (inc/global.glsl:48) that the
 * Material *_map_index fields index (pathtracer.rgen.glsl:122-146), which the
 * reference renderer fills from the scene's glTF images inside set_scene.
 * Each texture is width*height RGBA8 texels, row 0 at v = 0.  Sampling is
 * bilinear with REPEAT in u and v at level 0; base-colour and emission maps
 * are decoded from sRGB per texel, metallic-roughness (B = metallic,
 * G = roughness) and normal maps are read as linear.  Replaces the whole
 * table; the materials' indices are checked against it at hpt_commit. */
typedef struct hpt_texture_desc {
  uint32_t width;
  uint32_t height;
  const uint8_t* rgba8;
} hpt_texture_desc;

HPT_API hpt_status hpt_set_textures(hpt_renderer* r, const hpt_texture_desc* textures, uint32_t count);

/* HalaRenderer::commit (hala-pathtracer.rs:213): builds the BVH, uploads
 * everything and resets accumulation. */
HPT_API hpt_status hpt_commit(hpt_renderer* r);

/* The per-frame update()+render() pair (hala-pathtracer.rs:264-285) without
 * presentation: renders `frames` one-sample frames with the camera
 * `camera_index`, each folded into the running mean of
 * pathtracer.rgen.glsl:259-304.  Asynchronous on the handle's stream.
 * Frames beyond max_samples (when non-zero) are skipped. */
HPT_API hpt_status hpt_render(hpt_renderer* r, uint32_t frames, uint32_t camera_index);

/* Restart accumulation at frame 0 (what the renderer does when the camera or
 * scene changes); the accumulated images are cleared to zero. */
HPT_API hpt_status hpt_reset_accumulation(hpt_renderer* r);

/* How hpt_render submits a batch (no counterpart in the reference, which
 * records one command buffer per frame).  HPT_LAUNCH_GRAPH, the default,
 * captures a whole batch -- camera, the bounce loop, accumulation -- as a HIP
 * graph on first use and replays it with one launch for every later full
 * batch with the same parameters (a device frame counter that the graph
 * advances supplies the frame index); a change of scene, camera, settings,
 * batch size or buffers captures it again.  HPT_LAUNCH_EAGER launches each
 * kernel.  Both give the same images bit for bit.  Partial batches, profiled
 * renders and the TRANSPARENT permutation (whose bounce count depends on a
 * host read-back) always launch eagerly. */
enum { HPT_LAUNCH_EAGER = 0, HPT_LAUNCH_GRAPH = 1 };

/* Where hpt_commit builds the BVH (no counterpart in the reference, whose
 * acceleration structures the Vulkan driver builds).  HPT_BVH_GPU, the
 * default, runs the binned-SAH build and the 8-wide collapse on the device;
 * HPT_BVH_HOST runs the same algorithm on the CPU.  Both produce the same
 * bytes; the device build hands over to the host one when a node of more
 * than 1024 triangles has all centroids at one point. */
enum { HPT_BVH_HOST = 0, HPT_BVH_GPU = 1 };
HPT_API hpt_status hpt_set_bvh_builder(hpt_renderer* r, uint32_t builder);
/* The builder the last hpt_commit used and its wall time in milliseconds. */
HPT_API hpt_status hpt_get_bvh_build_stats(hpt_renderer* r, uint32_t* builder_used, double* build_ms);
HPT_API hpt_status hpt_set_launch_mode(hpt_renderer* r, uint32_t mode);

/* Material-sorted shading (no counterpart in the reference, whose ray
 * tracing pipeline dispatches hit shaders itself).  From bounce
 * `first_bounce` on, the extension kernel writes each hit's material bucket
 * and the paths of every queue segment are grouped by it before shading, so
 * a shading wave mostly runs one material's code and parameters.  Images are
 * the same bit for bit whatever the setting; it changes only speed, and the
 * autotuner measures it.  HPT_SHADE_SORT_OFF (the default) never sorts. */
#define HPT_SHADE_SORT_OFF 0xFFFFFFFFu
HPT_API hpt_status hpt_set_shade_sort(hpt_renderer* r, uint32_t first_bounce);

/* Volume-MIS shading (MEDIUM|VOL_MIS, inc/lighting.glsl:7-61) in phases (no
 * counterpart in the reference, where the walks are traceRayEXT calls inside
 * the raygen shader).  on = 1, the default: each bounce's shading runs as a
 * kernel up to the first transmittance walk, a kernel of the queued walks,
 * (for environment maps a second shading kernel and walk kernel,) and a
 * kernel that finishes the bounce, so no shading state is live during a
 * traversal; on = 0: one shading kernel with the walks inline.  The images
 * are the same bit for bit.  Other permutations ignore it. */
HPT_API hpt_status hpt_set_vmis_phases(hpt_renderer* r, uint32_t on);

/* Several jobs in one wavefront (no counterpart in the reference, which
 * renders one scene per process).  A group owns a path state, queues and a
 * stream of its own; hpt_group_render traces `frames[i]` frames of job
 * `jobs[i]` (camera `camera_indices[i]`, or 0 when that pointer is null) for
 * all jobs together: one camera and one accumulation launch per job over its
 * slice of the wavefront, and one extension, shading and shadow launch per
 * bounce for the paths of every job, each path shaded with its own job's
 * depth limits.  The jobs must share one committed scene (hpt_share_scene),
 * shader permutation and environment; resolution, depths, frame counts and
 * cameras may differ.  Every job's images and frame counter end exactly as
 * if it had rendered those frames alone with hpt_render.  The step is queued
 * on the group's stream after each job's earlier work, and later work on a
 * job's own stream (hpt_read_image, hpt_wait_idle, hpt_render) waits for it.
 * With HPT_LAUNCH_GRAPH (the default) a step that repeats the previous one
 * (same jobs, frames and settings) is captured once and replayed as a HIP
 * graph.  At most 256 jobs per step. */
typedef struct hpt_group hpt_group;
HPT_API hpt_status hpt_group_create(int32_t device, hpt_group** out);
HPT_API hpt_status hpt_group_render(hpt_group* g, hpt_renderer* const* jobs, const uint32_t* frames,
                                    const uint32_t* camera_indices, uint32_t count);
HPT_API hpt_status hpt_group_set_launch_mode(hpt_group* g, uint32_t mode);
HPT_API hpt_status hpt_group_set_shade_sort(hpt_group* g, uint32_t first_bounce);
/* hpt_set_vmis_phases for a group's steps (phased, 1, by default). */
HPT_API hpt_status hpt_group_set_vmis_phases(hpt_group* g, uint32_t on);
/* Steps rendered, steps replayed as graphs, device bytes the group holds. */
HPT_API hpt_status hpt_group_get_stats(hpt_group* g, uint64_t* steps, uint64_t* graph_launches,
                                       uint64_t* device_bytes);
HPT_API hpt_status hpt_group_wait_idle(hpt_group* g);
HPT_API void hpt_group_destroy(hpt_group* g);
/* The bounce from which the paths still live run to their end in one kernel,
 * each path on one lane (closest hit, shading and shadow rays in the order
 * of the wavefront: same images), instead of one wavefront launch per
 * bounce (no counterpart in the reference, whose rays run that way).  Late
 * bounces hold few paths, and each wavefront launch over them costs the
 * dependent-load latency of its slowest path; the tail pays one such chain
 * per path instead of one per bounce.  first_bounce >= 1;
 * HPT_TAIL_FROM_NONE (the default) keeps every bounce up to max_depth + 1
 * in the wavefront, and only TRANSPARENT paths that outlive those (a
 * pass-through does not count toward depth, pathtracer.rgen.glsl:196-200)
 * go to the tail.  The best value depends on scene and batch size
 * (autotune.py picks it). */
#define HPT_TAIL_FROM_NONE 0xFFFFFFFFu
HPT_API hpt_status hpt_set_tail_from(hpt_renderer* r, uint32_t first_bounce);
HPT_API hpt_status hpt_group_set_tail_from(hpt_group* g, uint32_t first_bounce);

/* Threads per block of the traversal kernels (closest-hit extension, shadow
 * rays, the volume-MIS transmittance walks): 64, 128 (the default) or 256
 * (no counterpart in the reference).  Same images; the autotuner measures
 * it. */
HPT_API hpt_status hpt_set_traversal_block(hpt_renderer* r, uint32_t threads);
HPT_API hpt_status hpt_group_set_traversal_block(hpt_group* g, uint32_t threads);

/* Lane refill in the traversal kernels (no counterpart in the reference):
 * with it on (the default), the extension and shadow rays, the
 * transmittance walks and the tail's paths run in blocks that stay
 * resident, and a wave whose lanes' rays have ended takes new queue
 * entries for them instead of stepping on with most lanes idle until its
 * longest ray ends.  Off: one queue entry per lane.  Same images either
 * way. */
HPT_API hpt_status hpt_set_traversal_refill(hpt_renderer* r, uint32_t on);
HPT_API hpt_status hpt_group_set_traversal_refill(hpt_group* g, uint32_t on);

/* Batches replayed as graphs and graphs captured since the handle was made. */
HPT_API hpt_status hpt_get_launch_stats(hpt_renderer* r, uint64_t* graph_launches, uint64_t* graph_captures);

/* Paths stopped at the bounce limit since the handle was made.  With
 * TRANSPARENT a pass-through does not count toward max_depth
 * (pathtracer.rgen.glsl:196-200), so the number of bounces a path takes has
 * no fixed bound; the library stops a path after 4096 bounces (closest-hit
 * queries), keeps the radiance it gathered and counts it here.  A non-zero
 * count means the images of those frames are biased.  Waits for the
 * handle's (the group's) queued work. */
HPT_API hpt_status hpt_get_truncated_paths(hpt_renderer* r, uint64_t* paths);
HPT_API hpt_status hpt_group_get_truncated_paths(hpt_group* g, uint64_t* paths);

/* Sets the frame counter (the UBO frame_index) the next frame renders with.
 * The RNG is a pure function of (pixel, frame) (common/wltcRS.glsl:9-19), so
 * a job resumes from its accumulated images plus this counter. */
HPT_API hpt_status hpt_set_frame_index(hpt_renderer* r, uint64_t frame_index);

/* Number of frames accumulated so far (the UBO frame_index). */
HPT_API hpt_status hpt_get_frame_index(hpt_renderer* r, uint64_t* frame_index);

/* Copies one resident image into `dst` (the data behind
 * HalaRenderer::save_images, hala-pathtracer.rs:251).  Synchronises. */
HPT_API hpt_status hpt_read_image(hpt_renderer* r, hpt_image which, void* dst, size_t bytes);

/* Device address and size of one resident image, for consumers on the same
 * GPU (the learned denoiser reads colour, albedo and normal in place instead
 * of the PFM files denoise.py:18-19 hands to OIDN).  No counterpart in the
 * reference.  The address stays valid until hpt_destroy; work queued on the
 * handle's stream must have finished (hpt_wait_idle) before it is read. */
HPT_API hpt_status hpt_get_image_pointer(hpt_renderer* r, hpt_image which, void** ptr, size_t* bytes);

/* Closest-hit query over the committed scene: the semantics of traceRayEXT
 * with gl_RayFlagsOpaqueEXT (pathtracer.rgen.glsl:56-68): triangles are
 * double-sided, analytic quad/sphere lights are hit through
 * lights.rint.glsl:16-37.  Test and parity entry point; synchronous. */
HPT_API hpt_status hpt_trace_closest(hpt_renderer* r, const hpt_ray* rays, uint32_t count,
                                     hpt_hit* hits);

/* Any-hit (shadow) query: traceRayEXT with TerminateOnFirstHit|Opaque|
 * SkipClosestHitShader (inc/lighting.glsl:119-130,205-216).  occluded[i] = 1
 * when anything, triangles or analytic lights, lies on [tmin, tmax]. */
HPT_API hpt_status hpt_trace_any(hpt_renderer* r, const hpt_ray* rays, uint32_t count,
                                 uint32_t* occluded);

/* Closest-hit queries that allocate nothing (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * `rays` and `hits` are device addresses on the handle's device, `count`
 * entries each; the hits equal hpt_trace_closest's.  Queued on the handle's
 * stream: the call returns before the work ends (hpt_wait_idle).  For trees
 * deeper than the traversal's 8-entry LDS stack, each ray's further stack
 * entries live in a 64 MiB device array the library holds per GPU (found at
 * the first call on that GPU and kept until the process ends):
 * count * (tree levels - 8) must not exceed 8,388,608 entries, else
 * HPT_ERROR_INVALID_ARGUMENT.  All calls on one GPU share that array: calls
 * that overlap in time on one GPU (several handles or streams) race, so the
 * caller serialises them. */
HPT_API hpt_status hpt_trace_closest_device(hpt_renderer* r, const hpt_ray* rays, uint32_t count, hpt_hit* hits);

/* Presentation: the hand-off of finished frames to a window loop (the
 * reference's update()/render() pair presents the final image through its
 * swapchain, hala-pathtracer.rs:236-289).  A presenter keeps `slots` (2..16)
 * pinned host copies of the handle's final rgba8 image (width*height*4
 * bytes each).  hpt_present_frame, called after hpt_render, queues a copy of
 * the image as it stands after the frames queued so far: a device copy on
 * the handle's stream, then the host copy on a stream of the presenter's
 * own, so the tracer never waits for it; when every slot is in flight or
 * held the frame is dropped (counted) instead.  hpt_present_acquire returns,
 * without waiting, the newest finished copy and the frame count it was
 * taken at (pixels = NULL when none has finished); that slot stays valid and
 * unchanged until the next hpt_present_acquire.  Destroy a presenter before
 * its renderer. */
typedef struct hpt_presenter hpt_presenter;
HPT_API hpt_status hpt_present_create(hpt_renderer* r, uint32_t slots, hpt_presenter** out);
HPT_API hpt_status hpt_present_frame(hpt_presenter* p);
HPT_API hpt_status hpt_present_acquire(hpt_presenter* p, const void** pixels, uint64_t* frame_index);
/* Frames queued for presentation and frames dropped because no slot was free. */
HPT_API hpt_status hpt_present_get_stats(hpt_presenter* p, uint64_t* presented, uint64_t* dropped);
HPT_API void hpt_present_destroy(hpt_presenter* p);

/* Width and height the handle was created with. */
HPT_API hpt_status hpt_get_image_size(hpt_renderer* r, uint32_t* width, uint32_t* height);

/* The learned denoiser (replaces the OIDN step of denoise.py:26-30 that the
 * Save button's images go through, hala-pathtracer.rs:246-252).
 * hpt_denoiser_create loads a trained network from the bytes of the .npz
 * file the training run writes (arrays `levels`, `params` float32,
 * `residual`; weights/unet_denoiser.npz ships with the library) onto
 * `device`.  hpt_denoise waits for the renderer's queued frames, runs the
 * network on its resident colour, albedo and normal images with its frame
 * count as the sample count, and writes the denoised colour as
 * width*height float4 (RGB, alpha 1) into `dst`: host memory, or device
 * memory on the denoiser's device when dst_on_device is non-zero; `bytes`
 * must be width*height*16.  Synchronous.  One handle is externally
 * synchronised; its activation buffers follow the last image size. */
typedef struct hpt_denoiser hpt_denoiser;
HPT_API hpt_status hpt_denoiser_create(const void* npz, size_t size, int32_t device, hpt_denoiser** out);
HPT_API hpt_status hpt_denoise(hpt_denoiser* d, hpt_renderer* r, float* dst, size_t bytes, int dst_on_device);
HPT_API void hpt_denoiser_destroy(hpt_denoiser* d);

/* HalaRenderer::wait_idle (hala-pathtracer.rs:228). */
HPT_API hpt_status hpt_wait_idle(hpt_renderer* r);

/* Drops the renderer and all its device memory. */
HPT_API void hpt_destroy(hpt_renderer* r);

/* Message of the last failed call on this thread ("" if none). */
HPT_API const char* hpt_last_error(void);

/* ------------------------------------------------------------------------ */
/* Introspection (no counterpart in the reference: the Vulkan driver keeps  */
/* its acceleration structure opaque).  Used by parity tests and the        */
/* roofline accounting, which need the exact node and triangle arrays.      */
/* ------------------------------------------------------------------------ */

/* Sizes of the committed BVH: 8-wide compressed nodes (80 bytes each) and
 * leaf triangles (48 bytes each, in leaf order). */
typedef struct hpt_bvh_info {
  uint32_t node_count;
  uint32_t triangle_count;
  uint32_t max_depth;   /* levels of the 8-wide tree */
  uint32_t _pad;
} hpt_bvh_info;

HPT_API hpt_status hpt_get_bvh_info(hpt_renderer* r, hpt_bvh_info* info);

/* Copies the node array (node_count*80 bytes) and leaf triangle array
 * (triangle_count*48 bytes) to host memory. */
HPT_API hpt_status hpt_export_bvh(hpt_renderer* r, void* nodes, size_t node_bytes,
                                  void* triangles, size_t triangle_bytes);

/* Per-ray BVH node visits and triangle tests of the closest-hit (any_hit = 0)
 * or shadow (any_hit = 1) walk: the unit counts behind the roofline's
 * algorithmic bytes, 80 B per node + 48 B per triangle + 48 B per ray. */
HPT_API hpt_status hpt_trace_counts(hpt_renderer* r, const hpt_ray* rays, uint32_t count, int any_hit,
                                    uint32_t* node_visits, uint32_t* tri_tests);

/* Kernel timing with HIP events recorded on the handle's stream around each
 * launch.  `stage` is one of the HPT_STAGE_* values.  Any non-zero value
 * resets the counters and records per-launch timings and ray counts;
 * HPT_PROFILE_COUNT_VISITS additionally makes the extend and shadow kernels
 * count BVH node visits and triangle tests (a slower kernel variant, for
 * roofline accounting only). */
enum { HPT_PROFILE_TIMING = 1, HPT_PROFILE_COUNT_VISITS = 2 };
enum {
  HPT_STAGE_CAMERA = 0,
  HPT_STAGE_EXTEND = 1,  /* closest-hit traversal of the path queue */
  HPT_STAGE_SHADE = 2,   /* surface/medium interaction, NEE, BSDF sampling */
  HPT_STAGE_SHADOW = 3,  /* any-hit traversal of the shadow queue */
  HPT_STAGE_ACCUMULATE = 4,
  HPT_STAGE_SORT = 5,    /* material sort of the shading queue (hpt_set_shade_sort) */
  HPT_STAGE_WALK = 6,    /* transmittance walks of the phased volume-MIS shading */
  HPT_STAGE_TAIL = 7,    /* TRANSPARENT: the paths left after max_depth + 1 bounces,
                            each run to its end by one lane (rays = such paths,
                            queries = the bounces they took) */
  HPT_STAGE_COUNT = 8
};
typedef struct hpt_stage_stats {
  uint64_t launches;
  double total_ms;
  uint64_t rays;        /* rays the stage consumed (device-side counts); for
                           HPT_STAGE_SHADE the shading events */
  uint64_t node_visits; /* with HPT_PROFILE_COUNT_VISITS; for HPT_STAGE_SHADE
                           those of the volume-MIS transmittance walks */
  uint64_t tri_tests;   /* with HPT_PROFILE_COUNT_VISITS (likewise) */
  uint64_t queries;     /* with HPT_PROFILE_COUNT_VISITS: closest-hit queries
                           of the transmittance walks (inc/lighting.glsl:19-32)
                           -- HPT_STAGE_SHADE's inline walks, or HPT_STAGE_WALK's
                           when shading is phased; 0 for other stages.  For
                           HPT_STAGE_WALK, rays = walks */
} hpt_stage_stats;

HPT_API hpt_status hpt_set_profiling(hpt_renderer* r, int enable);
HPT_API hpt_status hpt_get_stage_stats(hpt_renderer* r, int stage, hpt_stage_stats* out);

/* ------------------------------------------------------------------------ */
/* Serving (no counterpart in the reference, which renders one scene per   */
/* process).  Many jobs over a few resident scenes share each scene's      */
/* device copy; each job is a handle with its own stream, images and path  */
/* state.                                                                   */
/* ------------------------------------------------------------------------ */

/* Makes `dst` render the committed scene of `src` without another BVH build
 * or upload: the device copy (BVH, shading arrays, environment, textures,
 * blue noise) is shared and freed with the last handle using it.  `dst`
 * keeps its own size, depths, tonemap, features and exposure; the sky
 * colours and environment intensity are copied from `src` and may be set per
 * handle afterwards.  `dst` becomes committed with cleared images.  While a
 * scene is shared, hpt_set_env_map and hpt_set_blue_noise on any of its
 * handles fail with HPT_ERROR_STATE. */
HPT_API hpt_status hpt_share_scene(hpt_renderer* dst, const hpt_renderer* src);

/* *idle = 1 when all work queued on the handle's stream has finished. */
HPT_API hpt_status hpt_query_idle(hpt_renderer* r, int* idle);

/* Device bytes the handle holds for its frames (path state, queues, images)
 * and for its (possibly shared) scene. */
HPT_API hpt_status hpt_get_memory_usage(hpt_renderer* r, uint64_t* frame_bytes, uint64_t* scene_bytes);

/* Frames traced together by hpt_render (no counterpart in the reference,
 * whose render loop submits one frame per command buffer).  A batch of k
 * frames is one wavefront of k * width * height paths; the accumulation
 * images fold the frames in order with the reference's per-frame weights,
 * so every k gives the images of frame-by-frame rendering.  Batching
 * amortises the short, latency-bound tail bounces over k frames.  frames = 0
 * sizes the batch from a 24 GiB path-state budget (at most 32 frames); 1..64
 * forces it (HPT_ERROR_INVALID_ARGUMENT above 64 or when k * width * height
 * exceeds 2^31).  Takes effect at the next hpt_render. */
HPT_API hpt_status hpt_set_frames_in_flight(hpt_renderer* r, uint32_t frames);
HPT_API hpt_status hpt_get_frames_in_flight(hpt_renderer* r, uint32_t* frames);

/* The serving loop (no counterpart in the reference).  A server takes
 * progressive render jobs over scenes committed once (resident scenes), admits
 * them in submission order while the sum of their device footprints fits a
 * budget (a job that does not fit waits, and holds back the jobs behind it so
 * large jobs are not starved), runs them concurrently and hands back each
 * job's final image in host memory, denoised when the job asks for it.
 *
 *  - HPT_SERVE_STREAMS: an admitted job is a handle of its own (sharing its
 *    scene's device copy, hpt_share_scene) on its own stream, rendered in
 *    batches of its frames in flight from a host thread of the server; the
 *    device interleaves the kernels of the jobs in flight.
 *  - HPT_SERVE_WAVEFRONT: the jobs of one resident scene share one wavefront
 *    per step (hpt_group_render): a host thread per scene renders the next
 *    chunk of about step_paths paths of every active job of that scene
 *    together; finished jobs leave after their step and jobs admitted
 *    meanwhile join the next one.
 *
 * A job's image is a function of its scene, size, depths and frame range
 * (the RNG is seeded by pixel and frame), so it is the same alone and among
 * any other jobs, in either mode, bit for bit.  A job's footprint is
 * pixels * (frames * 288 B of path state and queues [+ 5 B with shading sort]
 * + 52 B of images); in wavefront mode its share of its scene's wavefront
 * (chunk frames + 1) * 288 B per pixel plus its images.  Timestamps are
 * seconds on the server's clock (0 at hpt_server_create).  A failing job ends
 * with its status and message and leaves the others running.  The server
 * borrows its scenes' template handles and its denoiser: they must outlive
 * it.  Calls on one server may come from any host thread. */
typedef struct hpt_server hpt_server;
enum { HPT_SERVE_STREAMS = 0, HPT_SERVE_WAVEFRONT = 1 };
#define HPT_SERVE_DEFAULT 0xFFFFFFFFu /* a job field that takes the server's setting */

typedef struct hpt_server_config {
  int32_t device;
  uint32_t mode;               /* HPT_SERVE_STREAMS (default) or HPT_SERVE_WAVEFRONT */
  uint64_t budget_bytes;       /* device bytes the admitted jobs' footprints may sum to */
  uint32_t max_frames;         /* frames in flight of a job that does not set them (32) */
  uint32_t keep_images;        /* 1: finished jobs keep their image for hpt_server_read_image */
  uint64_t step_paths;         /* wavefront mode: paths per job and step (8 Mi) */
  uint32_t launch_mode;        /* HPT_LAUNCH_* of jobs that do not set it (HPT_LAUNCH_GRAPH) */
  uint32_t max_running;        /* streams mode: jobs admitted to run at once, 0 = as the budget allows;
                                  8 by default (C4 serving on MI355X: 1133 Msamples/s and p99 6.25 s at 8,
                                  about 1050-1090 and 6.4-6.7 s without a cap) */
  /* Dry run: 0 (the default) renders on the device.  Non-zero runs the
   * admission and scheduling without touching a device: a batch of n frames
   * of a job takes n * pixels * dry_run_ns_per_path nanoseconds of host time
   * and its image is filled with the job id -- for planning a budget on a
   * host without a GPU, and for testing the scheduling. */
  uint64_t dry_run_ns_per_path;
  /* 1 (the default): a finished job's handle is kept for a later job of the
   * same class (scene, size, depths, tonemap, exposure, batch and tuning),
   * which then renders without making a handle, allocating its frame state
   * or capturing its batch graph again.  Kept handles hold their footprint
   * within the budget and are freed, least recently kept first, when a job
   * needs the room.  0 frees every handle when its job ends. */
  uint32_t reuse_handles;
  uint32_t _reserved;
} hpt_server_config;
HPT_API void hpt_server_config_init(hpt_server_config* c);

typedef struct hpt_serve_scene_desc {
  const hpt_renderer* scene;   /* a committed handle whose scene the jobs render */
  uint32_t features;           /* HPT_FEATURE_* of the jobs (the shader permutation) */
  /* Settings of this scene's wavefront steps (HPT_SERVE_DEFAULT: the
   * library's defaults).  Streams-mode jobs take their own (hpt_job_desc). */
  uint32_t step_launch_mode, step_shade_sort, step_tail_from, step_traversal_block;
  /* Dry runs only: 1 makes every job of the scene fail when it starts (its
   * handle cannot be made), for testing failure isolation. */
  uint32_t dry_run_fail;
} hpt_serve_scene_desc;

typedef struct hpt_job_desc {
  uint32_t scene;              /* from hpt_server_add_scene */
  uint32_t width, height;
  uint32_t spp;                /* frames of one sample per pixel */
  uint32_t max_depth, rr_depth;
  uint32_t camera_index;
  uint32_t enable_tonemap, enable_aces, use_simple_aces;
  float exposure_value;
  /* Per-job tuning (autotune.py measures these per scene and resolution);
   * HPT_SERVE_DEFAULT takes the server's setting or the library default. */
  uint32_t frames_in_flight;   /* streams mode: frames per batch (capped at spp) */
  uint32_t launch_mode, shade_sort, tail_from, traversal_block;
  uint32_t denoise;            /* 1: the result is the denoised colour (hpt_server_set_denoiser) */
} hpt_job_desc;
HPT_API void hpt_job_desc_init(hpt_job_desc* d);

typedef struct hpt_job_result {
  uint64_t job_id;
  hpt_status status;           /* HPT_OK, or why the job failed */
  uint32_t width, height;
  uint32_t frames;             /* frames rendered */
  uint32_t has_image;          /* the image is held for hpt_server_read_image */
  uint32_t _pad;
  uint64_t footprint_bytes;
  double submitted_s, admitted_s, finished_s;  /* finished = image in host memory */
  char error[256];
} hpt_job_result;

typedef struct hpt_server_stats {
  uint64_t submitted, finished, failed;
  uint64_t bytes_in_use, peak_bytes;
  uint64_t wavefront_steps, max_jobs_per_step;
  uint64_t jobs_in_steps;      /* sum over wavefront steps of the jobs each held */
  uint64_t handles_reused;     /* jobs that rendered on a kept handle */
  uint64_t cached_bytes;       /* footprint of the handles kept now */
} hpt_server_stats;

HPT_API hpt_status hpt_server_create(const hpt_server_config* config, hpt_server** out);
/* Adds a resident scene; *scene_id is the index jobs name it by. */
HPT_API hpt_status hpt_server_add_scene(hpt_server* s, const hpt_serve_scene_desc* scene, uint32_t* scene_id);
/* The network that denoises jobs with `denoise` set (borrowed; NULL clears). */
HPT_API hpt_status hpt_server_set_denoiser(hpt_server* s, hpt_denoiser* d);
/* What admitting `job` would charge against the budget and its frames per
 * batch (streams) or per step (wavefront). */
HPT_API hpt_status hpt_server_plan_job(hpt_server* s, const hpt_job_desc* job, uint64_t* footprint_bytes,
                                       uint32_t* chunk_frames);
/* Queues a job and returns at once; HPT_ERROR_INVALID_ARGUMENT when its
 * footprint exceeds the whole budget. */
HPT_API hpt_status hpt_server_submit(hpt_server* s, const hpt_job_desc* job, uint64_t* job_id);
/* Waits until `job_id` has finished and describes it.  Its result stays
 * listed for hpt_server_next_result. */
HPT_API hpt_status hpt_server_wait(hpt_server* s, uint64_t job_id, hpt_job_result* out);
/* The next finished job in completion order, waiting up to timeout_s
 * seconds (negative: no limit); *got = 0 when none finished in time. */
HPT_API hpt_status hpt_server_next_result(hpt_server* s, double timeout_s, hpt_job_result* out, int* got);
/* Copies a finished job's image (width*height float4: accumulated radiance,
 * or the denoised colour with alpha 1) and releases it. */
HPT_API hpt_status hpt_server_read_image(hpt_server* s, uint64_t job_id, float* dst, size_t bytes);
/* Waits until every submitted job has finished. */
HPT_API hpt_status hpt_server_drain(hpt_server* s);
/* Seconds on the server's clock. */
HPT_API hpt_status hpt_server_clock(hpt_server* s, double* seconds);
HPT_API hpt_status hpt_server_get_stats(hpt_server* s, hpt_server_stats* out);
/* Drains, then stops the server's threads and frees what it holds (kept
 * handles included). */
HPT_API void hpt_server_destroy(hpt_server* s);

#ifdef __cplusplus
} /* extern "C" */
#endif

/* ------------------------------------------------------------------------ */
/* Layout checks against the std140 offsets of inc/types.glsl.              */
/* ------------------------------------------------------------------------ */
#ifdef __cplusplus
#define HPT_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define HPT_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
HPT_STATIC_ASSERT(sizeof(hpt_camera) == 80, "Camera std140 stride");
HPT_STATIC_ASSERT(offsetof(hpt_camera, right) == 16, "Camera.right");
HPT_STATIC_ASSERT(offsetof(hpt_camera, forward) == 48, "Camera.forward");
HPT_STATIC_ASSERT(offsetof(hpt_camera, yfov) == 60, "Camera.yfov");
HPT_STATIC_ASSERT(offsetof(hpt_camera, type) == 72, "Camera.type");
HPT_STATIC_ASSERT(sizeof(hpt_light) == 80, "Light std140 stride");
HPT_STATIC_ASSERT(offsetof(hpt_light, position) == 16, "Light.position");
HPT_STATIC_ASSERT(offsetof(hpt_light, u) == 32, "Light.u");
HPT_STATIC_ASSERT(offsetof(hpt_light, v) == 48, "Light.v");
HPT_STATIC_ASSERT(offsetof(hpt_light, radius) == 60, "Light.radius");
HPT_STATIC_ASSERT(offsetof(hpt_light, area) == 64, "Light.area");
HPT_STATIC_ASSERT(offsetof(hpt_light, type) == 68, "Light.type");
HPT_STATIC_ASSERT(sizeof(hpt_medium) == 32, "Medium size");
HPT_STATIC_ASSERT(sizeof(hpt_material) == 144, "Material size");
HPT_STATIC_ASSERT(offsetof(hpt_material, base_color) == 32, "Material.base_color");
HPT_STATIC_ASSERT(offsetof(hpt_material, emission) == 48, "Material.emission");
HPT_STATIC_ASSERT(offsetof(hpt_material, metallic) == 64, "Material.metallic");
HPT_STATIC_ASSERT(offsetof(hpt_material, clearcoat_tint) == 96, "Material.clearcoat_tint");
HPT_STATIC_ASSERT(offsetof(hpt_material, ior) == 112, "Material.ior");
HPT_STATIC_ASSERT(offsetof(hpt_material, ax) == 116, "Material.ax");
HPT_STATIC_ASSERT(offsetof(hpt_material, base_color_map_index) == 124, "Material.maps");
HPT_STATIC_ASSERT(offsetof(hpt_material, type) == 140, "Material.type");
HPT_STATIC_ASSERT(sizeof(hpt_vertex) == 44, "Vertex std430 size");
HPT_STATIC_ASSERT(sizeof(hpt_ray) == 32, "ray size");
HPT_STATIC_ASSERT(sizeof(hpt_hit) == 24, "hit size");

#endif /* HPT_H */
