/*
 * oracle.h — C entry points of the CPU oracle.
 *
 * TEST INFRASTRUCTURE.  The oracle is a CPU restatement of the reference's
 * GLSL path tracer (shaders/src/) plus brute-force ray queries.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker; the HIP library never links or calls it.
 *
 * It takes the same scene description as the product (include/hpt.h).
 */
#ifndef HPT_ORACLE_H
#define HPT_ORACLE_H

#include "../include/hpt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* Render settings: the MainUBO values (inc/global.glsl:11-30) a frame uses. */
typedef struct or_settings {
  uint32_t width, height, max_depth, rr_depth;
  uint32_t enable_tonemap, enable_aces, use_simple_aces;
  uint32_t features; /* HPT_FEATURE_* */
  uint32_t camera_index;
  float exposure_value;
} or_settings;

or_scene* or_scene_create(const hpt_scene_desc* desc);
void or_scene_destroy(or_scene* s);
int or_scene_set_env_map(or_scene* s, const float* rgb, uint32_t w, uint32_t h, float rotation);
void or_scene_set_env_sky(or_scene* s, const float ground[3], const float sky[3]);
void or_scene_set_env_intensity(or_scene* s, float intensity);
int or_scene_set_blue_noise(or_scene* s, const uint8_t* rgba8, uint32_t w, uint32_t h);
/* Material texture table (hpt_set_textures semantics); 0 on success. */
int or_scene_set_textures(or_scene* s, const hpt_texture_desc* textures, uint32_t count);
/* One look-up of texture t at (u, v) into rgba[4]; srgb selects the colour-map decode. */
float or_texture_sample(const or_scene* s, uint32_t t, float u, float v, int srgb, float* rgba);
uint32_t or_scene_triangle_count(const or_scene* s);

/* Ray queries. */
void or_trace_closest_brute(const or_scene* s, const hpt_ray* rays, uint32_t n, hpt_hit* hits);
void or_trace_closest(const or_scene* s, const hpt_ray* rays, uint32_t n, hpt_hit* hits);
void or_trace_any(const or_scene* s, const hpt_ray* rays, uint32_t n, uint32_t* occluded);
void or_trace_any_brute(const or_scene* s, const hpt_ray* rays, uint32_t n, uint32_t* occluded);
/* FP64 reference of the triangle test: per ray, the closest t over all
 * triangles computed in double (ids as in hpt_hit.instance/primitive), and the
 * gap to the second-closest distinct triangle. */
void or_trace_closest_fp64(const or_scene* s, const hpt_ray* rays, uint32_t n, double* t,
                           uint32_t* global_tri, double* gap);

/* Walks the product's 8-wide BVH (arrays from hpt_export_bvh) exactly as the
 * GPU kernel does and counts node visits and triangle tests per ray.
 * any_hit != 0 gives the shadow-ray walk (stops at the first hit). */
void or_trace_bvh8(const or_scene* s, const void* nodes, uint32_t node_count, const void* tris,
                   uint32_t tri_count, const hpt_ray* rays, uint32_t n, int any_hit,
                   hpt_hit* hits, uint32_t* node_visits, uint32_t* tri_tests);

/* The camera callable for one pixel (consumes the frame's blue-noise draw). */
void or_camera_ray(const or_scene* s, const or_settings* st, uint32_t frame, uint32_t x,
                   uint32_t y, hpt_ray* out);

/* One path per pixel of frame `frame` (pathtracer.rgen.glsl main() before the
 * accumulation): radiance, first-hit albedo and normal, rgb each, for the
 * pixel rectangle [x0,x1)x[y0,y1), row-major.  threads <= 0 uses all cores. */
int or_trace_paths(const or_scene* s, const or_settings* st, uint32_t frame, uint32_t x0,
                   uint32_t y0, uint32_t x1, uint32_t y1, float* radiance, float* albedo,
                   float* normal, int threads);

/* Frames [frame_begin, frame_begin+count) accumulated into the full-image
 * rgba32f running means accum/albedo/normal (pathtracer.rgen.glsl:259-278)
 * and the rgba8 final image (:280-304). */
int or_render(const or_scene* s, const or_settings* st, uint32_t frame_begin, uint32_t count,
              float* accum, float* albedo, float* normal, uint8_t* final_rgba8, int threads);

/* Tonemap + sRGB of one accumulated colour (pathtracer.rgen.glsl:280-304). */
void or_final_color(const or_settings* st, const float rgb[3], float out[3]);

/* ---- single functions, for pinning tests ---- */
void or_pcg4d(uint32_t v[4]);
float or_u32_to_unit(uint32_t u);
void or_rng_init(uint32_t x, uint32_t y, uint32_t frame, uint32_t s0[4], uint32_t s1[4]);
void or_blue_noise_texel(const or_scene* s, uint32_t tx, uint32_t ty, uint8_t out[4]);

/* Shading frame for BxDF calls: tangent, bitangent, N (= ffnormal), eta,
 * depth and any_non_specular_bounce. */
typedef struct or_frame {
  float tangent[3], bitangent[3], n[3];
  float eta;
  uint32_t depth;
  uint32_t any_non_specular_bounce;
} or_frame;
void or_bxdf_eval(const hpt_material* m, const or_frame* fr, const float V[3], const float L[3],
                  float f[3], float* pdf);
/* rng[4] is the s0 state, advanced in place. */
void or_bxdf_sample(const hpt_material* m, const or_frame* fr, const float V[3], uint32_t rng[4],
                    float L[3], float f[3], float* pdf, uint32_t* flags);

/* Light sampling for lights[index] seen from `p`; num_lights is the UBO count. */
void or_light_sample(const hpt_light* light, uint32_t num_lights, const float p[3], uint32_t rng[4],
                     float dir[3], float normal[3], float emission[3], float* dist, float* pdf);
/* lights.rint.glsl for one light: returns t (1e6 when missed), pdf. */
float or_light_intersect(const hpt_light* light, const float org[3], const float dir[3], float* pdf);

/* Environment helpers on the scene's map. */
double or_build_alias(const float* lum, uint32_t n, float* prob, uint32_t* alias);
void or_env_sample(const or_scene* s, uint32_t rng[4], float dir[3], float color[3], float* pdf);
void or_env_eval(const or_scene* s, const float dir[3], uint32_t depth, float bsdf_pdf, float rgb[3]);
float or_env_pdf_dir(const or_scene* s, const float dir[3]);
void or_env_alias(const or_scene* s, float* prob, uint32_t* alias, float* lum, float* total);

/* Sampling helpers. */
void or_sample_hg(const float V[3], float g, float r1, float r2, float out[3]);
float or_phase_hg(float cos_theta, float g);
void or_sample_ggx_vndf(const float wi[3], float ax, float ay, float r1, float r2, float out[3]);
void or_cosine_sample_hemisphere(float r1, float r2, float out[3]);
void or_sample_gtr1(float roughness, float r1, float r2, float out[3]);
void or_tonemap_aces_fitted(const float in[3], float out[3]);
void or_tonemap_aces(const float in[3], float out[3]);
void or_tonemap_reinhard(const float in[3], float limit, float out[3]);
float or_linear_2_srgb(float x);

#ifdef __cplusplus
}
#endif

#endif
