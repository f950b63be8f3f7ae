// shading.h — CPU restatement of the reference's shading GLSL, function by
// function.  TEST INFRASTRUCTURE (see glsl.h).  Every function cites the GLSL
// file:line it restates (paths under shaders/src/).
#pragma once

#include <cstdint>
#include <vector>

#include "glsl.h"

namespace oracle {

// ---------------------------------------------------------------------------
// inc/constants.glsl:8-44
// ---------------------------------------------------------------------------
constexpr uint32_t INVALID_INDEX = 0xFFFFFFFFu;
constexpr uint32_t RAY_FLAGS_HIT = 0x1;
constexpr uint32_t RAY_FLAGS_IS_EMITTER = 0x2;
constexpr uint32_t RAY_FLAGS_SPECULAR = 0x4;
constexpr uint32_t RAY_FLAGS_REFLECTION = 0x8;
constexpr uint32_t RAY_FLAGS_TRANSMISSION = 0x10;
constexpr int LIGHT_TYPE_POINT = 0, LIGHT_TYPE_DIRECTIONAL = 1, LIGHT_TYPE_SPOT = 2,
              LIGHT_TYPE_QUAD = 3, LIGHT_TYPE_SPHERE = 4;
constexpr uint32_t ENV_TYPE_SKY = 0, ENV_TYPE_MAP = 1;
constexpr uint32_t MEDIUM_NONE = 0, MEDIUM_ABSORB = 1, MEDIUM_SCATTER = 2, MEDIUM_EMISSIVE = 3;

// ---------------------------------------------------------------------------
// inc/types.glsl
// ---------------------------------------------------------------------------
struct Ray {  // common/common.glsl:18-21
  vec3 origin, direction;
};
struct Camera {  // types.glsl:2-11
  vec3 position, right, up, forward;
  float yfov, focal_distance_or_xmag, aperture_or_ymag;
  uint32_t type;
};
struct Light {  // types.glsl:14-39
  vec3 intensity, position, u, v;
  float radius, area;
  int type;
};
struct Medium {  // types.glsl:42-49
  vec3 color;
  float density, anisotropy;
  uint32_t type;
};
struct Material {  // types.glsl:52-83
  Medium medium;
  vec3 base_color;
  float opacity;
  vec3 emission;
  float anisotropic, metallic, roughness, subsurface, specular_tint;
  float sheen, sheen_tint, clearcoat, clearcoat_roughness;
  vec3 clearcoat_tint;
  float specular_transmission, ior, ax, ay;
  uint32_t base_color_map_index, normal_map_index, metallic_roughness_map_index,
      emission_map_index, type;
};
struct State {  // types.glsl:128-144
  vec3 first_hit_position, ffnormal, normal, tangent, bitangent;
  vec2 tex_coord;
  float hit_distance = 0, pdf = 0, eta = 0;
  uint32_t flags = 0, depth = 0, material_index = 0;
  Medium medium{};
};

// The values of MainUBO (inc/global.glsl:11-30) that shading reads.
struct Globals {
  vec3 ground_color, sky_color;
  vec2 resolution;
  uint32_t max_depth = 1, rr_depth = 0, frame_index = 0, camera_index = 0, env_type = 0;
  uint32_t env_map_width = 0, env_map_height = 0;
  float env_total_sum = 0, env_rotation = 0, env_intensity = 1, exposure_value = 0;
  bool enable_tonemap = false, enable_aces = false, use_simple_aces = false;
  uint32_t num_of_lights = 0;
};

// Environment map resources: the texels (linear RGB) and the luminance alias
// table that replaces g_env_map_dist (inc/global.glsl:7-8).
struct EnvMap {
  uint32_t width = 0, height = 0;
  std::vector<float> rgb;        // width*height*3
  std::vector<float> lum;        // luminance(texel), width*height
  std::vector<float> prob;       // alias acceptance probability
  std::vector<uint32_t> alias;   // alias index
};

// Blue-noise texture read by rand4blue (common/wltcRS.glsl:46-53); RGBA8.
struct BlueNoise {
  uint32_t width = 1024, height = 1024;
  std::vector<uint8_t> rgba;
};

// ---------------------------------------------------------------------------
// common/common.glsl
// ---------------------------------------------------------------------------
float linear_2_srgb(float x);                // :26-32
vec3 linear_2_srgb(vec3 x);                  // :37-39
float luminance(vec3 c);                     // :69-71
vec3 to_world(vec3 X, vec3 Y, vec3 Z, vec3 v);  // :79-81
vec3 to_local(vec3 X, vec3 Y, vec3 Z, vec3 v);  // :89-91
float max_(vec3 v);                          // :186

// ---------------------------------------------------------------------------
// common/wltcRS.glsl
// ---------------------------------------------------------------------------
struct RNGState {
  uvec4 s0, s1;
  int px, py;
};
RNGState rng_initialize(uint32_t x, uint32_t y, uint32_t frame);  // :9-19
void pcg4d(uvec4& v);                                               // :22-27
float rand(RNGState& s);                                            // :29-31
vec2 rand2(RNGState& s);                                            // :33-35
void shift2(RNGState& s, int& tx, int& ty);                         // :46-49
void rand4blue(RNGState& s, const BlueNoise& bn, float out[4]);     // :51-53
// float(u) / float(0xffffffffu): the divisor rounds to 2^32.
inline float u32_to_unit(uint32_t u) { return (float)u / 4294967296.0f; }

// ---------------------------------------------------------------------------
// common/sampling.glsl
// ---------------------------------------------------------------------------
float gtr1(float n_dot_h, float a);                                 // :8-18
vec3 sample_gtr1(float roughness, float r1, float r2);              // :25-44
vec3 sample_ggx_vndf(vec3 wi, float ax, float ay, float r1, float r2);  // :120-136
float gtr2_aniso(float n_dot_h, float h_dot_x, float h_dot_y, float ax, float ay);  // :146-154
float smith_g(float n_dot_v, float alpha_g);                        // :181-188
float smith_g_aniso(float n_dot_v, float v_dot_x, float v_dot_y, float ax, float ay);  // :193-202
float schlick_weight(float u);                                      // :207-211
float dielectric_fresnel(float cos_theta_i, float eta);             // :218-236
vec3 cosine_sample_hemisphere(float r1, float r2);                  // :252-260
vec3 uniform_sample_hemisphere(float r1, float r2);                 // :266-270
float power_heuristic(float a, float b);                            // :286-290
void onb(vec3 N, vec3& T, vec3& B);                                 // :296-300
vec3 sample_hg(vec3 V, float g, float r1, float r2);                // :308-327
float phase_hg(float cos_theta, float g);                           // :333-336

// ---------------------------------------------------------------------------
// common/intersection.glsl
// ---------------------------------------------------------------------------
float sphere_intersect(float rad, vec3 pos, const Ray& r);          // :6-33
float rect_intersect(vec3 pos, vec3 u, vec3 v, float plane[4], const Ray& r);  // :43-72

// ---------------------------------------------------------------------------
// common/tonemap.glsl
// ---------------------------------------------------------------------------
vec3 aces_fitted(vec3 color);  // :21-33
vec3 aces(vec3 c);             // :35-43
vec3 tonemap(vec3 c, float limit);  // :45-47

// ---------------------------------------------------------------------------
// BxDFs: inc/bxdf/{lambert,oren_nayar,disney}.glsl and the diffuse/disney
// callables (diffuse_bxdf_{eval,sample}.rcall.glsl, disney_bxdf_*.rcall.glsl).
// ---------------------------------------------------------------------------
vec3 lambert_eval(const State& state, const Material& mat, vec3 V, vec3 N, vec3 L, float& pdf);
vec3 oren_nayar_eval(const State& state, const Material& mat, vec3 V, vec3 N, vec3 L,
                     float& pdf);
vec3 disney_eval(bool any_non_specular_bounce, const State& state, const Material& mat, vec3 V,
                 vec3 N, vec3 L, float& pdf);
vec3 disney_sample(RNGState& rng, bool& any_non_specular_bounce, const State& state,
                   const Material& mat, vec3 V, vec3 N, vec3& L, float& pdf, uint32_t& flags);
// Callable dispatch CALLABLE_MATERIAL_BXDF_BEGIN + mat.type*2 (+1).
vec3 bxdf_eval(bool any_non_specular_bounce, const State& state, const Material& mat, vec3 V,
               vec3 N, vec3 L, float& pdf);
vec3 bxdf_sample(RNGState& rng, bool& any_non_specular_bounce, const State& state,
                 const Material& mat, vec3 V, vec3 N, vec3& L, float& pdf, uint32_t& flags);

// ---------------------------------------------------------------------------
// Lights: inc/light/*.glsl via the *_light_sample callables.
// ---------------------------------------------------------------------------
struct LightSample {
  vec3 normal, emission, direction;
  float dist = 0, pdf = 0;
};
void light_sample(RNGState& rng, const Globals& g, const State& state, const Light& light,
                  LightSample& out);
float quad_intersect(const Light& light, const Ray& ray, vec3& normal, float& pdf);   // quad.glsl:42-60
float sphere_light_intersect(const Light& light, const Ray& ray, float& pdf);          // sphere.glsl:50-66

// ---------------------------------------------------------------------------
// Environment: inc/env_map.glsl, env_map_eval.rcall.glsl, env_sky_eval.rcall.glsl.
// ---------------------------------------------------------------------------
// Alias table over texel luminance (Vose's method, double precision, stacks
// filled in increasing index order); returns the luminance total.
double build_env_alias(const float* lum, uint32_t n, float* prob, uint32_t* alias);
vec3 env_texture(const EnvMap& env, vec2 uv);  // textureLod(g_env_map, uv, 0)
vec3 sample_env_map(RNGState& rng, const Globals& g, const EnvMap& env, vec3& color, float& pdf);
float env_pdf(const Globals& g, const EnvMap& env, vec2 uv, float theta);
vec3 env_eval(const Globals& g, const EnvMap& env, uint32_t flags, uint32_t depth, float pdf,
              vec3 direction, bool use_scatter_flag, bool is_surface_scatter);

// ---------------------------------------------------------------------------
// Cameras: perspective_ray.rcall.glsl:12-64, orthographic_ray.rcall.glsl:12-45.
// ---------------------------------------------------------------------------
Ray camera_ray(RNGState& rng, const Globals& g, const Camera& cam, const BlueNoise& bn,
               uint32_t x, uint32_t y);

}  // namespace oracle
