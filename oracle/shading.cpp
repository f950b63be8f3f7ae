// shading.cpp — CPU restatement of the reference shading GLSL.
// TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load anything under oracle/.
// Built with -O2 -ffp-contract=off; every expression keeps the GLSL's order.
#include "shading.h"

#include <algorithm>
#include <cmath>

namespace oracle {

// ------------------------------------------------------------ common.glsl
float linear_2_srgb(float x) {  // common/common.glsl:26-32
  if (x <= 0.0031308f) {
    return 12.92f * x;
  } else {
    return 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
  }
}
vec3 linear_2_srgb(vec3 x) {  // :37-39
  return vec3(linear_2_srgb(x.x), linear_2_srgb(x.y), linear_2_srgb(x.z));
}
float luminance(vec3 c) {  // :69-71 (BT.709)
  return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z;
}
vec3 to_world(vec3 X, vec3 Y, vec3 Z, vec3 v) {  // :79-81
  return v.x * X + v.y * Y + v.z * Z;
}
vec3 to_local(vec3 X, vec3 Y, vec3 Z, vec3 v) {  // :89-91
  return vec3(dot(v, X), dot(v, Y), dot(v, Z));
}
// cos_theta(w) = w.z (:96) is used inline.
static float cos2_theta(vec3 w) { return w.z * w.z; }                   // :101
static float abs_cos_theta(vec3 w) { return std::fabs(w.z); }           // :106
static float sin2_theta(vec3 w) { return gmax(0.f, 1.f - cos2_theta(w)); }  // :111
static float sin_theta(vec3 w) { return std::sqrt(sin2_theta(w)); }     // :116
static float cos_phi(vec3 w) {                                          // :131-134
  const float sin_theta_ = sin_theta(w);
  return (sin_theta_ == 0.f) ? 1.f : clampf(w.x / sin_theta_, -1.f, 1.f);
}
static float sin_phi(vec3 w) {  // :139-142
  const float sin_theta_ = sin_theta(w);
  return (sin_theta_ == 0.f) ? 0.f : clampf(w.y / sin_theta_, -1.f, 1.f);
}
float max_(vec3 v) { return gmax(gmax(v.x, v.y), v.z); }  // :186

// ------------------------------------------------------------ wltcRS.glsl
RNGState rng_initialize(uint32_t x, uint32_t y, uint32_t frame) {  // :9-19
  RNGState s;
  s.px = (int)x;
  s.py = (int)y;
  s.s0 = uvec4{x, y, frame, x + y};
  s.s1 = uvec4{frame, frame * 15843u, frame * 31u + 4566u, frame * 2345u + 58585u};
  return s;
}
void pcg4d(uvec4& v) {  // :22-27
  v.x = v.x * 1664525u + 1013904223u;
  v.y = v.y * 1664525u + 1013904223u;
  v.z = v.z * 1664525u + 1013904223u;
  v.w = v.w * 1664525u + 1013904223u;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
  v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u; v.w ^= v.w >> 16u;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
}
float rand(RNGState& s) {  // :29-31
  pcg4d(s.s0);
  return u32_to_unit(s.s0.x);
}
vec2 rand2(RNGState& s) {  // :33-35
  pcg4d(s.s0);
  return vec2(u32_to_unit(s.s0.x), u32_to_unit(s.s0.y));
}
void shift2(RNGState& s, int& tx, int& ty) {  // :46-49
  pcg4d(s.s1);
  tx = (s.px + (int)(s.s1.x % 0x0fffffffu)) % 1024;
  ty = (s.py + (int)(s.s1.y % 0x0fffffffu)) % 1024;
}
void rand4blue(RNGState& s, const BlueNoise& bn, float out[4]) {  // :51-53
  int tx, ty;
  shift2(s, tx, ty);
  // texelFetch of an RGBA8 UNORM texture: byte / 255.
  const uint8_t* t = &bn.rgba[((size_t)ty * bn.width + (size_t)tx) * 4];
  for (int i = 0; i < 4; ++i) out[i] = (float)t[i] / 255.0f;
}

// ------------------------------------------------------------ sampling.glsl
float gtr1(float n_dot_h, float a) {  // :8-18
  if (a >= 1.0f) return INV_PI;
  const float a2 = a * a;
  const float t = 1.0f + (a2 - 1.0f) * n_dot_h * n_dot_h;
  return (a2 - 1.0f) / (PI * std::log(a2) * t);
}
vec3 sample_gtr1(float roughness, float r1, float r2) {  // :25-44
  const float a = gmax(0.001f, roughness);
  const float a2 = a * a;
  const float phi = r1 * TWO_PI;
  const float cos_theta_ = std::sqrt((1.0f - std::pow(a2, 1.0f - r2)) / (1.0f - a2));
  const float sin_theta_ = clampf(std::sqrt(1.0f - (cos_theta_ * cos_theta_)), 0.0f, 1.0f);
  const float sin_phi_ = std::sin(phi);
  const float cos_phi_ = std::cos(phi);
  (void)sin_phi_;
  // The reference returns sin_theta*sin_theta as y (sampling.glsl:43).
  return vec3(sin_theta_ * cos_phi_, sin_theta_ * sin_theta_, cos_theta_);
}
vec3 sample_ggx_vndf(vec3 wi, float ax, float ay, float r1, float r2) {  // :120-136
  const vec3 wiStd = normalize(vec3(wi.x * ax, wi.y * ay, wi.z));
  const float phi = (2.0f * r1 - 1.0f) * PI;
  const float z = std::fma((1.0f - r2), (1.0f + wiStd.z), -wiStd.z);
  const float sinTheta = std::sqrt(clampf(1.0f - z * z, 0.0f, 1.0f));
  const float x = sinTheta * std::cos(phi);
  const float y = sinTheta * std::sin(phi);
  const vec3 c = vec3(x, y, z);
  const vec3 wmStd = c + wiStd;
  const vec3 wm = normalize(vec3(wmStd.x * ax, wmStd.y * ay, wmStd.z));
  return wm;
}
float gtr2_aniso(float n_dot_h, float h_dot_x, float h_dot_y, float ax, float ay) {  // :146-154
  const float a = h_dot_x / ax;
  const float b = h_dot_y / ay;
  const float c = a * a + b * b + n_dot_h * n_dot_h;
  return 1.0f / (PI * ax * ay * c * c);
}
float smith_g(float n_dot_v, float alpha_g) {  // :181-188
  const float a = alpha_g * alpha_g;
  const float b = n_dot_v * n_dot_v;
  return (2.0f * n_dot_v) / (n_dot_v + std::sqrt(a + b - a * b));
}
float smith_g_aniso(float n_dot_v, float v_dot_x, float v_dot_y, float ax, float ay) {  // :193-202
  const float a = v_dot_x * ax;
  const float b = v_dot_y * ay;
  const float c = n_dot_v;
  return (2.0f * n_dot_v) / (n_dot_v + std::sqrt(a * a + b * b + c * c));
}
float schlick_weight(float u) {  // :207-211
  const float m = clampf(1.0f - u, 0.0f, 1.0f);
  const float m2 = m * m;
  return m2 * m2 * m;
}
float dielectric_fresnel(float cos_theta_i, float eta) {  // :218-236
  const float sin_theat_t_sq = eta * eta * (1.0f - cos_theta_i * cos_theta_i);
  if (sin_theat_t_sq > 1.0f) return 1.0f;
  const float cos_theta_t = std::sqrt(gmax(1.0f - sin_theat_t_sq, 0.0f));
  const float rs = (eta * cos_theta_t - cos_theta_i) / (eta * cos_theta_t + cos_theta_i);
  const float rp = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
  return 0.5f * (rs * rs + rp * rp);
}
vec3 cosine_sample_hemisphere(float r1, float r2) {  // :252-260
  vec3 dir;
  const float r = std::sqrt(r1);
  const float phi = TWO_PI * r2;
  dir.x = r * std::cos(phi);
  dir.y = r * std::sin(phi);
  dir.z = std::sqrt(gmax(0.0f, 1.0f - dir.x * dir.x - dir.y * dir.y));
  return dir;
}
vec3 uniform_sample_hemisphere(float r1, float r2) {  // :266-270
  const float r = std::sqrt(gmax(0.0f, 1.0f - r1 * r1));
  const float phi = TWO_PI * r2;
  return vec3(r * std::cos(phi), r * std::sin(phi), r1);
}
float power_heuristic(float a, float b) {  // :286-290
  a *= a;
  b *= b;
  return a / (a + b);
}
void onb(vec3 N, vec3& T, vec3& B) {  // :296-300
  vec3 up = std::fabs(N.z) < 0.9999999f ? vec3(0, 0, 1) : vec3(1, 0, 0);
  T = normalize(cross(up, N));
  B = cross(N, T);
}
vec3 sample_hg(vec3 V, float g, float r1, float r2) {  // :308-327
  float cos_theta_;
  if (std::fabs(g) < 0.001f)
    cos_theta_ = 1 - 2 * r2;
  else {
    const float sqr_term = (1 - g * g) / (1 + g - 2 * g * r2);
    cos_theta_ = -(1 + g * g - sqr_term * sqr_term) / (2 * g);
  }
  const float phi = r1 * TWO_PI;
  const float sin_theta_ = clampf(std::sqrt(1.0f - (cos_theta_ * cos_theta_)), 0.0f, 1.0f);
  const float sin_phi_ = std::sin(phi);
  const float cos_phi_ = std::cos(phi);
  vec3 v1, v2;
  onb(V, v1, v2);
  return sin_theta_ * cos_phi_ * v1 + sin_theta_ * sin_phi_ * v2 + cos_theta_ * V;
}
float phase_hg(float cos_theta_, float g) {  // :333-336
  const float denom = gmax(1 + g * g + 2 * g * cos_theta_, EPS);
  return INV_FOUR_PI * (1 - g * g) / (denom * std::sqrt(denom));
}

// ------------------------------------------------------------ intersection.glsl
float sphere_intersect(float rad, vec3 pos, const Ray& r) {  // :6-33
  const vec3 op = pos - r.origin;
  const float b = dot(op, r.direction);
  float det = b * b - dot(op, op) + rad * rad;
  if (det < 0.0f) return INF;
  det = std::sqrt(det);
  const float t1 = b - det;
  if (t1 > EPS) return t1;
  const float t2 = b + det;
  if (t2 > EPS) return t2;
  return INF;
}
float rect_intersect(vec3 pos, vec3 u, vec3 v, float plane[4], const Ray& r) {  // :43-72
  const vec3 n = vec3(plane[0], plane[1], plane[2]);
  const float dt = dot(r.direction, n);
  const float t = (plane[3] - dot(n, r.origin)) / dt;
  if (t > EPS) {
    const vec3 p = r.origin + r.direction * t;
    const vec3 vi = p - pos;
    const float a1 = dot(u, vi);
    if (a1 >= 0.0f && a1 <= 1.0f) {
      const float a2 = dot(v, vi);
      if (a2 >= 0.0f && a2 <= 1.0f) return t;
    }
  }
  return INF;
}

// ------------------------------------------------------------ tonemap.glsl
// mat3(c0, c1, c2) takes columns; `v * M` multiplies the row vector v, so
// result[j] = dot(v, column j): the listed vec3s act as rows of the transform
// applied to v (tonemap.glsl:2-13,22,27).
static vec3 row_mul(vec3 v, const float m[3][3]) {
  return vec3(dot(v, vec3(m[0][0], m[0][1], m[0][2])), dot(v, vec3(m[1][0], m[1][1], m[1][2])),
              dot(v, vec3(m[2][0], m[2][1], m[2][2])));
}
static const float ACES_INPUT_MATRIX[3][3] = {{0.59719f, 0.35458f, 0.04823f},
                                              {0.07600f, 0.90834f, 0.01566f},
                                              {0.02840f, 0.13383f, 0.83777f}};
static const float ACES_OUTPUT_MATRIX[3][3] = {{1.60475f, -0.53108f, -0.07367f},
                                               {-0.10208f, 1.10813f, -0.00605f},
                                               {-0.00327f, -0.07276f, 1.07602f}};
static vec3 rrt_odt_fit(vec3 v) {  // :15-19
  vec3 a = v * (v + 0.0245786f) - 0.000090537f;
  vec3 b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
  return a / b;
}
vec3 aces_fitted(vec3 color) {  // :21-33
  color = row_mul(color, ACES_INPUT_MATRIX);
  color = rrt_odt_fit(color);
  color = row_mul(color, ACES_OUTPUT_MATRIX);
  color = clamp3(color, 0.0f, 1.0f);
  return color;
}
vec3 aces(vec3 c) {  // :35-43
  const float a = 2.51f, b = 0.03f, y = 2.43f, d = 0.59f, e = 0.14f;
  return clamp3((c * (a * c + b)) / (c * (y * c + d) + e), 0.0f, 1.0f);
}
vec3 tonemap(vec3 c, float limit) {  // :45-47
  return c * 1.0f / (1.0f + luminance(c) / limit);
}

// ------------------------------------------------------------ lambert.glsl
vec3 lambert_eval(const State&, const Material& mat, vec3, vec3 N, vec3 L, float& pdf) {  // :37-42
  const float n_dot_l = dot(N, L);
  pdf = n_dot_l * INV_PI;
  return mat.base_color * n_dot_l * INV_PI;
}
static vec3 lambert_sample(RNGState& rng, bool& any_non_specular_bounce, const State& state,
                           const Material& mat, vec3 V, vec3 N, vec3& L, float& pdf,
                           uint32_t& flags) {  // :15-27
  const float r1 = rand(rng);
  const float r2 = rand(rng);
  L = cosine_sample_hemisphere(r1, r2);
  L = to_world(state.tangent, state.bitangent, N, L);
  any_non_specular_bounce = true;
  flags = RAY_FLAGS_REFLECTION;
  return lambert_eval(state, mat, V, N, L, pdf);
}

// ------------------------------------------------------------ oren_nayar.glsl
static vec3 oren_nayar_f(const State&, const Material& mat, vec3 wo, vec3 wi) {  // :7-51
  const float sin_theta_i = sin_theta(wi);
  const float sin_theta_o = sin_theta(wo);
  float max_cos = 0.0f;
  if (sin_theta_i > EPS && sin_theta_o > EPS) {
    const float sin_phi_i = sin_phi(wi), cos_phi_i = cos_phi(wi);
    const float sin_phi_o = sin_phi(wo), cos_phi_o = cos_phi(wo);
    const float d_cos = cos_phi_i * cos_phi_o + sin_phi_i * sin_phi_o;
    max_cos = gmax(0.0f, d_cos);
  }
  float sin_alpha, tan_beta;
  if (abs_cos_theta(wi) > abs_cos_theta(wo)) {
    sin_alpha = sin_theta_o;
    tan_beta = sin_theta_i / abs_cos_theta(wi);
  } else {
    sin_alpha = sin_theta_i;
    tan_beta = sin_theta_o / abs_cos_theta(wo);
  }
  const float A = mat.ax;
  const float B = mat.ay;
  return mat.base_color * INV_PI * (A + B * max_cos * sin_alpha * tan_beta);
}
vec3 oren_nayar_eval(const State& state, const Material& mat, vec3 V, vec3 N, vec3 L,
                     float& pdf) {  // :89-96
  L = to_local(state.tangent, state.bitangent, N, L);
  V = to_local(state.tangent, state.bitangent, N, V);
  pdf = L.z * INV_PI;
  return oren_nayar_f(state, mat, V, L) * L.z;
}
static vec3 oren_nayar_sample(RNGState& rng, bool& any_non_specular_bounce, const State& state,
                              const Material& mat, vec3 V, vec3 N, vec3& L, float& pdf,
                              uint32_t& flags) {  // :67-79
  const float r1 = rand(rng);
  const float r2 = rand(rng);
  L = cosine_sample_hemisphere(r1, r2);
  L = to_world(state.tangent, state.bitangent, N, L);
  any_non_specular_bounce = true;
  flags = RAY_FLAGS_REFLECTION;
  return oren_nayar_eval(state, mat, V, N, L, pdf);
}

// ------------------------------------------------------------ disney.glsl
static void tint_colors(vec3 base_color, float specular_tint, float sheen_tint, float eta,
                        float& F0, vec3& c_sheen, vec3& c_spec0) {  // :9-32
  const float lum = luminance(base_color);
  const vec3 ctint = lum > 0.0f ? base_color / lum : vec3(1.0f);
  F0 = (1.0f - eta) / (1.0f + eta);
  F0 *= F0;
  c_spec0 = F0 * mix3(vec3(1.0f), ctint, specular_tint);
  c_sheen = mix3(vec3(1.0f), ctint, sheen_tint);
}
static float diffuse_directional_albedo(float theta, float ax, float ay, float e0) {  // :34-37
  return mixf(e0 + (1.0f - e0) * std::pow(std::fabs(1.0f - theta), 5.0f), 0.04762f + 0.95238f * e0,
              1.0f - std::pow(std::fabs(1.0f - ax * ay), 5.0f));
}
static float diffuse_average_albedo(float ax, float ay, float e0) {  // :39-41
  return e0 + (-0.33263f * ax * ay - 0.072359f) * (1.0f - e0) * e0;
}
static float coupled_diffuse(float ax, float ay, float dot_wi_n, float dot_wo_n, float e0) {  // :43-48
  float Ewi = diffuse_directional_albedo(dot_wi_n, ax, ay, e0);
  float Ewo = diffuse_directional_albedo(dot_wo_n, ax, ay, e0);
  float Eavg = diffuse_average_albedo(ax, ay, e0);
  return (1.0f - Ewo) * (1.0f - Ewi) / (PI * (1.0f - Eavg));
}
static float sheen_directional_albedo(float cos_theta_, float alpha) {  // :50-54
  float c = 1.0f - cos_theta_;
  float c3 = c * c * c;
  return 0.65584461f * c3 + 1.0f / (4.16526551f + std::exp(-7.97291361f * std::sqrt(alpha) + 6.33516894f));
}
static vec3 eval_disney_diffuse(float ax, float ay, float F0, float specular_tint, vec3 base_color,
                                float roughness, float subsurface, float sheen, vec3 c_sheen,
                                vec3 V, vec3 L, vec3 H, float& pdf) {  // :71-113
  pdf = 0.0f;
  if (L.z <= 0.0f) return vec3(0.0f);
  const float L_dot_H = dot(L, H);
  const float retro_reflection = 2.0f * roughness * L_dot_H * L_dot_H;
  const float f_L = schlick_weight(L.z);
  const float f_V = schlick_weight(V.z);
  const float f_retro = retro_reflection * (f_L + f_V + f_L * f_V * (retro_reflection - 1.0f));
  const float f_d = (1.0f - 0.5f * f_L) * (1.0f - 0.5f * f_V);
  const float coupled = coupled_diffuse(ax, ay, std::fabs(V.z), std::fabs(L.z), F0 * specular_tint);
  const float f_ss90 = 0.5f * retro_reflection;
  const float f_ss = mixf(1.0f, f_ss90, f_L) * mixf(1.0f, f_ss90, f_V);
  const float ss = 1.25f * (f_ss * (1.0f / (L.z + V.z) - 0.5f) + 0.5f);
  const float f_H = schlick_weight(L_dot_H);
  const vec3 f_sheen = f_H * sheen * c_sheen;
  float alpha = gmax(roughness, 0.07f);
  alpha = alpha * alpha;
  const float Ewi = max_(c_sheen) * sheen_directional_albedo(std::fabs(V.z), alpha);
  const float Ewo = max_(c_sheen) * sheen_directional_albedo(std::fabs(L.z), alpha);
  const float sheen_weight = gmin(1.0f - Ewi, 1.0f - Ewo);
  pdf = L.z * INV_PI;
  return INV_PI * base_color * mixf(f_d * coupled * PI + f_retro, ss, subsurface) +
         f_sheen * sheen_weight;
}
static vec3 eval_microfacet_reflection(float ax, float ay, vec3 V, vec3 L, vec3 H, vec3 F,
                                       float& pdf) {  // :124-143
  pdf = 0.0f;
  if (L.z <= 0.0f) return vec3(0.0f);
  const float D = gtr2_aniso(H.z, H.x, H.y, ax, ay);
  const float G1 = smith_g_aniso(std::fabs(V.z), V.x, V.y, ax, ay);
  const float G2 = G1 * smith_g_aniso(std::fabs(L.z), L.x, L.y, ax, ay);
  pdf = G1 * D / (4.0f * V.z);
  return F * D * G2 / (4.0f * L.z * V.z);
}
static float specular_directional_albedo(float alpha, float cosTheta) {  // :145-149
  return 1.0f - 1.45940f * alpha * (-0.20276f + alpha * (2.77203f + (-2.61748f + 0.73343f * alpha) * alpha)) *
                     cosTheta *
                     (3.09507f + cosTheta * (-9.11368f + cosTheta * (15.88844f + cosTheta * (-13.70343f + 4.51786f * cosTheta))));
}
static float specular_average_albedo(float alpha) {  // :151-153
  return 1.0f + alpha * (-0.11304f + alpha * (-1.86947f + (2.22682f - 0.83397f * alpha) * alpha));
}
static vec3 specular_average_fresnel(vec3 f0, vec3 f90) {  // :155-157
  return 20.0f / 21.0f * f0 + 1.0f / 21.0f * f90;
}
static vec3 eval_microfacet_ms(float ax, float ay, vec3 V, vec3 L, vec3, vec3 F0) {  // :168-177
  float alpha = std::sqrt(ax * ay);
  float Ewi = specular_directional_albedo(alpha, std::fabs(V.z));
  float Ewo = specular_directional_albedo(alpha, std::fabs(L.z));
  float Eavg = specular_average_albedo(alpha);
  float ms = (1.0f - Ewo) * (1.0f - Ewi) / (PI * (1.0f - Eavg));
  vec3 Favg = specular_average_fresnel(F0, vec3(1.0f));
  vec3 f = (Favg * Favg * Eavg) / (1.0f - Favg * (1.0f - Eavg));
  return ms * f;
}
static vec3 eval_microfacet_refraction(float ax, float ay, vec3 base_color, float eta, vec3 V,
                                       vec3 L, vec3 H, vec3 F, float& pdf) {  // :190-219
  pdf = 0.0f;
  if (L.z >= 0.0f) return vec3(0.0f);
  const float l_dot_h = dot(L, H);
  const float v_dot_h = dot(V, H);
  const float D = gtr2_aniso(H.z, H.x, H.y, ax, ay);
  const float G1 = smith_g_aniso(std::fabs(V.z), V.x, V.y, ax, ay);
  const float G2 = G1 * smith_g_aniso(std::fabs(L.z), L.x, L.y, ax, ay);
  float denom = l_dot_h + v_dot_h * eta;
  denom *= denom;
  const float eta2 = eta * eta;
  const float jacobian = std::fabs(l_dot_h) / denom;
  pdf = G1 * gmax(0.0f, v_dot_h) * D * jacobian / V.z;
  return pow3(base_color, vec3(0.5f)) * (1.0f - F) * D * G2 * std::fabs(v_dot_h) * jacobian * eta2 /
         std::fabs(L.z * V.z);
}
static vec3 eval_clearcoat(float clearcoat_roughness, vec3 V, vec3 L, vec3 H, float& pdf) {  // :229-252
  pdf = 0.0f;
  if (L.z <= 0.0f) return vec3(0.0f);
  const float v_dot_h = dot(V, H);
  const float F = mixf(0.04f, 1.0f, schlick_weight(v_dot_h));
  const float D = gtr1(H.z, clearcoat_roughness);
  const float G = smith_g(L.z, 0.25f) * smith_g(V.z, 0.25f);
  const float jacobian = 1.0f / (4.0f * v_dot_h);
  pdf = D * H.z * jacobian;
  return vec3(F) * D * G;
}
vec3 disney_eval(bool any_non_specular_bounce, const State& state, const Material& mat, vec3 V,
                 vec3 N, vec3 L, float& pdf) {  // :264-390
  float ax = mat.ax;
  float ay = mat.ay;
  // ROUGHNESS_REGULARIZE_EVAL is defined (make_shaders.yaml:4).
  if (state.depth > 0 && any_non_specular_bounce) {
    if (ax < 0.3f) ax = clampf(2 * ax, 0.1f, 0.3f);
    if (ay < 0.3f) ay = clampf(2 * ay, 0.1f, 0.3f);
  }
  pdf = 0.0f;
  vec3 f = vec3(0.0f);
  V = to_local(state.tangent, state.bitangent, N, V);
  L = to_local(state.tangent, state.bitangent, N, L);
  vec3 H;
  if (L.z > 0.0f)
    H = normalize(L + V);
  else
    H = normalize(L + V * state.eta);
  if (H.z < 0.0f) H = -H;
  vec3 c_sheen, c_spec0;
  float F0;
  tint_colors(mat.base_color, mat.specular_tint, mat.sheen_tint, state.eta, F0, c_sheen, c_spec0);
  const float dielectric_wt = (1.0f - mat.metallic) * (1.0f - mat.specular_transmission);
  const float metal_wt = mat.metallic;
  const float glass_wt = (1.0f - mat.metallic) * mat.specular_transmission;
  const float schlick_wt = schlick_weight(V.z);
  float diffuse_pr = dielectric_wt * luminance(mat.base_color);
  float dielectric_pr = dielectric_wt * luminance(mix3(c_spec0, vec3(1.0f), schlick_wt));
  float metal_pr = metal_wt * luminance(mix3(mat.base_color, vec3(1.0f), schlick_wt));
  float glass_pr = glass_wt;
  float clearcoat_pr = 0.25f * mat.clearcoat;
  const float inv_total_wt = 1.0f / (diffuse_pr + dielectric_pr + metal_pr + glass_pr + clearcoat_pr);
  diffuse_pr *= inv_total_wt;
  dielectric_pr *= inv_total_wt;
  metal_pr *= inv_total_wt;
  glass_pr *= inv_total_wt;
  clearcoat_pr *= inv_total_wt;
  const bool reflect_ = L.z * V.z > 0;
  float tmp_pdf = 0.0f;
  const float v_dot_h = std::fabs(dot(V, H));
  if (diffuse_pr > 0.0f && reflect_) {
    f += eval_disney_diffuse(ax, ay, F0, mat.specular_tint, mat.base_color, mat.roughness,
                             mat.subsurface, mat.sheen, c_sheen, V, L, H, tmp_pdf) *
         dielectric_wt;
    pdf += tmp_pdf * diffuse_pr;
  }
  if (dielectric_pr > 0.0f && reflect_) {
    const float F = (dielectric_fresnel(v_dot_h, state.eta) - F0) / (1.0f - F0);
    f += eval_microfacet_reflection(ax, ay, V, L, H, mix3(c_spec0, vec3(1.0f), F), tmp_pdf) * dielectric_wt;
    f += eval_microfacet_ms(ax, ay, V, L, H, c_spec0) * dielectric_wt;
    pdf += tmp_pdf * dielectric_pr;
  }
  if (metal_pr > 0.0f && reflect_) {
    const vec3 F = mix3(mat.base_color, vec3(1.0f), schlick_weight(v_dot_h));
    f += eval_microfacet_reflection(ax, ay, V, L, H, F, tmp_pdf) * metal_wt;
    f += eval_microfacet_ms(ax, ay, V, L, H, mat.base_color) * metal_wt;
    pdf += tmp_pdf * metal_pr;
  }
  if (glass_pr > 0.0f) {
    const float F = dielectric_fresnel(v_dot_h, state.eta);
    if (reflect_) {
      f += eval_microfacet_reflection(ax, ay, V, L, H, vec3(F), tmp_pdf) * glass_wt;
      f += eval_microfacet_ms(ax, ay, V, L, H, c_spec0) * glass_wt;
      pdf += tmp_pdf * glass_pr * F;
    } else {
      f += eval_microfacet_refraction(ax, ay, mat.base_color, state.eta, V, L, H, vec3(F), tmp_pdf) * glass_wt;
      pdf += tmp_pdf * glass_pr * (1.0f - F);
    }
  }
  if (clearcoat_pr > 0.0f && reflect_) {
    f += eval_clearcoat(mat.clearcoat_roughness, V, L, H, tmp_pdf) * mat.clearcoat_tint * 0.25f * mat.clearcoat;
    pdf += tmp_pdf * clearcoat_pr;
  }
  return f * std::fabs(L.z);
}
vec3 disney_sample(RNGState& rng, bool& any_non_specular_bounce, const State& state,
                   const Material& mat, vec3 V, vec3 N, vec3& L, float& pdf,
                   uint32_t& flags) {  // :405-531
  float ax = mat.ax;
  float ay = mat.ay;
  // ROUGHNESS_REGULARIZE_SAMPLE is not defined (make_shaders.yaml:5).
  any_non_specular_bounce = true;
  pdf = 0.0f;
  const float r1 = rand(rng);
  const float r2 = rand(rng);
  V = to_local(state.tangent, state.bitangent, N, V);
  vec3 c_sheen, c_spec0;
  float F0;
  tint_colors(mat.base_color, mat.specular_tint, mat.sheen_tint, state.eta, F0, c_sheen, c_spec0);
  const float dielectric_wt = (1.0f - mat.metallic) * (1.0f - mat.specular_transmission);
  const float metal_wt = mat.metallic;
  const float glass_wt = (1.0f - mat.metallic) * mat.specular_transmission;
  const float schlick_wt = schlick_weight(V.z);
  float diffuse_pr = dielectric_wt * luminance(mat.base_color);
  float dielectric_pr = dielectric_wt * luminance(mix3(c_spec0, vec3(1.0f), schlick_wt));
  float metal_pr = metal_wt * luminance(mix3(mat.base_color, vec3(1.0f), schlick_wt));
  float glass_pr = glass_wt;
  float clearcoat_pr = 0.25f * mat.clearcoat;
  const float inv_total_wt = 1.0f / (diffuse_pr + dielectric_pr + metal_pr + glass_pr + clearcoat_pr);
  diffuse_pr *= inv_total_wt;
  dielectric_pr *= inv_total_wt;
  metal_pr *= inv_total_wt;
  glass_pr *= inv_total_wt;
  clearcoat_pr *= inv_total_wt;
  float cdf[5];
  cdf[0] = diffuse_pr;
  cdf[1] = cdf[0] + dielectric_pr;
  cdf[2] = cdf[1] + metal_pr;
  cdf[3] = cdf[2] + glass_pr;
  cdf[4] = cdf[3] + clearcoat_pr;
  float r3 = rand(rng);
  if (r3 < cdf[0]) {
    L = cosine_sample_hemisphere(r1, r2);
    flags = RAY_FLAGS_REFLECTION;
  } else if (r3 < cdf[2]) {
    vec3 H = sample_ggx_vndf(V, ax, ay, r1, r2);
    if (H.z < 0.0f) H = -H;
    L = normalize(reflect(-V, H));
    flags = RAY_FLAGS_REFLECTION;
  } else if (r3 < cdf[3]) {
    vec3 H = sample_ggx_vndf(V, ax, ay, r1, r2);
    const float F = dielectric_fresnel(std::fabs(dot(V, H)), state.eta);
    if (H.z < 0.0f) H = -H;
    r3 = (r3 - cdf[2]) / (cdf[3] - cdf[2]);
    if (r3 < F) {
      L = normalize(reflect(-V, H));
      flags = RAY_FLAGS_REFLECTION;
    } else {
      L = normalize(refract(-V, H, state.eta));
      flags = RAY_FLAGS_TRANSMISSION;
    }
  } else {
    vec3 H = sample_gtr1(mat.clearcoat_roughness, r1, r2);
    if (H.z < 0.0f) H = -H;
    L = normalize(reflect(-V, H));
    flags = RAY_FLAGS_REFLECTION;
  }
  L = to_world(state.tangent, state.bitangent, N, L);
  V = to_world(state.tangent, state.bitangent, N, V);
  return disney_eval(any_non_specular_bounce, state, mat, V, N, L, pdf);
}

// Callables 11..14 (inc/constants.glsl:39; src/bin/hala-pathtracer.rs:170-186).
vec3 bxdf_eval(bool anb, const State& state, const Material& mat, vec3 V, vec3 N, vec3 L,
               float& pdf) {
  if (mat.type == 1) return disney_eval(anb, state, mat, V, N, L, pdf);  // disney_bxdf_eval.rcall.glsl:14-22
  // diffuse_bxdf_eval.rcall.glsl:15-21
  if (mat.roughness > EPS) return oren_nayar_eval(state, mat, V, N, L, pdf);
  return lambert_eval(state, mat, V, N, L, pdf);
}
vec3 bxdf_sample(RNGState& rng, bool& anb, const State& state, const Material& mat, vec3 V,
                 vec3 N, vec3& L, float& pdf, uint32_t& flags) {
  if (mat.type == 1) return disney_sample(rng, anb, state, mat, V, N, L, pdf, flags);  // disney_bxdf_sample.rcall.glsl:14-24
  // diffuse_bxdf_sample.rcall.glsl:15-38
  if (mat.roughness > EPS) return oren_nayar_sample(rng, anb, state, mat, V, N, L, pdf, flags);
  return lambert_sample(rng, anb, state, mat, V, N, L, pdf, flags);
}

// ------------------------------------------------------------ lights
static vec3 point_sample(const Globals& g, const State& state, const Light& light, LightSample& o) {  // point.glsl:10-20
  vec3 light_dir = light.position - state.first_hit_position;
  const float dist_sq = dot(light_dir, light_dir);
  o.dist = std::sqrt(dist_sq);
  light_dir = light_dir / o.dist;
  o.normal = -light_dir;
  o.emission = light.intensity * (float)g.num_of_lights / dist_sq;
  o.pdf = 1.0f;
  return light_dir;
}
static vec3 distant_sample(RNGState& rng, const Globals& g, const Light& light, LightSample& o) {  // distant.glsl:10-34
  vec3 direction = -light.u;
  const float cos_alpha = light.v.x;
  if (cos_alpha > EPS) {
    const vec3 N = direction;
    vec3 T, B;
    onb(N, T, B);
    const float r = rand(rng);
    const float phi = TWO_PI * rand(rng);
    const float z = (1 - r * (1 - cos_alpha));
    const float radius = std::sqrt(1 - z * z);
    const float x = radius * std::cos(phi);
    const float y = radius * std::sin(phi);
    direction = to_world(T, B, N, vec3(x, y, z));
  }
  o.normal = light.u;
  o.emission = light.intensity * (float)g.num_of_lights;
  o.dist = WORLD_SIZE;
  o.pdf = 1.0f;
  return direction;
}
static vec3 spot_sample(const Globals& g, const State& state, const Light& light, LightSample& o) {  // spot.glsl:10-20
  vec3 light_dir = light.position - state.first_hit_position;
  const float dist_sq = dot(light_dir, light_dir);
  o.dist = std::sqrt(dist_sq);
  light_dir = light_dir / o.dist;
  o.normal = -light_dir;
  o.emission = smoothstep(light.v.y, light.v.x, dot(-light_dir, light.u)) * light.intensity *
               (float)g.num_of_lights / dist_sq;
  o.pdf = 1.0f;
  return light_dir;
}
static vec3 quad_sample(RNGState& rng, const Globals& g, const State& state, const Light& light, LightSample& o) {  // quad.glsl:10-34
  const float r1 = rand(rng);
  const float r2 = rand(rng);
  const vec3 light_surface_pos = light.position + light.u * r1 + light.v * r2;
  vec3 light_dir = light_surface_pos - state.first_hit_position;
  o.normal = normalize(cross(light.v, light.u));
  const float cos_theta_ = dot(-light_dir, o.normal);
  if (cos_theta_ < 0.0f) {
    o.emission = vec3(0.0f);
    o.dist = 0.0f;
    o.pdf = 0.0f;
    return vec3(0.0f);
  }
  const float dist_sq = dot(light_dir, light_dir);
  o.dist = std::sqrt(dist_sq);
  light_dir /= o.dist;
  o.emission = light.intensity * (float)g.num_of_lights;
  o.pdf = dist_sq / (light.area * std::fabs(cos_theta_));
  return light_dir;
}
static vec3 sphere_sample(RNGState& rng, const Globals& g, const State& state, const Light& light, LightSample& o) {  // sphere.glsl:10-42
  const float r1 = rand(rng);
  const float r2 = rand(rng);
  vec3 sphere_center_to_surface = state.first_hit_position - light.position;
  const float dist_to_sphere_center = length(sphere_center_to_surface);
  sphere_center_to_surface /= dist_to_sphere_center;
  if (dist_to_sphere_center < light.radius) {
    o.normal = sphere_center_to_surface;
    o.emission = vec3(0.0f);
    o.dist = dist_to_sphere_center;
    o.pdf = 0.0f;
    return vec3(0.0f);
  }
  vec3 sampled_dir = uniform_sample_hemisphere(r1, r2);
  vec3 T, B;
  onb(sphere_center_to_surface, T, B);
  sampled_dir = T * sampled_dir.x + B * sampled_dir.y + sphere_center_to_surface * sampled_dir.z;
  vec3 light_surface_pos = light.position + sampled_dir * light.radius;
  vec3 light_dir = light_surface_pos - state.first_hit_position;
  const float dist_sq = dot(light_dir, light_dir);
  o.dist = std::sqrt(dist_sq);
  light_dir = light_dir / o.dist;
  o.normal = normalize(light_surface_pos - light.position);
  o.emission = light.intensity * (float)g.num_of_lights;
  o.pdf = dist_sq / (light.area * 0.5f * std::fabs(dot(o.normal, light_dir)));
  return light_dir;
}
void light_sample(RNGState& rng, const Globals& g, const State& state, const Light& light,
                  LightSample& o) {
  // Callable CALLABLE_LIGHT_BEGIN + light.type (inc/lighting.glsl:163).
  o = LightSample{};
  switch (light.type) {
    case LIGHT_TYPE_POINT: o.direction = point_sample(g, state, light, o); break;
    case LIGHT_TYPE_DIRECTIONAL: o.direction = distant_sample(rng, g, light, o); break;
    case LIGHT_TYPE_SPOT: o.direction = spot_sample(g, state, light, o); break;
    case LIGHT_TYPE_QUAD: o.direction = quad_sample(rng, g, state, light, o); break;
    case LIGHT_TYPE_SPHERE: o.direction = sphere_sample(rng, g, state, light, o); break;
    default: break;
  }
}
float quad_intersect(const Light& light, const Ray& ray, vec3& normal, float& pdf) {  // quad.glsl:42-60
  normal = normalize(cross(light.v, light.u));
  const float cos_theta_ = dot(-ray.direction, normal);
  if (cos_theta_ < 0.0f) return INF;
  float plane[4] = {normal.x, normal.y, normal.z, dot(normal, light.position)};
  const vec3 u = light.u / dot(light.u, light.u);
  const vec3 v = light.v / dot(light.v, light.v);
  const float t = rect_intersect(light.position, u, v, plane, ray);
  if (t < 0.0f) return INF;
  pdf = (t * t) / (light.area * cos_theta_);
  return t;
}
float sphere_light_intersect(const Light& light, const Ray& ray, float& pdf) {  // sphere.glsl:50-66
  const float dist = length(light.position - ray.origin);
  if (dist < light.radius) return INF;
  const float t = sphere_intersect(light.radius, light.position, ray);
  if (t < 0.0f) return INF;
  const vec3 hit_pt = ray.origin + t * ray.direction;
  const float cos_theta_ = dot(-ray.direction, normalize(hit_pt - light.position));
  pdf = (t * t) / (light.area * cos_theta_ * 0.5f);
  return t;
}

// ------------------------------------------------------------ environment
double build_env_alias(const float* lum, uint32_t n, float* prob, uint32_t* alias) {
  double total = 0.0;
  for (uint32_t i = 0; i < n; ++i) total += (double)lum[i];
  if (!(total > 0.0)) {
    for (uint32_t i = 0; i < n; ++i) { prob[i] = 1.0f; alias[i] = i; }
    return total;
  }
  std::vector<double> scaled(n);
  std::vector<uint32_t> small_, large_;
  small_.reserve(n);
  large_.reserve(n);
  for (uint32_t i = 0; i < n; ++i) {
    scaled[i] = (double)lum[i] * (double)n / total;
    if (scaled[i] < 1.0) small_.push_back(i); else large_.push_back(i);
  }
  while (!small_.empty() && !large_.empty()) {
    uint32_t s = small_.back(); small_.pop_back();
    uint32_t l = large_.back(); large_.pop_back();
    prob[s] = (float)scaled[s];
    alias[s] = l;
    scaled[l] = (scaled[l] + scaled[s]) - 1.0;
    if (scaled[l] < 1.0) small_.push_back(l); else large_.push_back(l);
  }
  for (uint32_t i : large_) { prob[i] = 1.0f; alias[i] = i; }
  for (uint32_t i : small_) { prob[i] = 1.0f; alias[i] = i; }
  return total;
}

// Bilinear lookup with REPEAT in u and CLAMP_TO_EDGE in v, texel centres at
// (i+0.5)/W: the sampler state of g_env_map is set up by the renderer
// (unvendored), this is the filter this project fixes for it.
vec3 env_texture(const EnvMap& env, vec2 uv) {
  const float W = (float)env.width, H = (float)env.height;
  const float x = uv.x * W - 0.5f;
  const float y = uv.y * H - 0.5f;
  const float fx0 = std::floor(x), fy0 = std::floor(y);
  const float fx = x - fx0, fy = y - fy0;
  long i0 = (long)fx0, j0 = (long)fy0;
  long w = (long)env.width, h = (long)env.height;
  long i0w = ((i0 % w) + w) % w, i1w = (((i0 + 1) % w) + w) % w;
  long j0c = std::min(std::max(j0, 0L), h - 1), j1c = std::min(std::max(j0 + 1, 0L), h - 1);
  auto tx = [&](long i, long j) {
    const float* p = &env.rgb[((size_t)j * env.width + (size_t)i) * 3];
    return vec3(p[0], p[1], p[2]);
  };
  vec3 top = mix3(tx(i0w, j0c), tx(i1w, j0c), fx);
  vec3 bot = mix3(tx(i0w, j1c), tx(i1w, j1c), fx);
  return mix3(top, bot, fy);
}

// Texel of the alias table that holds uv (REPEAT in u, CLAMP in v).
static uint32_t env_texel_index(const EnvMap& env, vec2 uv) {
  float fu = uv.x - std::floor(uv.x);
  long i = (long)(fu * (float)env.width);
  long j = (long)std::floor(uv.y * (float)env.height);
  i = std::min(std::max(i, 0L), (long)env.width - 1);
  j = std::min(std::max(j, 0L), (long)env.height - 1);
  return (uint32_t)(j * (long)env.width + i);
}

// inc/env_map.glsl:6-37 with the two distribution-texture lookups replaced by
// one alias-table draw.  It consumes one pcg4d step of s0, as rand2 does:
// x picks the texel, y decides texel/alias, z and w place the sample inside
// the texel.  pdf has the reference's form lum/total * W*H/(2*pi^2*sin(theta))
// with lum the drawn texel's luminance (the exact density of this sampler).
vec3 sample_env_map(RNGState& rng, const Globals& g, const EnvMap& env, vec3& color, float& pdf) {
  pcg4d(rng.s0);
  const uint32_t n = env.width * env.height;
  uint32_t idx = (uint32_t)(((uint64_t)rng.s0.x * (uint64_t)n) >> 32);
  if (!(u32_to_unit(rng.s0.y) < env.prob[idx])) idx = env.alias[idx];
  const uint32_t ti = idx % env.width, tj = idx / env.width;
  const float u = ((float)ti + u32_to_unit(rng.s0.z)) / (float)env.width;
  const float v = ((float)tj + u32_to_unit(rng.s0.w)) / (float)env.height;
  color = env_texture(env, vec2(u, v)) * g.env_intensity;
  pdf = env.lum[idx] / g.env_total_sum;
  const float phi = (u + g.env_rotation) * TWO_PI;
  const float theta = v * PI;
  const float sin_theta_ = std::sin(theta);
  if (sin_theta_ <= EPS) pdf = 0.0f;
  const vec3 direction = normalize(vec3(-sin_theta_ * std::cos(phi), sin_theta_ * std::sin(phi), std::cos(theta)));
  pdf = (pdf * (float)g.env_map_width * (float)g.env_map_height) / (TWO_PI * PI * sin_theta_);
  return direction;
}

float env_pdf(const Globals& g, const EnvMap& env, vec2 uv, float theta) {  // inc/env_map.glsl:43-55
  const float sin_theta_ = std::sin(theta);
  if (sin_theta_ > EPS) {
    const float pdf = env.lum[env_texel_index(env, uv)] / g.env_total_sum;
    if (pdf > EPS)
      return (pdf * (float)g.env_map_width * (float)g.env_map_height) / (TWO_PI * PI * sin_theta_);
    else
      return 0.0f;
  } else {
    return 0.0f;
  }
}

vec3 env_eval(const Globals& g, const EnvMap& env, uint32_t flags, uint32_t depth, float pdf,
              vec3 dir, bool use_scatter_flag, bool is_surface_scatter) {
  if (g.env_type == ENV_TYPE_SKY) {  // env_sky_eval.rcall.glsl:12-15
    const float a = gmax(dir.z, 0.0f);
    return mix3(g.ground_color, g.sky_color, a) * g.env_intensity;
  }
  // env_map_eval.rcall.glsl:14-35.  acos is undefined above 1 in GLSL; a
  // normalised direction can carry z = 1 + 2^-23, so z is clamped to [-1, 1]
  // (otherwise the NaN would poison the running mean of the pixel for good).
  const float theta = std::acos(gmin(gmax(dir.z, -1.0f), 1.0f));
  const vec2 uv = vec2(-(PI + std::atan2(dir.y, dir.x)) * INV_TWO_PI - g.env_rotation, theta * INV_PI);
  float mis_weight = 1.0f;
  if (depth > 0 && (flags & RAY_FLAGS_SPECULAR) == 0) {
    const float p = env_pdf(g, env, uv, theta);
    mis_weight = power_heuristic(pdf, p);
  }
  if (use_scatter_flag && !is_surface_scatter) mis_weight = 1.0f;
  return mis_weight * env_texture(env, uv) * g.env_intensity;
}

// ------------------------------------------------------------ cameras
Ray camera_ray(RNGState& rng, const Globals& g, const Camera& camera, const BlueNoise& bn,
               uint32_t x, uint32_t y) {
  Ray ray;
  float r[4];
  rand4blue(rng, bn, r);
  float r1 = 2.0f * r[0];
  float r2 = 2.0f * r[1];
  vec2 jitter;
  jitter.x = r1 < 1.0f ? std::sqrt(r1) - 1.0f : 1.0f - std::sqrt(2.0f - r1);
  jitter.y = r2 < 1.0f ? std::sqrt(r2) - 1.0f : 1.0f - std::sqrt(2.0f - r2);
  jitter /= (g.resolution * 0.5f);
  vec2 screen_uv = vec2((float)x, g.resolution.y - (float)y) / g.resolution;
  vec2 d = (2.0f * screen_uv - 1.0f) + jitter;
  if (camera.type == 1) {  // orthographic_ray.rcall.glsl:12-45
    const float xmag = camera.focal_distance_or_xmag;
    const float ymag = camera.aperture_or_ymag;
    ray.direction = normalize(camera.forward);
    ray.origin = camera.position + camera.right * (d.x * xmag) + camera.up * (d.y * ymag);
    return ray;
  }
  // perspective_ray.rcall.glsl:12-64
  const float focal_distance = camera.focal_distance_or_xmag;
  const float aperture = camera.aperture_or_ymag;
  float scale = std::tan(camera.yfov * 0.5f);
  d.x *= scale * g.resolution.x / g.resolution.y;
  d.y *= scale;
  vec3 ray_dir = normalize(camera.right * d.x + camera.up * d.y + camera.forward);
  vec3 final_ray_dir = ray_dir;
  vec3 final_ray_origin = camera.position;
  if (aperture > 0.0f) {
    vec3 focal_point = focal_distance * ray_dir;
    float cam_r1 = r[2] * TWO_PI;
    float cam_r2 = r[3] * aperture;
    vec3 rnd_aperture_pos = (std::cos(cam_r1) * camera.right + std::sin(cam_r1) * camera.up) * std::sqrt(cam_r2);
    final_ray_dir = normalize(focal_point - rnd_aperture_pos);
    final_ray_origin += rnd_aperture_pos;
  }
  ray.direction = final_ray_dir;
  ray.origin = final_ray_origin;
  return ray;
}

}  // namespace oracle
