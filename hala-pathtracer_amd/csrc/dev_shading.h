// dev_shading.h — the reference's shading GLSL as gfx950 device code: RNG,
// cameras, BxDFs, light sampling and the environment.  Each function cites
// the GLSL (paths under shaders/src/) whose arithmetic it keeps.
#pragma once

#include "common.h"

namespace hpt {

// ------------------------------------------------------------------- RNG
// common/wltcRS.glsl:22-27.
__device__ __forceinline__ void pcg4d(uint4& v) {
  v.x = v.x * 1664525u + 1013904223u;
  v.y = v.y * 1664525u + 1013904223u;
  v.z = v.z * 1664525u + 1013904223u;
  v.w = v.w * 1664525u + 1013904223u;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
  v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u; v.w ^= v.w >> 16u;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
}
// float(u) / float(0xffffffffu); the divisor rounds to 2^32 (wltcRS.glsl:30).
__device__ __forceinline__ float u2f(uint32_t u) { return (float)u * 2.3283064365386963e-10f; }
__device__ __forceinline__ float rnd(uint4& s) {  // rand(), :29-31
  pcg4d(s);
  return u2f(s.x);
}

// ------------------------------------------------------------------ common
__device__ __forceinline__ float luminance(f3 c) {  // common/common.glsl:69-71
  return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z;
}
__device__ __forceinline__ f3 to_world(f3 X, f3 Y, f3 Z, f3 v) { return v.x * X + v.y * Y + v.z * Z; }
__device__ __forceinline__ f3 to_local(f3 X, f3 Y, f3 Z, f3 v) { return mk3(dot(v, X), dot(v, Y), dot(v, Z)); }
__device__ __forceinline__ float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ f3 mix3(f3 x, f3 y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float max3(f3 v) { return fmaxf(fmaxf(v.x, v.y), v.z); }
__device__ __forceinline__ f3 reflect3(f3 I, f3 N) { return I - 2.0f * dot(N, I) * N; }
__device__ __forceinline__ f3 refract3(f3 I, f3 N, float eta) {
  const float d = dot(N, I);
  const float k = 1.0f - eta * eta * (1.0f - d * d);
  if (k < 0.0f) return mk3(0.0f);
  return eta * I - (eta * d + sqrtf(k)) * N;
}
__device__ __forceinline__ float sin_theta_l(f3 w) { return sqrtf(fmaxf(0.f, 1.f - w.z * w.z)); }  // common.glsl:111-116
__device__ __forceinline__ float cos_phi_l(f3 w) {  // :131-134
  const float s = sin_theta_l(w);
  return (s == 0.f) ? 1.f : clampf(w.x / s, -1.f, 1.f);
}
__device__ __forceinline__ float sin_phi_l(f3 w) {  // :139-142
  const float s = sin_theta_l(w);
  return (s == 0.f) ? 0.f : clampf(w.y / s, -1.f, 1.f);
}

// ---------------------------------------------------------------- sampling
__device__ __forceinline__ float gtr1(float n_dot_h, float a) {  // common/sampling.glsl:8-18
  if (a >= 1.0f) return INV_PI;
  const float a2 = a * a;
  const float t = 1.0f + (a2 - 1.0f) * n_dot_h * n_dot_h;
  return (a2 - 1.0f) / (PI * logf(a2) * t);
}
__device__ __forceinline__ f3 sample_gtr1(float roughness, float r1, float r2) {  // :25-44
  const float a = fmaxf(0.001f, roughness);
  const float a2 = a * a;
  const float phi = r1 * TWO_PI;
  const float cos_theta = sqrtf((1.0f - powf(a2, 1.0f - r2)) / (1.0f - a2));
  const float sin_theta = clampf(sqrtf(1.0f - (cos_theta * cos_theta)), 0.0f, 1.0f);
  const float cos_phi = cosf(phi);
  // y = sin_theta * sin_theta exactly as the reference writes it (:43).
  return mk3(sin_theta * cos_phi, sin_theta * sin_theta, cos_theta);
}
__device__ __forceinline__ f3 sample_ggx_vndf(f3 wi, float ax, float ay, float r1, float r2) {  // :120-136
  const f3 wiStd = normalize(mk3(wi.x * ax, wi.y * ay, wi.z));
  const float phi = (2.0f * r1 - 1.0f) * PI;
  const float z = fmaf((1.0f - r2), (1.0f + wiStd.z), -wiStd.z);
  const float sinTheta = sqrtf(clampf(1.0f - z * z, 0.0f, 1.0f));
  float sp, cp;
  sincosf(phi, &sp, &cp);
  const f3 c = mk3(sinTheta * cp, sinTheta * sp, z);
  const f3 wmStd = c + wiStd;
  return normalize(mk3(wmStd.x * ax, wmStd.y * ay, wmStd.z));
}
__device__ __forceinline__ float gtr2_aniso(float n_dot_h, float h_dot_x, float h_dot_y, float ax, float ay) {  // :146-154
  const float a = h_dot_x / ax;
  const float b = h_dot_y / ay;
  const float c = a * a + b * b + n_dot_h * n_dot_h;
  return 1.0f / (PI * ax * ay * c * c);
}
__device__ __forceinline__ float smith_g(float n_dot_v, float alpha_g) {  // :181-188
  const float a = alpha_g * alpha_g;
  const float b = n_dot_v * n_dot_v;
  return (2.0f * n_dot_v) / (n_dot_v + sqrtf(a + b - a * b));
}
__device__ __forceinline__ float smith_g_aniso(float n_dot_v, float v_dot_x, float v_dot_y, float ax, float ay) {  // :193-202
  const float a = v_dot_x * ax;
  const float b = v_dot_y * ay;
  const float c = n_dot_v;
  return (2.0f * n_dot_v) / (n_dot_v + sqrtf(a * a + b * b + c * c));
}
__device__ __forceinline__ float schlick_weight(float u) {  // :207-211
  const float m = clampf(1.0f - u, 0.0f, 1.0f);
  const float m2 = m * m;
  return m2 * m2 * m;
}
__device__ __forceinline__ float dielectric_fresnel(float cos_theta_i, float eta) {  // :218-236
  const float s2 = eta * eta * (1.0f - cos_theta_i * cos_theta_i);
  if (s2 > 1.0f) return 1.0f;
  const float cos_theta_t = sqrtf(fmaxf(1.0f - s2, 0.0f));
  const float rs = (eta * cos_theta_t - cos_theta_i) / (eta * cos_theta_t + cos_theta_i);
  const float rp = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
  return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ f3 cosine_sample_hemisphere(float r1, float r2) {  // :252-260
  const float r = sqrtf(r1);
  const float phi = TWO_PI * r2;
  float sp, cp;
  sincosf(phi, &sp, &cp);
  const float x = r * cp, y = r * sp;
  return mk3(x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}
__device__ __forceinline__ f3 uniform_sample_hemisphere(float r1, float r2) {  // :266-270
  const float r = sqrtf(fmaxf(0.0f, 1.0f - r1 * r1));
  const float phi = TWO_PI * r2;
  float sp, cp;
  sincosf(phi, &sp, &cp);
  return mk3(r * cp, r * sp, r1);
}
__device__ __forceinline__ float power_heuristic(float a, float b) {  // :286-290
  a *= a;
  b *= b;
  return a / (a + b);
}
__device__ __forceinline__ void onb(f3 N, f3& T, f3& B) {  // :296-300
  const f3 up = fabsf(N.z) < 0.9999999f ? mk3(0.f, 0.f, 1.f) : mk3(1.f, 0.f, 0.f);
  T = normalize(cross(up, N));
  B = cross(N, T);
}
__device__ __forceinline__ f3 sample_hg(f3 V, float g, float r1, float r2) {  // :308-327
  float cos_theta;
  if (fabsf(g) < 0.001f) {
    cos_theta = 1.0f - 2.0f * r2;
  } else {
    const float sqr_term = (1.0f - g * g) / (1.0f + g - 2.0f * g * r2);
    cos_theta = -(1.0f + g * g - sqr_term * sqr_term) / (2.0f * g);
  }
  const float phi = r1 * TWO_PI;
  const float sin_theta = clampf(sqrtf(1.0f - (cos_theta * cos_theta)), 0.0f, 1.0f);
  float sp, cp;
  sincosf(phi, &sp, &cp);
  f3 v1, v2;
  onb(V, v1, v2);
  return sin_theta * cp * v1 + sin_theta * sp * v2 + cos_theta * V;
}
__device__ __forceinline__ float phase_hg(float cos_theta, float g) {  // :333-336
  const float denom = fmaxf(1.0f + g * g + 2.0f * g * cos_theta, EPS);
  return INV_FOUR_PI * (1.0f - g * g) / (denom * sqrtf(denom));
}

// -------------------------------------------------------------- materials
// The Material fields the BxDFs read, loaded once per shading event.
struct Mat {
  f3 base_color, emission, clearcoat_tint, medium_color;
  float opacity, metallic, roughness, subsurface, specular_tint, sheen, sheen_tint;
  float clearcoat, clearcoat_roughness, specular_transmission, ior, ax, ay;
  float medium_density, medium_anisotropy;
  uint32_t medium_type, type;
  uint32_t base_color_map, normal_map, metallic_roughness_map, emission_map;
};

__device__ __forceinline__ Mat load_material(const hpt_material* mp) {
  const float4* q = reinterpret_cast<const float4*>(mp);
  const float4 m0 = q[0], m1 = q[1], m2 = q[2], m3 = q[3], m4 = q[4], m5 = q[5], m6 = q[6], m7 = q[7],
               m8 = q[8];
  Mat m;
  m.medium_color = mk3(m0.x, m0.y, m0.z);
  m.medium_density = m0.w;
  m.medium_anisotropy = m1.x;
  m.medium_type = __float_as_uint(m1.y);
  m.base_color = mk3(m2.x, m2.y, m2.z);
  m.opacity = m2.w;
  m.emission = mk3(m3.x, m3.y, m3.z);
  m.metallic = m4.x;
  m.roughness = m4.y;
  m.subsurface = m4.z;
  m.specular_tint = m4.w;
  m.sheen = m5.x;
  m.sheen_tint = m5.y;
  m.clearcoat = m5.z;
  m.clearcoat_roughness = m5.w;
  m.clearcoat_tint = mk3(m6.x, m6.y, m6.z);
  m.specular_transmission = m6.w;
  m.ior = m7.x;
  m.ax = m7.y;
  m.ay = m7.z;
  m.type = __float_as_uint(m8.w);
  m.base_color_map = __float_as_uint(m7.w);
  m.normal_map = __float_as_uint(m8.x);
  m.metallic_roughness_map = __float_as_uint(m8.y);
  m.emission_map = __float_as_uint(m8.z);
  return m;
}

// textureLod(g_textures[t], uv, 0) (pathtracer.rgen.glsl:122-146): bilinear,
// REPEAT in u and v, texel centres at (i + 0.5) / W.  Colour maps (base
// colour, emission) decode each texel from sRGB before filtering, as an sRGB
// image format does; data maps (metallic-roughness, normal) read c / 255.
__device__ __forceinline__ float4 tex_sample(const SceneView& sv, uint32_t t, float u, float v, bool srgb) {
  const TexInfo ti = sv.tex_info[t];
  const int w = (int)ti.width, h = (int)ti.height;
  const float x = u * (float)w - 0.5f;
  const float y = v * (float)h - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float fx = x - fx0, fy = y - fy0;
  int i0 = (int)fx0 % w;
  if (i0 < 0) i0 += w;
  int j0 = (int)fy0 % h;
  if (j0 < 0) j0 += h;
  const int i1 = i0 + 1 == w ? 0 : i0 + 1;
  const int j1 = j0 + 1 == h ? 0 : j0 + 1;
  const uchar4* px = sv.tex_texels + ti.offset;
  auto texel = [&](int i, int j) {
    const uchar4 c = px[(size_t)j * w + i];
    if (srgb) return make_float4(sv.srgb_lut[c.x], sv.srgb_lut[c.y], sv.srgb_lut[c.z], (float)c.w / 255.0f);
    return make_float4((float)c.x / 255.0f, (float)c.y / 255.0f, (float)c.z / 255.0f, (float)c.w / 255.0f);
  };
  const float4 a = texel(i0, j0), b = texel(i1, j0), c = texel(i0, j1), d = texel(i1, j1);
  auto mix4 = [](float4 p, float4 q, float s) {
    return make_float4(mixf(p.x, q.x, s), mixf(p.y, q.y, s), mixf(p.z, q.z, s), mixf(p.w, q.w, s));
  };
  return mix4(mix4(a, b, fx), mix4(c, d, fx), fy);
}

// Shading frame of one interaction (the State fields the BxDFs read).
struct Frame {
  f3 tangent, bitangent;
  float eta;
  uint32_t depth;
};

// ------------------------------------------------------- diffuse (callables 11/12)
// inc/bxdf/oren_nayar.glsl:7-51.
__device__ __forceinline__ f3 oren_nayar_f(const Mat& mat, f3 wo, f3 wi) {
  const float sin_theta_i = sin_theta_l(wi);
  const float sin_theta_o = sin_theta_l(wo);
  float max_cos = 0.0f;
  if (sin_theta_i > EPS && sin_theta_o > EPS) {
    const float d_cos = cos_phi_l(wi) * cos_phi_l(wo) + sin_phi_l(wi) * sin_phi_l(wo);
    max_cos = fmaxf(0.0f, d_cos);
  }
  float sin_alpha, tan_beta;
  if (fabsf(wi.z) > fabsf(wo.z)) {
    sin_alpha = sin_theta_o;
    tan_beta = sin_theta_i / fabsf(wi.z);
  } else {
    sin_alpha = sin_theta_i;
    tan_beta = sin_theta_o / fabsf(wo.z);
  }
  return mat.base_color * INV_PI * (mat.ax + mat.ay * max_cos * sin_alpha * tan_beta);
}
// diffuse_bxdf_eval.rcall.glsl:15-21: Oren-Nayar when roughness > EPS
// (oren_nayar.glsl:89-96) else Lambert (lambert.glsl:37-42).
__device__ __forceinline__ f3 diffuse_eval(const Frame& fr, const Mat& mat, f3 V, f3 N, f3 L, float& pdf) {
  if (mat.roughness > EPS) {
    const f3 Ll = to_local(fr.tangent, fr.bitangent, N, L);
    const f3 Vl = to_local(fr.tangent, fr.bitangent, N, V);
    pdf = Ll.z * INV_PI;
    return oren_nayar_f(mat, Vl, Ll) * Ll.z;
  }
  const float n_dot_l = dot(N, L);
  pdf = n_dot_l * INV_PI;
  return mat.base_color * n_dot_l * INV_PI;
}

// --------------------------------------------------------- Disney (13/14)
// inc/bxdf/disney.glsl:9-32.
__device__ __forceinline__ void tint_colors(f3 base_color, float specular_tint, float sheen_tint, float eta,
                                            float& F0, f3& c_sheen, f3& c_spec0) {
  const float lum = luminance(base_color);
  const f3 ctint = lum > 0.0f ? base_color / lum : mk3(1.0f);
  F0 = (1.0f - eta) / (1.0f + eta);
  F0 *= F0;
  c_spec0 = F0 * mix3(mk3(1.0f), ctint, specular_tint);
  c_sheen = mix3(mk3(1.0f), ctint, sheen_tint);
}
__device__ __forceinline__ float diffuse_directional_albedo(float theta, float ax, float ay, float e0) {  // :34-37
  return mixf(e0 + (1.0f - e0) * powf(fabsf(1.0f - theta), 5.0f), 0.04762f + 0.95238f * e0,
              1.0f - powf(fabsf(1.0f - ax * ay), 5.0f));
}
__device__ __forceinline__ float coupled_diffuse(float ax, float ay, float dot_wi_n, float dot_wo_n, float e0) {  // :39-48
  const float Ewi = diffuse_directional_albedo(dot_wi_n, ax, ay, e0);
  const float Ewo = diffuse_directional_albedo(dot_wo_n, ax, ay, e0);
  const float Eavg = e0 + (-0.33263f * ax * ay - 0.072359f) * (1.0f - e0) * e0;
  return (1.0f - Ewo) * (1.0f - Ewi) / (PI * (1.0f - Eavg));
}
__device__ __forceinline__ float sheen_directional_albedo(float cos_theta, float alpha) {  // :50-54
  const float c = 1.0f - cos_theta;
  const float c3 = c * c * c;
  return 0.65584461f * c3 + 1.0f / (4.16526551f + expf(-7.97291361f * sqrtf(alpha) + 6.33516894f));
}
// :71-113
__device__ __forceinline__ f3 eval_disney_diffuse(float ax, float ay, float F0, const Mat& m, f3 c_sheen,
                                                  f3 V, f3 L, f3 H, float& pdf) {
  pdf = 0.0f;
  if (L.z <= 0.0f) return mk3(0.0f);
  const float L_dot_H = dot(L, H);
  const float retro = 2.0f * m.roughness * L_dot_H * L_dot_H;
  const float f_L = schlick_weight(L.z);
  const float f_V = schlick_weight(V.z);
  const float f_retro = retro * (f_L + f_V + f_L * f_V * (retro - 1.0f));
  const float f_d = (1.0f - 0.5f * f_L) * (1.0f - 0.5f * f_V);
  const float coupled = coupled_diffuse(ax, ay, fabsf(V.z), fabsf(L.z), F0 * m.specular_tint);
  const float f_ss90 = 0.5f * retro;
  const float f_ss = mixf(1.0f, f_ss90, f_L) * mixf(1.0f, f_ss90, f_V);
  const float ss = 1.25f * (f_ss * (1.0f / (L.z + V.z) - 0.5f) + 0.5f);
  const float f_H = schlick_weight(L_dot_H);
  const f3 f_sheen = f_H * m.sheen * c_sheen;
  float alpha = fmaxf(m.roughness, 0.07f);
  alpha = alpha * alpha;
  const float Ewi = max3(c_sheen) * sheen_directional_albedo(fabsf(V.z), alpha);
  const float Ewo = max3(c_sheen) * sheen_directional_albedo(fabsf(L.z), alpha);
  const float sheen_weight = fminf(1.0f - Ewi, 1.0f - Ewo);
  pdf = L.z * INV_PI;
  return INV_PI * m.base_color * mixf(f_d * coupled * PI + f_retro, ss, m.subsurface) + f_sheen * sheen_weight;
}
// :124-143
__device__ __forceinline__ f3 eval_microfacet_reflection(float ax, float ay, f3 V, f3 L, f3 H, f3 F, float& pdf) {
  pdf = 0.0f;
  if (L.z <= 0.0f) return mk3(0.0f);
  const float D = gtr2_aniso(H.z, H.x, H.y, ax, ay);
  const float G1 = smith_g_aniso(fabsf(V.z), V.x, V.y, ax, ay);
  const float G2 = G1 * smith_g_aniso(fabsf(L.z), L.x, L.y, ax, ay);
  pdf = G1 * D / (4.0f * V.z);
  return F * D * G2 / (4.0f * L.z * V.z);
}
__device__ __forceinline__ float specular_directional_albedo(float alpha, float c) {  // :145-149
  return 1.0f - 1.45940f * alpha * (-0.20276f + alpha * (2.77203f + (-2.61748f + 0.73343f * alpha) * alpha)) * c *
                    (3.09507f + c * (-9.11368f + c * (15.88844f + c * (-13.70343f + 4.51786f * c))));
}
// :151-177
__device__ __forceinline__ f3 eval_microfacet_ms(float ax, float ay, f3 V, f3 L, f3 F0) {
  const float alpha = sqrtf(ax * ay);
  const float Ewi = specular_directional_albedo(alpha, fabsf(V.z));
  const float Ewo = specular_directional_albedo(alpha, fabsf(L.z));
  const float Eavg = 1.0f + alpha * (-0.11304f + alpha * (-1.86947f + (2.22682f - 0.83397f * alpha) * alpha));
  const float ms = (1.0f - Ewo) * (1.0f - Ewi) / (PI * (1.0f - Eavg));
  const f3 Favg = (20.0f / 21.0f) * F0 + (1.0f / 21.0f) * mk3(1.0f);
  const f3 f = (Favg * Favg * Eavg) / (1.0f - Favg * (1.0f - Eavg));
  return ms * f;
}
// :190-219
__device__ __forceinline__ f3 eval_microfacet_refraction(float ax, float ay, f3 base_color, float eta, f3 V, f3 L,
                                                         f3 H, float F, float& pdf) {
  pdf = 0.0f;
  if (L.z >= 0.0f) return mk3(0.0f);
  const float l_dot_h = dot(L, H);
  const float v_dot_h = dot(V, H);
  const float D = gtr2_aniso(H.z, H.x, H.y, ax, ay);
  const float G1 = smith_g_aniso(fabsf(V.z), V.x, V.y, ax, ay);
  const float G2 = G1 * smith_g_aniso(fabsf(L.z), L.x, L.y, ax, ay);
  float denom = l_dot_h + v_dot_h * eta;
  denom *= denom;
  const float eta2 = eta * eta;
  const float jacobian = fabsf(l_dot_h) / denom;
  pdf = G1 * fmaxf(0.0f, v_dot_h) * D * jacobian / V.z;
  const f3 sq = mk3(sqrtf(base_color.x), sqrtf(base_color.y), sqrtf(base_color.z));  // pow(base_color, 0.5)
  return sq * (1.0f - F) * D * G2 * fabsf(v_dot_h) * jacobian * eta2 / fabsf(L.z * V.z);
}
// :229-252
__device__ __forceinline__ float eval_clearcoat(float clearcoat_roughness, f3 V, f3 L, f3 H, float& pdf) {
  pdf = 0.0f;
  if (L.z <= 0.0f) return 0.0f;
  const float v_dot_h = dot(V, H);
  const float F = mixf(0.04f, 1.0f, schlick_weight(v_dot_h));
  const float D = gtr1(H.z, clearcoat_roughness);
  const float G = smith_g(L.z, 0.25f) * smith_g(V.z, 0.25f);
  const float jacobian = 1.0f / (4.0f * v_dot_h);
  pdf = D * H.z * jacobian;
  return F * D * G;
}

struct LobePr {
  float diffuse, dielectric, metal, glass, clearcoat;
  float dielectric_wt, metal_wt, glass_wt;
  float F0;
  f3 c_sheen, c_spec0;
};
// :300-331 / :433-464: lobe weights and normalised probabilities.
__device__ __forceinline__ LobePr lobe_probs(const Mat& m, float eta, float Vz) {
  LobePr p;
  tint_colors(m.base_color, m.specular_tint, m.sheen_tint, eta, p.F0, p.c_sheen, p.c_spec0);
  p.dielectric_wt = (1.0f - m.metallic) * (1.0f - m.specular_transmission);
  p.metal_wt = m.metallic;
  p.glass_wt = (1.0f - m.metallic) * m.specular_transmission;
  const float schlick_wt = schlick_weight(Vz);
  float d = p.dielectric_wt * luminance(m.base_color);
  float s = p.dielectric_wt * luminance(mix3(p.c_spec0, mk3(1.0f), schlick_wt));
  float me = p.metal_wt * luminance(mix3(m.base_color, mk3(1.0f), schlick_wt));
  float g = p.glass_wt;
  float c = 0.25f * m.clearcoat;
  const float inv_total = 1.0f / (d + s + me + g + c);
  p.diffuse = d * inv_total;
  p.dielectric = s * inv_total;
  p.metal = me * inv_total;
  p.glass = g * inv_total;
  p.clearcoat = c * inv_total;
  return p;
}

// :264-390; V, L in world space.
__device__ __forceinline__ f3 disney_eval(bool anb, const Frame& fr, const Mat& m, f3 V, f3 N, f3 L, float& pdf) {
  float ax = m.ax, ay = m.ay;
  // ROUGHNESS_REGULARIZE_EVAL (make_shaders.yaml:4).
  if (fr.depth > 0 && anb) {
    if (ax < 0.3f) ax = clampf(2.0f * ax, 0.1f, 0.3f);
    if (ay < 0.3f) ay = clampf(2.0f * ay, 0.1f, 0.3f);
  }
  pdf = 0.0f;
  f3 f = mk3(0.0f);
  V = to_local(fr.tangent, fr.bitangent, N, V);
  L = to_local(fr.tangent, fr.bitangent, N, L);
  f3 H = L.z > 0.0f ? normalize(L + V) : normalize(L + V * fr.eta);
  if (H.z < 0.0f) H = -H;
  const LobePr pr = lobe_probs(m, fr.eta, V.z);
  const bool refl = L.z * V.z > 0.0f;
  float tmp = 0.0f;
  const float v_dot_h = fabsf(dot(V, H));
  if (pr.diffuse > 0.0f && refl) {
    f += eval_disney_diffuse(ax, ay, pr.F0, m, pr.c_sheen, V, L, H, tmp) * pr.dielectric_wt;
    pdf += tmp * pr.diffuse;
  }
  if (pr.dielectric > 0.0f && refl) {
    const float F = (dielectric_fresnel(v_dot_h, fr.eta) - pr.F0) / (1.0f - pr.F0);
    f += eval_microfacet_reflection(ax, ay, V, L, H, mix3(pr.c_spec0, mk3(1.0f), F), tmp) * pr.dielectric_wt;
    f += eval_microfacet_ms(ax, ay, V, L, pr.c_spec0) * pr.dielectric_wt;
    pdf += tmp * pr.dielectric;
  }
  if (pr.metal > 0.0f && refl) {
    const f3 F = mix3(m.base_color, mk3(1.0f), schlick_weight(v_dot_h));
    f += eval_microfacet_reflection(ax, ay, V, L, H, F, tmp) * pr.metal_wt;
    f += eval_microfacet_ms(ax, ay, V, L, m.base_color) * pr.metal_wt;
    pdf += tmp * pr.metal;
  }
  if (pr.glass > 0.0f) {
    const float F = dielectric_fresnel(v_dot_h, fr.eta);
    if (refl) {
      f += eval_microfacet_reflection(ax, ay, V, L, H, mk3(F), tmp) * pr.glass_wt;
      f += eval_microfacet_ms(ax, ay, V, L, pr.c_spec0) * pr.glass_wt;
      pdf += tmp * pr.glass * F;
    } else {
      f += eval_microfacet_refraction(ax, ay, m.base_color, fr.eta, V, L, H, F, tmp) * pr.glass_wt;
      pdf += tmp * pr.glass * (1.0f - F);
    }
  }
  if (pr.clearcoat > 0.0f && refl) {
    f += mk3(eval_clearcoat(m.clearcoat_roughness, V, L, H, tmp)) * m.clearcoat_tint * 0.25f * m.clearcoat;
    pdf += tmp * pr.clearcoat;
  }
  return f * fabsf(L.z);
}

// :405-531
__device__ __forceinline__ f3 disney_sample(uint4& rng, bool& anb, const Frame& fr, const Mat& m, f3 V, f3 N,
                                            f3& L, float& pdf, uint32_t& flags) {
  const float ax = m.ax, ay = m.ay;  // ROUGHNESS_REGULARIZE_SAMPLE is off (make_shaders.yaml:5)
  anb = true;
  pdf = 0.0f;
  const float r1 = rnd(rng);
  const float r2 = rnd(rng);
  V = to_local(fr.tangent, fr.bitangent, N, V);
  const LobePr pr = lobe_probs(m, fr.eta, V.z);
  const float cdf0 = pr.diffuse;
  const float cdf1 = cdf0 + pr.dielectric;
  const float cdf2 = cdf1 + pr.metal;
  const float cdf3 = cdf2 + pr.glass;
  float r3 = rnd(rng);
  if (r3 < cdf0) {
    L = cosine_sample_hemisphere(r1, r2);
    flags = RAY_FLAGS_REFLECTION;
  } else if (r3 < cdf2) {
    f3 H = sample_ggx_vndf(V, ax, ay, r1, r2);
    if (H.z < 0.0f) H = -H;
    L = normalize(reflect3(-V, H));
    flags = RAY_FLAGS_REFLECTION;
  } else if (r3 < cdf3) {
    f3 H = sample_ggx_vndf(V, ax, ay, r1, r2);
    const float F = dielectric_fresnel(fabsf(dot(V, H)), fr.eta);
    if (H.z < 0.0f) H = -H;
    r3 = (r3 - cdf2) / (cdf3 - cdf2);
    if (r3 < F) {
      L = normalize(reflect3(-V, H));
      flags = RAY_FLAGS_REFLECTION;
    } else {
      L = normalize(refract3(-V, H, fr.eta));
      flags = RAY_FLAGS_TRANSMISSION;
    }
  } else {
    f3 H = sample_gtr1(m.clearcoat_roughness, r1, r2);
    if (H.z < 0.0f) H = -H;
    L = normalize(reflect3(-V, H));
    flags = RAY_FLAGS_REFLECTION;
  }
  L = to_world(fr.tangent, fr.bitangent, N, L);
  V = to_world(fr.tangent, fr.bitangent, N, V);
  return disney_eval(anb, fr, m, V, N, L, pdf);
}

// Callable dispatch CALLABLE_MATERIAL_BXDF_BEGIN + mat.type*2 (+1).
__device__ __forceinline__ f3 bxdf_eval(bool anb, const Frame& fr, const Mat& m, f3 V, f3 N, f3 L, float& pdf) {
  if (m.type == HPT_MATERIAL_DISNEY) return disney_eval(anb, fr, m, V, N, L, pdf);
  return diffuse_eval(fr, m, V, N, L, pdf);
}
__device__ __forceinline__ f3 bxdf_sample(uint4& rng, bool& anb, const Frame& fr, const Mat& m, f3 V, f3 N, f3& L,
                                          float& pdf, uint32_t& flags) {
  if (m.type == HPT_MATERIAL_DISNEY) return disney_sample(rng, anb, fr, m, V, N, L, pdf, flags);
  // diffuse_bxdf_sample.rcall.glsl:15-38 -> lambert.glsl:15-27 / oren_nayar.glsl:67-79
  const float r1 = rnd(rng);
  const float r2 = rnd(rng);
  L = to_world(fr.tangent, fr.bitangent, N, cosine_sample_hemisphere(r1, r2));
  anb = true;
  flags = RAY_FLAGS_REFLECTION;
  return diffuse_eval(fr, m, V, N, L, pdf);
}

// ------------------------------------------------------------------ lights
struct LightSample {
  f3 normal, emission, direction;
  float dist, pdf;
};
// The *_light_sample callables (inc/light/{point,distant,spot,quad,sphere}.glsl).
__device__ __forceinline__ void light_sample(uint4& rng, uint32_t num_lights, f3 pos, const hpt_light& Ls,
                                             LightSample& o) {
  const f3 inten = ld3(Ls.intensity), lp = ld3(Ls.position), lu = ld3(Ls.u), lv = ld3(Ls.v);
  const float nl = (float)num_lights;
  o.normal = mk3(0.0f);
  o.emission = mk3(0.0f);
  o.direction = mk3(0.0f);
  o.dist = 0.0f;
  o.pdf = 0.0f;
  switch (Ls.type) {
    case HPT_LIGHT_POINT: {  // point.glsl:10-20
      f3 d = lp - pos;
      const float dist_sq = dot(d, d);
      o.dist = sqrtf(dist_sq);
      d = d / o.dist;
      o.normal = -d;
      o.emission = inten * nl / dist_sq;
      o.pdf = 1.0f;
      o.direction = d;
      break;
    }
    case HPT_LIGHT_DIRECTIONAL: {  // distant.glsl:10-34
      f3 direction = -lu;
      const float cos_alpha = lv.x;
      if (cos_alpha > EPS) {
        const f3 N = direction;
        f3 T, B;
        onb(N, T, B);
        const float r = rnd(rng);
        const float phi = TWO_PI * rnd(rng);
        const float z = (1.0f - r * (1.0f - cos_alpha));
        const float radius = sqrtf(1.0f - z * z);
        float sp, cp;
        sincosf(phi, &sp, &cp);
        direction = to_world(T, B, N, mk3(radius * cp, radius * sp, z));
      }
      o.normal = lu;
      o.emission = inten * nl;
      o.dist = WORLD_SIZE;
      o.pdf = 1.0f;
      o.direction = direction;
      break;
    }
    case HPT_LIGHT_SPOT: {  // spot.glsl:10-20
      f3 d = lp - pos;
      const float dist_sq = dot(d, d);
      o.dist = sqrtf(dist_sq);
      d = d / o.dist;
      o.normal = -d;
      const float e0 = lv.y, e1 = lv.x;
      const float tt = clampf((dot(-d, lu) - e0) / (e1 - e0), 0.0f, 1.0f);
      const float ss = tt * tt * (3.0f - 2.0f * tt);
      o.emission = ss * inten * nl / dist_sq;
      o.pdf = 1.0f;
      o.direction = d;
      break;
    }
    case HPT_LIGHT_QUAD: {  // quad.glsl:10-34
      const float r1 = rnd(rng);
      const float r2 = rnd(rng);
      const f3 sp = lp + lu * r1 + lv * r2;
      f3 d = sp - pos;
      o.normal = normalize(cross(lv, lu));
      const float cos_theta = dot(-d, o.normal);
      if (cos_theta < 0.0f) break;
      const float dist_sq = dot(d, d);
      o.dist = sqrtf(dist_sq);
      d /= o.dist;
      o.emission = inten * nl;
      o.pdf = dist_sq / (Ls.area * fabsf(cos_theta));
      o.direction = d;
      break;
    }
    case HPT_LIGHT_SPHERE: {  // sphere.glsl:10-42
      const float r1 = rnd(rng);
      const float r2 = rnd(rng);
      f3 c2s = pos - lp;
      const float dc = length(c2s);
      c2s /= dc;
      if (dc < Ls.radius) {
        o.normal = c2s;
        o.dist = dc;
        break;
      }
      f3 sd = uniform_sample_hemisphere(r1, r2);
      f3 T, B;
      onb(c2s, T, B);
      sd = T * sd.x + B * sd.y + c2s * sd.z;
      const f3 lsp = lp + sd * Ls.radius;
      f3 d = lsp - pos;
      const float dist_sq = dot(d, d);
      o.dist = sqrtf(dist_sq);
      d = d / o.dist;
      o.normal = normalize(lsp - lp);
      o.emission = inten * nl;
      o.pdf = dist_sq / (Ls.area * 0.5f * fabsf(dot(o.normal, d)));
      o.direction = d;
      break;
    }
    default:
      break;
  }
}

// ----------------------------------------------------------- environment
// Bilinear fetch, REPEAT in u, CLAMP_TO_EDGE in v, texel centres (i+0.5)/W.
__device__ __forceinline__ f3 env_texture(const SceneView& sv, uint32_t W, uint32_t H, float u, float v) {
  const float x = u * (float)W - 0.5f;
  const float y = v * (float)H - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float fx = x - fx0, fy = y - fy0;
  const int w = (int)W, h = (int)H;
  int i0 = (int)fx0 % w;
  if (i0 < 0) i0 += w;
  const int i1 = i0 + 1 == w ? 0 : i0 + 1;
  const int j = (int)fy0;
  const int j0 = min(max(j, 0), h - 1), j1 = min(max(j + 1, 0), h - 1);
  const float4 a = sv.env_rgb[j0 * w + i0], b = sv.env_rgb[j0 * w + i1];
  const float4 c = sv.env_rgb[j1 * w + i0], d = sv.env_rgb[j1 * w + i1];
  const f3 top = mix3(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), fx);
  const f3 bot = mix3(mk3(c.x, c.y, c.z), mk3(d.x, d.y, d.z), fx);
  return mix3(top, bot, fy);
}

// env_texture through the record of texel (ti, tj): the same four texels
// and the same arithmetic when the bilinear footprint lies in the record's
// 3x3 neighbourhood (always, for points inside or on the edge of the texel),
// the texture itself otherwise.
__device__ __forceinline__ f3 env_texture_at(const SceneView& sv, uint32_t W, uint32_t H, float u, float v,
                                             uint32_t ti, uint32_t tj) {
  const float x = u * (float)W - 0.5f;
  const float y = v * (float)H - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float fx = x - fx0, fy = y - fy0;
  const int w = (int)W;
  int i0 = (int)fx0 % w;
  if (i0 < 0) i0 += w;
  const int j = (int)fy0;
  const int di = i0 - (int)ti;
  const int dx = di == 0 ? 0 : ((di == -1 || di == w - 1) ? -1 : 2);
  const int dy = j - (int)tj;
  if (dx > 0 || dy < -1 || dy > 0) return env_texture(sv, W, H, u, v);
  const float* t = sv.env_texels[tj * W + ti].rgb[0][0];
  const float* r0 = t + (dy + 1) * 9 + (dx + 1) * 3;  // row j0, column i0
  const float* r1 = r0 + 9;                           // row j1
  const f3 a = ld3(r0), b = ld3(r0 + 3), c = ld3(r1), d = ld3(r1 + 3);
  const f3 top = mix3(a, b, fx);
  const f3 bot = mix3(c, d, fx);
  return mix3(top, bot, fy);
}

__device__ __forceinline__ uint32_t env_texel(uint32_t W, uint32_t H, float u, float v) {
  const float fu = u - floorf(u);
  int i = (int)(fu * (float)W);
  int j = (int)floorf(v * (float)H);
  i = min(max(i, 0), (int)W - 1);
  j = min(max(j, 0), (int)H - 1);
  return (uint32_t)j * W + (uint32_t)i;
}

// inc/env_map.glsl:6-37 with the marginal/conditional textures replaced by a
// luminance alias table: one pcg4d step (as rand2), x picks the texel, y
// accepts it or its alias, z and w place the point inside the texel.
__device__ __forceinline__ f3 sample_env_map(uint4& rng, const SceneView& sv, const FrameParams& fp, f3& color,
                                             float& pdf) {
  pcg4d(rng);
  const uint32_t W = fp.env_w, H = fp.env_h;
  uint32_t idx = (uint32_t)(((uint64_t)rng.x * (uint64_t)(W * H)) >> 32);
  const float4 e = *reinterpret_cast<const float4*>(sv.env_texels + idx);
  float lum = e.z;
  if (!(u2f(rng.y) < e.x)) {
    idx = __float_as_uint(e.y);
    lum = e.w;
  }
  const uint32_t ti = idx % W, tj = idx / W;
  const float u = ((float)ti + u2f(rng.z)) / (float)W;
  const float v = ((float)tj + u2f(rng.w)) / (float)H;
  color = env_texture_at(sv, W, H, u, v, ti, tj) * fp.env_intensity;
  pdf = lum / fp.env_total_sum;
  const float phi = (u + fp.env_rotation) * TWO_PI;
  const float theta = v * PI;
  float st, ct, sp, cp;
  sincosf(theta, &st, &ct);
  sincosf(phi, &sp, &cp);
  if (st <= EPS) pdf = 0.0f;
  const f3 dir = normalize(mk3(-st * cp, st * sp, ct));
  pdf = (pdf * (float)W * (float)H) / (TWO_PI * PI * st);
  return dir;
}

// env_map_eval.rcall.glsl:14-35 (MAP) and env_sky_eval.rcall.glsl:12-15 (SKY).
__device__ __forceinline__ f3 env_eval(const SceneView& sv, const FrameParams& fp, uint32_t depth, float bsdf_pdf,
                                       f3 dir, bool mis_off) {
  if (fp.env_type == HPT_ENV_SKY) {
    const float a = fmaxf(dir.z, 0.0f);
    return mix3(ld3(fp.ground), ld3(fp.sky), a) * fp.env_intensity;
  }
  // acos is undefined above 1 in GLSL; a normalised direction can carry
  // z = 1 + 2^-23, so z is clamped to [-1, 1] (as the oracle does).
  const float theta = acosf(fminf(fmaxf(dir.z, -1.0f), 1.0f));
  const float u = -(PI + atan2f(dir.y, dir.x)) * INV_TWO_PI - fp.env_rotation;
  const float v = theta * INV_PI;
  const uint32_t texel = env_texel(fp.env_w, fp.env_h, u, v);
  float mis_weight = 1.0f;
  if (depth > 0) {
    // env_pdf, inc/env_map.glsl:43-55.
    float p = 0.0f;
    const float st = sinf(theta);
    if (st > EPS) {
      const float pp = sv.env_texels[texel].lum / fp.env_total_sum;
      if (pp > EPS) p = (pp * (float)fp.env_w * (float)fp.env_h) / (TWO_PI * PI * st);
    }
    mis_weight = power_heuristic(bsdf_pdf, p);
  }
  if (mis_off) mis_weight = 1.0f;
  const f3 L = env_texture_at(sv, fp.env_w, fp.env_h, u, v, texel % fp.env_w, texel / fp.env_w);
  return mis_weight * L * fp.env_intensity;
}

// ----------------------------------------------------------------- cameras
// perspective_ray.rcall.glsl:12-64 / orthographic_ray.rcall.glsl:12-45 with
// rand4blue (common/wltcRS.glsl:46-53) over the 1024x1024 RGBA8 texture.
__device__ __forceinline__ void camera_ray(const SceneView& sv, const FrameParams& fp, uint32_t frame, uint32_t x,
                                           uint32_t y, f3& org, f3& dir) {
#pragma clang fp contract(off)
  uint4 s1 = make_uint4(frame, frame * 15843u, frame * 31u + 4566u, frame * 2345u + 58585u);
  pcg4d(s1);
  const int tx = ((int)x + (int)(s1.x % 0x0fffffffu)) % 1024;
  const int ty = ((int)y + (int)(s1.y % 0x0fffffffu)) % 1024;
  const uchar4 bn = sv.blue_noise[ty * 1024 + tx];
  const float rx = (float)bn.x / 255.0f, ry = (float)bn.y / 255.0f;
  const float rz = (float)bn.z / 255.0f, rw = (float)bn.w / 255.0f;
  const hpt_camera& cam = sv.cameras[fp.camera_index];
  const float r1 = 2.0f * rx, r2 = 2.0f * ry;
  float jx = r1 < 1.0f ? sqrtf(r1) - 1.0f : 1.0f - sqrtf(2.0f - r1);
  float jy = r2 < 1.0f ? sqrtf(r2) - 1.0f : 1.0f - sqrtf(2.0f - r2);
  jx = jx / (fp.res_x * 0.5f);
  jy = jy / (fp.res_y * 0.5f);
  const float sux = (float)x / fp.res_x;
  const float suy = (fp.res_y - (float)y) / fp.res_y;
  float dx = (2.0f * sux - 1.0f) + jx;
  float dy = (2.0f * suy - 1.0f) + jy;
  const f3 right = ld3(cam.right), up = ld3(cam.up), fwd = ld3(cam.forward), pos = ld3(cam.position);
  if (cam.type == HPT_CAMERA_ORTHOGRAPHIC) {
    dir = normalize(fwd);
    org = pos + right * (dx * cam.focal_distance_or_xmag) + up * (dy * cam.aperture_or_ymag);
    return;
  }
  const float scale = fp.tan_half_fov;
  dx *= scale * fp.res_x / fp.res_y;
  dy *= scale;
  const f3 rd = normalize(right * dx + up * dy + fwd);
  dir = rd;
  org = pos;
  const float aperture = cam.aperture_or_ymag;
  if (aperture > 0.0f) {
    const f3 focal_point = cam.focal_distance_or_xmag * rd;
    const float cam_r1 = rz * TWO_PI;
    const float cam_r2 = rw * aperture;
    float sp, cp;
    sincosf(cam_r1, &sp, &cp);
    const f3 ap = (cp * right + sp * up) * sqrtf(cam_r2);
    dir = normalize(focal_point - ap);
    org = org + ap;
  }
}

}  // namespace hpt
