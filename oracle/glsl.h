// glsl.h — the GLSL builtins the reference shaders use, restated for the CPU
// oracle.  TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load anything under oracle/.
//
// Semantics follow the GLSL 4.60 specification: normalize(v) = v/length(v),
// mix(x,y,a) = x*(1-a)+y*a, clamp(x,l,h) = min(max(x,l),h), reflect/refract
// and smoothstep as in §8.3-8.5.  Expression order is kept exactly as written
// in the shaders (the oracle is built with -ffp-contract=off, so every
// product and sum rounds separately unless fmaf is written out).
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

struct vec2 {
  float x, y;
  vec2() : x(0), y(0) {}
  vec2(float a) : x(a), y(a) {}
  vec2(float a, float b) : x(a), y(b) {}
};

struct vec3 {
  float x, y, z;
  vec3() : x(0), y(0), z(0) {}
  explicit vec3(float a) : x(a), y(a), z(a) {}
  vec3(float a, float b, float c) : x(a), y(b), z(c) {}
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct uvec4 {
  uint32_t x, y, z, w;
};

struct vec4 {
  float x, y, z, w;
  vec4() : x(0), y(0), z(0), w(0) {}
  vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
};

inline vec3 operator+(vec3 a, vec3 b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline vec3 operator-(vec3 a, vec3 b) { return vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline vec3 operator*(vec3 a, vec3 b) { return vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline vec3 operator/(vec3 a, vec3 b) { return vec3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline vec3 operator*(vec3 a, float s) { return vec3(a.x * s, a.y * s, a.z * s); }
inline vec3 operator*(float s, vec3 a) { return vec3(s * a.x, s * a.y, s * a.z); }
inline vec3 operator/(vec3 a, float s) { return vec3(a.x / s, a.y / s, a.z / s); }
inline vec3 operator+(vec3 a, float s) { return vec3(a.x + s, a.y + s, a.z + s); }
inline vec3 operator-(vec3 a, float s) { return vec3(a.x - s, a.y - s, a.z - s); }
inline vec3 operator-(float s, vec3 a) { return vec3(s - a.x, s - a.y, s - a.z); }
inline vec3 operator-(vec3 a) { return vec3(-a.x, -a.y, -a.z); }
inline vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
inline vec3& operator*=(vec3& a, vec3 b) { a = a * b; return a; }
inline vec3& operator*=(vec3& a, float s) { a = a * s; return a; }
inline vec3& operator/=(vec3& a, float s) { a = a / s; return a; }

inline vec2 operator+(vec2 a, vec2 b) { return vec2(a.x + b.x, a.y + b.y); }
inline vec2 operator-(vec2 a, float s) { return vec2(a.x - s, a.y - s); }
inline vec2 operator*(float s, vec2 a) { return vec2(s * a.x, s * a.y); }
inline vec2 operator*(vec2 a, float s) { return vec2(a.x * s, a.y * s); }
inline vec2 operator*(vec2 a, vec2 b) { return vec2(a.x * b.x, a.y * b.y); }
inline vec2 operator/(vec2 a, vec2 b) { return vec2(a.x / b.x, a.y / b.y); }
inline vec2& operator/=(vec2& a, vec2 b) { a = a / b; return a; }

// dot(x,y) = x0*y0 + x1*y1 + x2*y2, summed left to right.
inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b) {
  return vec3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline float length(vec3 a) { return std::sqrt(dot(a, a)); }
inline vec3 normalize(vec3 a) { return a / length(a); }

inline float gmin(float a, float b) { return std::fmin(a, b); }
inline float gmax(float a, float b) { return std::fmax(a, b); }
inline float clampf(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
inline vec3 clamp3(vec3 v, float lo, float hi) {
  return vec3(clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi));
}
inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline vec3 mix3(vec3 x, vec3 y, float a) { return x * (1.0f - a) + y * a; }
inline vec3 exp3(vec3 v) { return vec3(std::exp(v.x), std::exp(v.y), std::exp(v.z)); }
inline vec3 pow3(vec3 v, vec3 e) {
  return vec3(std::pow(v.x, e.x), std::pow(v.y, e.y), std::pow(v.z, e.z));
}
inline float smoothstep(float e0, float e1, float x) {
  float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
// reflect(I, N) = I - 2.0 * dot(N, I) * N
inline vec3 reflect(vec3 I, vec3 N) { return I - 2.0f * dot(N, I) * N; }
// refract(I, N, eta): k = 1 - eta*eta*(1 - dot(N,I)^2); k<0 -> 0,
// else eta*I - (eta*dot(N,I) + sqrt(k))*N
inline vec3 refract(vec3 I, vec3 N, float eta) {
  float d = dot(N, I);
  float k = 1.0f - eta * eta * (1.0f - d * d);
  if (k < 0.0f) return vec3(0.0f);
  return eta * I - (eta * d + std::sqrt(k)) * N;
}

// Constants of common/common.glsl:6-15.  PI is the float literal; the
// derived ones are float expressions folded at compile time.
constexpr float PI = 3.14159265358979323846264338327950288f;
constexpr float FRAC_PI_2 = PI * 0.5f;
constexpr float FRAC_PI_4 = PI * 0.25f;
constexpr float INV_PI = 1.0f / PI;
constexpr float TWO_PI = 2.0f * PI;
constexpr float INV_TWO_PI = 1.0f / TWO_PI;
constexpr float FOUR_PI = 4.0f * PI;
constexpr float INV_FOUR_PI = 1.0f / FOUR_PI;
constexpr float EPS = 0.0003f;
constexpr float INF = 1000000.0f;
constexpr float WORLD_SIZE = 10000.0f;  // inc/constants.glsl:6

}  // namespace oracle
