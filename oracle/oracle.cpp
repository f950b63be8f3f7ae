// oracle.cpp — CPU oracle: scene, ray queries and the path integrator of
// shaders/src/hala-pathtracer/pathtracer.rgen.glsl restated for the CPU.
//
// TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker.
//
// Ray queries.  The reference hands traversal and the ray/triangle test to
// the Vulkan driver (traceRayEXT, pathtracer.rgen.glsl:56-68), whose
// arithmetic no source in the reference fixes.  This oracle fixes it as:
//  * triangles: the watertight test of Woop, Benthin and Wald, "Watertight
//    Ray/Triangle Intersection", JCGT 2(1), 2013, double-sided, with the
//    paper's double-precision fallback when an edge function is 0;
//  * closest hit = the lexicographic minimum of (t, global triangle id) over
//    triangles with t in [tmin, tmax]; an analytic light (lights.rint.glsl)
//    replaces it only when strictly closer; among lights the lowest index wins;
//  * any hit = some triangle or analytic light with t in [tmin, tmax].
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <thread>
#include <vector>

#include "shading.h"

using namespace oracle;

namespace {

constexpr float BOX_MARGIN = 1.0f + 1.0f / 262144.0f;  // 1 + 2^-18 on the far distance

struct Prim {
  float it[9];  // transpose(inverse(mat3(transform))), column-major
  uint32_t material_index;
};

struct BNode {
  float lo[3], hi[3];
  uint32_t left, right;  // internal: children; leaf: right = ~0u
  uint32_t first, count;
};

}  // namespace

struct or_scene {
  std::vector<Camera> cameras;
  std::vector<Light> lights;  // padded with zero entries up to MAX_LIGHTS + 1
  uint32_t num_lights = 0;
  std::vector<Material> materials;
  std::vector<Prim> prims;
  std::vector<hpt_vertex> verts;   // all vertices, primitive-major
  std::vector<vec3> wpos;          // world-space positions
  std::vector<uint32_t> tri_vidx;  // 3 global vertex indices per triangle
  std::vector<uint32_t> tri_prim;
  std::vector<uint32_t> tri_local;
  std::vector<BNode> bvh;
  std::vector<uint32_t> bvh_tris;
  EnvMap env;
  uint32_t env_type = ENV_TYPE_SKY;
  vec3 ground{0.5f, 0.5f, 0.5f}, sky{0.5f, 0.5f, 0.5f};
  float env_rotation = 0.0f, env_intensity = 1.0f, env_total = 0.0f;
  BlueNoise bn;
  // Material textures (g_textures[], inc/global.glsl:48): RGBA8 rows, v = 0 first.
  std::vector<std::vector<uint8_t>> tex;
  std::vector<uint32_t> tex_w, tex_h;
};

namespace {

// ----------------------------------------------------------------- triangles
struct RayPre {
  vec3 org, dir;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

RayPre prepare_ray(vec3 org, vec3 dir) {
  RayPre r;
  r.org = org;
  r.dir = dir;
  const float ax = std::fabs(dir.x), ay = std::fabs(dir.y), az = std::fabs(dir.z);
  r.kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
  r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
  r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
  if (dir[r.kz] < 0.0f) std::swap(r.kx, r.ky);
  r.Sx = dir[r.kx] / dir[r.kz];
  r.Sy = dir[r.ky] / dir[r.kz];
  r.Sz = 1.0f / dir[r.kz];
  return r;
}

// Woop-Benthin-Wald watertight test (double-sided).  Returns false on a miss;
// otherwise t = T/det and the barycentric weights of B and C.
bool tri_test(const RayPre& r, vec3 A0, vec3 B0, vec3 C0, float& t, float& u, float& v) {
  const vec3 A = A0 - r.org, B = B0 - r.org, C = C0 - r.org;
  const float Ax = A[r.kx] - r.Sx * A[r.kz];
  const float Ay = A[r.ky] - r.Sy * A[r.kz];
  const float Bx = B[r.kx] - r.Sx * B[r.kz];
  const float By = B[r.ky] - r.Sy * B[r.kz];
  const float Cx = C[r.kx] - r.Sx * C[r.kz];
  const float Cy = C[r.ky] - r.Sy * C[r.kz];
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = r.Sz * A[r.kz];
  const float Bz = r.Sz * B[r.kz];
  const float Cz = r.Sz * C[r.kz];
  const float T = U * Az + V * Bz + W * Cz;
  t = T / det;
  u = V / det;
  v = W / det;
  return true;
}

struct HitRec {
  float t, u, v;
  uint32_t kind, id;  // id: global triangle or light index
};

inline bool better(float t, uint32_t id, float bt, uint32_t bid) {
  return t < bt || (t == bt && id < bid);
}

vec3 tri_v(const or_scene& s, uint32_t tri, int k) { return s.wpos[s.tri_vidx[tri * 3 + k]]; }

// ---------------------------------------------------------------- own BVH2
inline vec3 inv_dir(vec3 d) {
  auto f = [](float x) {
    float c = std::fabs(x) < 1e-30f ? std::copysign(1e-30f, x) : x;
    return 1.0f / c;
  };
  return vec3(f(d.x), f(d.y), f(d.z));
}

bool box_hit(const float lo[3], const float hi[3], vec3 o, vec3 inv, float tmin, float tmax,
             float& tnear) {
  float n = tmin, f = tmax;
  for (int a = 0; a < 3; ++a) {
    float t0 = (lo[a] - o[a]) * inv[a];
    float t1 = (hi[a] - o[a]) * inv[a];
    if (inv[a] < 0.0f) std::swap(t0, t1);
    n = std::fmax(n, t0);
    f = std::fmin(f, t1);
  }
  tnear = n;
  return n <= f * BOX_MARGIN;
}

void build_bvh2(or_scene& s) {
  const uint32_t n = (uint32_t)s.tri_prim.size();
  s.bvh.clear();
  s.bvh_tris.resize(n);
  for (uint32_t i = 0; i < n; ++i) s.bvh_tris[i] = i;
  if (n == 0) return;
  std::vector<float> cen(n * 3), blo(n * 3), bhi(n * 3);
  for (uint32_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      float x0 = tri_v(s, i, 0)[a], x1 = tri_v(s, i, 1)[a], x2 = tri_v(s, i, 2)[a];
      blo[i * 3 + a] = std::min(x0, std::min(x1, x2));
      bhi[i * 3 + a] = std::max(x0, std::max(x1, x2));
      cen[i * 3 + a] = 0.5f * (blo[i * 3 + a] + bhi[i * 3 + a]);
    }
  }
  struct Item { uint32_t node, first, count; };
  std::vector<Item> work;
  s.bvh.push_back(BNode{});
  work.push_back({0, 0, n});
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    BNode nd{};
    for (int a = 0; a < 3; ++a) { nd.lo[a] = INFINITY; nd.hi[a] = -INFINITY; }
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = it.first; k < it.first + it.count; ++k) {
      uint32_t t = s.bvh_tris[k];
      for (int a = 0; a < 3; ++a) {
        nd.lo[a] = std::min(nd.lo[a], blo[t * 3 + a]);
        nd.hi[a] = std::max(nd.hi[a], bhi[t * 3 + a]);
        clo[a] = std::min(clo[a], cen[t * 3 + a]);
        chi[a] = std::max(chi[a], cen[t * 3 + a]);
      }
    }
    // Pad the box so the float slab test cannot cull a touching triangle.
    for (int a = 0; a < 3; ++a) {
      float pad = 1e-6f * (std::fabs(nd.lo[a]) + std::fabs(nd.hi[a])) + 1e-30f;
      nd.lo[a] -= pad;
      nd.hi[a] += pad;
    }
    if (it.count <= 4) {
      nd.right = ~0u;
      nd.first = it.first;
      nd.count = it.count;
      s.bvh[it.node] = nd;
      continue;
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
      if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    uint32_t mid = it.first + it.count / 2;
    std::nth_element(s.bvh_tris.begin() + it.first, s.bvh_tris.begin() + mid,
                     s.bvh_tris.begin() + it.first + it.count, [&](uint32_t a, uint32_t b) {
                       float ca = cen[a * 3 + axis], cb = cen[b * 3 + axis];
                       return ca < cb || (ca == cb && a < b);
                     });
    nd.left = (uint32_t)s.bvh.size();
    nd.right = nd.left + 1;
    s.bvh[it.node] = nd;
    s.bvh.push_back(BNode{});
    s.bvh.push_back(BNode{});
    work.push_back({nd.left, it.first, mid - it.first});
    work.push_back({nd.right, mid, it.first + it.count - mid});
  }
}

void closest_triangles_bvh2(const or_scene& s, const RayPre& r, float tmin, float tmax,
                            HitRec& h) {
  if (s.bvh.empty()) return;
  const vec3 inv = inv_dir(r.dir);
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode& nd = s.bvh[stack[--sp]];
    float tn;
    float cur = h.kind ? h.t : tmax;
    if (!box_hit(nd.lo, nd.hi, r.org, inv, tmin, cur, tn)) continue;
    if (nd.right == ~0u) {
      for (uint32_t k = nd.first; k < nd.first + nd.count; ++k) {
        uint32_t tri = s.bvh_tris[k];
        float t, u, v;
        if (!tri_test(r, tri_v(s, tri, 0), tri_v(s, tri, 1), tri_v(s, tri, 2), t, u, v)) continue;
        if (!(t >= tmin && t <= tmax)) continue;
        if (h.kind && !better(t, tri, h.t, h.id)) continue;
        h = HitRec{t, u, v, HPT_HIT_TRIANGLE, tri};
      }
    } else {
      float tl, tr;
      float cur2 = h.kind ? h.t : tmax;
      bool hl = box_hit(s.bvh[nd.left].lo, s.bvh[nd.left].hi, r.org, inv, tmin, cur2, tl);
      bool hr = box_hit(s.bvh[nd.right].lo, s.bvh[nd.right].hi, r.org, inv, tmin, cur2, tr);
      if (hl && hr) {
        if (tl <= tr) { stack[sp++] = nd.right; stack[sp++] = nd.left; }
        else { stack[sp++] = nd.left; stack[sp++] = nd.right; }
      } else if (hl) {
        stack[sp++] = nd.left;
      } else if (hr) {
        stack[sp++] = nd.right;
      }
    }
  }
}

bool any_triangle_bvh2(const or_scene& s, const RayPre& r, float tmin, float tmax) {
  if (s.bvh.empty()) return false;
  const vec3 inv = inv_dir(r.dir);
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode& nd = s.bvh[stack[--sp]];
    float tn;
    if (!box_hit(nd.lo, nd.hi, r.org, inv, tmin, tmax, tn)) continue;
    if (nd.right == ~0u) {
      for (uint32_t k = nd.first; k < nd.first + nd.count; ++k) {
        uint32_t tri = s.bvh_tris[k];
        float t, u, v;
        if (tri_test(r, tri_v(s, tri, 0), tri_v(s, tri, 1), tri_v(s, tri, 2), t, u, v) &&
            t >= tmin && t <= tmax)
          return true;
      }
    } else {
      stack[sp++] = nd.left;
      stack[sp++] = nd.right;
    }
  }
  return false;
}

// Analytic lights as lights.rint.glsl reports them.
float light_t(const Light& L, const Ray& ray, vec3& normal, float& pdf) {
  pdf = 1.0f;
  normal = vec3(0.0f);
  if (L.type == LIGHT_TYPE_QUAD) return quad_intersect(L, ray, normal, pdf);
  if (L.type == LIGHT_TYPE_SPHERE) return sphere_light_intersect(L, ray, pdf);
  return INF;
}

void closest_lights(const or_scene& s, const Ray& ray, float tmin, float tmax, HitRec& h) {
  for (uint32_t i = 0; i < s.num_lights; ++i) {
    vec3 nrm;
    float pdf;
    float t = light_t(s.lights[i], ray, nrm, pdf);
    if (!(t >= tmin && t <= tmax)) continue;
    if (h.kind && !(t < h.t)) continue;
    h = HitRec{t, 0.0f, 0.0f, HPT_HIT_LIGHT, i};
  }
}

bool any_light(const or_scene& s, const Ray& ray, float tmin, float tmax) {
  for (uint32_t i = 0; i < s.num_lights; ++i) {
    vec3 nrm;
    float pdf;
    float t = light_t(s.lights[i], ray, nrm, pdf);
    if (t >= tmin && t <= tmax) return true;
  }
  return false;
}

HitRec closest(const or_scene& s, const Ray& ray, float tmin, float tmax) {
  HitRec h{0, 0, 0, HPT_HIT_NONE, 0};
  RayPre r = prepare_ray(ray.origin, ray.direction);
  closest_triangles_bvh2(s, r, tmin, tmax, h);
  closest_lights(s, ray, tmin, tmax, h);
  return h;
}

bool occluded(const or_scene& s, const Ray& ray, float tmin, float tmax) {
  RayPre r = prepare_ray(ray.origin, ray.direction);
  if (any_triangle_bvh2(s, r, tmin, tmax)) return true;
  return any_light(s, ray, tmin, tmax);
}

void to_hpt_hit(const or_scene& s, const HitRec& h, hpt_hit& o) {
  o.t = h.t;
  o.u = h.u;
  o.v = h.v;
  o.kind = h.kind;
  if (h.kind == HPT_HIT_TRIANGLE) {
    o.instance = s.tri_prim[h.id];
    o.primitive = s.tri_local[h.id];
  } else if (h.kind == HPT_HIT_LIGHT) {
    o.instance = h.id;
    o.primitive = INVALID_INDEX;
  } else {
    o.t = 0.0f;
    o.u = o.v = 0.0f;
    o.instance = o.primitive = INVALID_INDEX;
  }
}

Ray to_ray(const hpt_ray& r) {
  return Ray{vec3(r.origin[0], r.origin[1], r.origin[2]),
             vec3(r.direction[0], r.direction[1], r.direction[2])};
}

// ------------------------------------------------------- hit shaders
// triangles.rchit.glsl:13-76 and lights.rchit.glsl:13-26 writing into State;
// pathtracer.rmiss.glsl:12-14 clearing RAY_FLAGS_HIT on a miss.
void trace_payload(const or_scene& s, const Ray& r, State& st) {
  HitRec h = closest(s, r, 0.0f, WORLD_SIZE);
  if (h.kind == HPT_HIT_NONE) {
    st.flags &= ~RAY_FLAGS_HIT;
    return;
  }
  if (h.kind == HPT_HIT_LIGHT) {
    vec3 nrm;
    float pdf;
    light_t(s.lights[h.id], r, nrm, pdf);
    st.flags |= (RAY_FLAGS_HIT | RAY_FLAGS_IS_EMITTER);
    st.material_index = h.id;
    st.hit_distance = h.t;
    st.first_hit_position = r.origin + r.direction * st.hit_distance;
    st.normal = nrm;
    st.ffnormal = dot(st.normal, r.direction) <= 0.0f ? st.normal : -st.normal;
    st.tangent = vec3(0.0f);
    st.bitangent = vec3(0.0f);
    st.tex_coord = vec2(pdf, 0.0f);
    return;
  }
  const Prim& prim = s.prims[s.tri_prim[h.id]];
  const hpt_vertex& v0 = s.verts[s.tri_vidx[h.id * 3 + 0]];
  const hpt_vertex& v1 = s.verts[s.tri_vidx[h.id * 3 + 1]];
  const hpt_vertex& v2 = s.verts[s.tri_vidx[h.id * 3 + 2]];
  auto n3 = [](const float* p) { return vec3(p[0], p[1], p[2]); };
  const vec3 bary = vec3(1.0f - h.u - h.v, h.u, h.v);
  const vec3 c0(prim.it[0], prim.it[1], prim.it[2]), c1(prim.it[3], prim.it[4], prim.it[5]),
      c2(prim.it[6], prim.it[7], prim.it[8]);
  auto mat3mul = [&](vec3 x) { return c0 * x.x + c1 * x.y + c2 * x.z; };
  vec3 normal = n3(v0.normal) * bary.x + n3(v1.normal) * bary.y + n3(v2.normal) * bary.z;
  normal = normalize(mat3mul(normal));
  vec3 tangent = n3(v0.tangent) * bary.x + n3(v1.tangent) * bary.y + n3(v2.tangent) * bary.z;
  tangent = normalize(mat3mul(tangent));
  const vec3 bitangent = cross(normal, tangent);
  const vec2 tex_coord = vec2(v0.tex_coord[0], v0.tex_coord[1]) * bary.x +
                         vec2(v1.tex_coord[0], v1.tex_coord[1]) * bary.y +
                         vec2(v2.tex_coord[0], v2.tex_coord[1]) * bary.z;
  st.flags |= RAY_FLAGS_HIT;
  st.material_index = prim.material_index;
  st.hit_distance = h.t;
  st.first_hit_position = r.origin + r.direction * st.hit_distance;
  st.normal = normal;
  st.ffnormal = dot(normal, r.direction) <= 0.0f ? normal : -normal;
  st.tangent = tangent;
  st.bitangent = bitangent;
  st.tex_coord = tex_coord;
}

// ------------------------------------------------------------ textures
// 8-bit sRGB code -> linear (IEC 61966-2-1 in double, rounded to float).
float srgb_code_to_linear(int code) {
  const double c = code / 255.0;
  return (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
}

// textureLod(g_textures[t], uv, 0): bilinear, REPEAT in u and v, texel
// centres at (i + 0.5) / W; colour maps decode sRGB per texel before
// filtering, data maps read c / 255.  (The sampler state lives in the
// unvendored renderer; this is the build's stated choice.)
vec4 texture_lod(const or_scene& s, uint32_t t, vec2 uv, bool srgb) {
  const int w = (int)s.tex_w[t], h = (int)s.tex_h[t];
  const float x = uv.x * (float)w - 0.5f;
  const float y = uv.y * (float)h - 0.5f;
  const float fx0 = std::floor(x), fy0 = std::floor(y);
  const float fx = x - fx0, fy = y - fy0;
  int i0 = (int)fx0 % w;
  if (i0 < 0) i0 += w;
  int j0 = (int)fy0 % h;
  if (j0 < 0) j0 += h;
  const int i1 = i0 + 1 == w ? 0 : i0 + 1;
  const int j1 = j0 + 1 == h ? 0 : j0 + 1;
  auto texel = [&](int i, int j) {
    const uint8_t* c = s.tex[t].data() + ((size_t)j * w + i) * 4;
    if (srgb)
      return vec4(srgb_code_to_linear(c[0]), srgb_code_to_linear(c[1]), srgb_code_to_linear(c[2]),
                  (float)c[3] / 255.0f);
    return vec4((float)c[0] / 255.0f, (float)c[1] / 255.0f, (float)c[2] / 255.0f, (float)c[3] / 255.0f);
  };
  auto mix4 = [](vec4 p, vec4 q, float a) {
    return vec4(mixf(p.x, q.x, a), mixf(p.y, q.y, a), mixf(p.z, q.z, a), mixf(p.w, q.w, a));
  };
  return mix4(mix4(texel(i0, j0), texel(i1, j0), fx), mix4(texel(i0, j1), texel(i1, j1), fx), fy);
}

// pathtracer.rgen.glsl:121-146: base colour, metallic (b) / roughness (g),
// emission and the normal map (z * 5 before the remap; the bitangent
// follows the new normal).
void apply_texture_maps(const or_scene& s, vec3 dir, Material& mat, State& st) {
  if (mat.base_color_map_index != INVALID_INDEX) {
    const vec4 c = texture_lod(s, mat.base_color_map_index, st.tex_coord, true);
    mat.base_color = vec3(c.x, c.y, c.z);
  }
  if (mat.metallic_roughness_map_index != INVALID_INDEX) {
    const vec4 mr = texture_lod(s, mat.metallic_roughness_map_index, st.tex_coord, false);
    mat.metallic = mr.z;
    mat.roughness = mr.y;
  }
  if (mat.emission_map_index != INVALID_INDEX) {
    const vec4 c = texture_lod(s, mat.emission_map_index, st.tex_coord, true);
    mat.emission = vec3(c.x, c.y, c.z);
  }
  if (mat.normal_map_index != INVALID_INDEX) {
    const vec4 c = texture_lod(s, mat.normal_map_index, st.tex_coord, false);
    vec3 nts = vec3(c.x, c.y, c.z);
    nts.z *= 5.0f;
    nts = normalize(nts * 2.0f - 1.0f);
    const vec3 original = st.normal;
    st.normal = normalize(st.tangent * nts.x + st.bitangent * nts.y + st.normal * nts.z);
    st.ffnormal = dot(original, dir) <= 0.0f ? st.normal : -st.normal;
    st.bitangent = cross(st.normal, st.tangent);
  }
}

// ------------------------------------------------------- integrator
struct Variant {
  bool medium, vol_mis, transparent;
};

struct Payload {  // g_ray_payload (inc/rayload.glsl:1)
  RNGState rng;
  State state;
  Ray ray;
};

Globals make_globals(const or_scene& s, const or_settings& st, uint32_t frame) {
  Globals g;
  g.ground_color = s.ground;
  g.sky_color = s.sky;
  g.resolution = vec2((float)st.width, (float)st.height);
  g.max_depth = st.max_depth;
  g.rr_depth = st.rr_depth;
  g.frame_index = frame;
  g.camera_index = st.camera_index;
  g.env_type = s.env_type;
  g.env_map_width = s.env.width;
  g.env_map_height = s.env.height;
  g.env_total_sum = s.env_total;
  g.env_rotation = s.env_rotation;
  g.env_intensity = s.env_intensity;
  g.exposure_value = st.exposure_value;
  g.enable_tonemap = st.enable_tonemap != 0;
  g.enable_aces = st.enable_aces != 0;
  g.use_simple_aces = st.use_simple_aces != 0;
  g.num_of_lights = s.num_lights;
  return g;
}

// inc/lighting.glsl:7-61 (MEDIUM + VOL_MIS).  Transparency draws come from
// the main path's RNG, as in the reference (:42).
vec3 eval_transmittance(const or_scene& s, const Globals& g, const Variant& var, Payload& pl,
                        Ray ray) {
  Payload tp;
  tp.ray = ray;
  tp.state.flags = 0;
  tp.state.pdf = 1.0f;
  tp.state.eta = 1.0f;
  tp.state.material_index = INVALID_INDEX;
  tp.state.medium.type = MEDIUM_NONE;
  tp.state.medium.density = 0.0f;
  tp.state.medium.anisotropy = 1.0f;
  vec3 transmittance = vec3(1.0f);
  for (uint32_t depth = 0; depth < g.max_depth; depth++) {
    trace_payload(s, tp.ray, tp.state);
    if ((tp.state.flags & RAY_FLAGS_HIT) == 0 || (tp.state.flags & RAY_FLAGS_IS_EMITTER) != 0) break;
    if (tp.state.material_index == INVALID_INDEX) return vec3(0.0f);
    const Material& mat = s.materials[tp.state.material_index];
    bool is_transparent = var.transparent ? (mat.opacity == 0.0f || rand(pl.rng) > mat.opacity) : false;
    bool is_refractive = (1.0f - mat.metallic) * mat.specular_transmission > 0.0f;
    if (!(is_transparent || is_refractive)) return vec3(0.0f);
    if (dot(tp.ray.direction, tp.state.normal) > 0 && mat.medium.type != MEDIUM_NONE) {
      vec3 color = mat.medium.type == MEDIUM_ABSORB ? vec3(1.0f) - mat.medium.color : vec3(1.0f);
      transmittance *= exp3(-color * mat.medium.density * tp.state.hit_distance);
    }
    tp.ray.origin = tp.state.first_hit_position + tp.ray.direction * EPS;
  }
  return transmittance;
}

// inc/lighting.glsl:69-244.
vec3 direct_lighting(const or_scene& s, const Globals& g, const Variant& var, Payload& pl,
                     const Material& mat, bool is_surface, bool anb) {
  vec3 Ld = vec3(0.0f);
  vec3 Li = vec3(0.0f);
  const vec3 scatter_pos = pl.state.first_hit_position + pl.state.ffnormal * EPS;
  const bool vmis = var.medium && var.vol_mis;

  if (g.env_type != ENV_TYPE_SKY) {
    vec3 color;
    float light_pdf;
    const vec3 light_dir = sample_env_map(pl.rng, g, s.env, color, light_pdf);
    Li = color;
    const Ray shadow_ray{scatter_pos, light_dir};
    if (light_pdf > 0.0f) {
      if (vmis) {
        const vec3 transmittance = eval_transmittance(s, g, var, pl, shadow_ray);
        if (luminance(transmittance) > 0.0f) {
          Li *= transmittance;
          float pdf = 0.0f;
          vec3 f = vec3(0.0f);
          if (is_surface) {
            f = bxdf_eval(anb, pl.state, mat, -pl.ray.direction, pl.state.ffnormal, light_dir, pdf);
          } else {
            pdf = phase_hg(dot(-pl.ray.direction, light_dir), pl.state.medium.anisotropy);
            f = vec3(pdf);
          }
          if (pdf > 0.0f) {
            const float mis_weight = power_heuristic(light_pdf, pdf);
            if (mis_weight > 0.0f) Ld += mis_weight * Li * f / light_pdf;
          }
        }
      } else {
        if (!occluded(s, shadow_ray, 0.0f, WORLD_SIZE)) {
          float pdf;
          const vec3 f = bxdf_eval(anb, pl.state, mat, -pl.ray.direction, pl.state.ffnormal, light_dir, pdf);
          if (pdf > 0.0f) {
            const float mis_weight = power_heuristic(light_pdf, pdf);
            if (mis_weight > 0.0f) Ld += mis_weight * Li * f / light_pdf;
          }
        }
      }
    }
  }

  if (g.num_of_lights > 0) {
    const int light_index = (int)(rand(pl.rng) * (float)g.num_of_lights);
    // rand() can return 1.0 (common/wltcRS.glsl:30), giving light_index ==
    // num_of_lights: the reference then reads the zero-filled slot after the
    // last light (a black point light at the origin); s.lights keeps that slot.
    const Light& light = s.lights[light_index];
    LightSample ls;
    light_sample(pl.rng, g, pl.state, light, ls);
    const vec3 light_dir = ls.direction;
    const Ray shadow_ray{scatter_pos, light_dir};
    const float light_pdf = ls.pdf;
    Li = ls.emission;
    if (light_pdf > 0.0f && dot(light_dir, ls.normal) < 0.0f) {
      if (vmis) {
        const vec3 transmittance = eval_transmittance(s, g, var, pl, shadow_ray);
        if (luminance(transmittance) > 0.0f) {
          Li *= transmittance;
          float pdf = 0.0f;
          vec3 f = vec3(0.0f);
          if (is_surface) {
            f = bxdf_eval(anb, pl.state, mat, -pl.ray.direction, pl.state.ffnormal, light_dir, pdf);
          } else {
            pdf = phase_hg(dot(-pl.ray.direction, light_dir), pl.state.medium.anisotropy);
            f = vec3(pdf);
          }
          if (pdf > 0.0f) {
            float mis_weight = 1.0f;
            if (light.area > 0.0f) mis_weight = power_heuristic(light_pdf, pdf);
            if (mis_weight > 0.0f) Ld += mis_weight * Li * f / light_pdf;
          }
        }
      } else {
        if (!occluded(s, shadow_ray, 0.0f, ls.dist - 2.0f * EPS)) {
          float pdf;
          const vec3 f = bxdf_eval(anb, pl.state, mat, -pl.ray.direction, pl.state.ffnormal, light_dir, pdf);
          if (pdf > 0.0f) {
            float mis_weight = 1.0f;
            if (light.area > 0.0f) mis_weight = power_heuristic(light_pdf, pdf);
            if (mis_weight > 0.0f) Ld += mis_weight * Li * f / light_pdf;
          }
        }
      }
    }
  }
  return Ld;
}

struct PathOut {
  vec3 radiance, albedo, normal;
};

// pathtracer.rgen.glsl:22-257.
PathOut trace_path(const or_scene& s, const Globals& g, const Variant& var, uint32_t x, uint32_t y) {
  Payload pl;
  pl.rng = rng_initialize(x, y, g.frame_index);
  const Camera& camera = s.cameras[g.camera_index];
  pl.ray = camera_ray(pl.rng, g, camera, s.bn, x, y);
  pl.state.flags = 0;
  pl.state.pdf = 1.0f;
  pl.state.eta = 1.0f;
  pl.state.material_index = INVALID_INDEX;
  pl.state.medium.type = MEDIUM_NONE;
  pl.state.medium.density = 0.0f;
  pl.state.medium.anisotropy = 1.0f;

  vec3 radiance = vec3(0.0f);
  vec3 throughput = vec3(1.0f);
  bool is_in_medium = false;
  bool is_medium_sampled = false;
  bool is_surface_scatter = false;  // only read when MEDIUM && !VOL_MIS
  const bool surface_flag = var.medium && !var.vol_mis;
  bool any_non_specular_bounce = false;
  vec3 albedo = vec3(0.0f);
  vec3 normal = vec3(0.0f, 0.0f, 1.0f);
  for (pl.state.depth = 0;; pl.state.depth++) {
    trace_payload(s, pl.ray, pl.state);

    if ((pl.state.flags & RAY_FLAGS_HIT) == 0) {
      const vec3 env = env_eval(g, s.env, pl.state.flags, pl.state.depth, pl.state.pdf,
                                pl.ray.direction, surface_flag, is_surface_scatter);
      radiance += throughput * env;
      if (pl.state.depth == 0) {
        albedo = radiance;
        normal = -pl.ray.direction;
      }
      break;
    }

    if ((pl.state.flags & RAY_FLAGS_IS_EMITTER) != 0) {
      const Light& light = s.lights[pl.state.material_index];
      const float light_pdf = pl.state.tex_coord.x;
      float mis_weight = 1.0f;
      if (pl.state.depth > 0) mis_weight = power_heuristic(pl.state.pdf, light_pdf);
      if (surface_flag && !is_surface_scatter) mis_weight = 1.0f;
      radiance += mis_weight * light.intensity * throughput;
      break;
    }

    if (pl.state.material_index == INVALID_INDEX) {
      radiance = vec3(0.7f, 0.0f, 0.7f);
      break;
    }

    Material mat = s.materials[pl.state.material_index];
    pl.state.eta = dot(pl.ray.direction, pl.state.normal) < 0.0f ? (1.0f / mat.ior) : mat.ior;
    apply_texture_maps(s, pl.ray.direction, mat, pl.state);  // pathtracer.rgen.glsl:121-146

    if (pl.state.depth == 0) {
      albedo = mat.base_color;
      normal = pl.state.normal;
    }

    radiance += throughput * mat.emission;

    if (var.medium) {
      is_medium_sampled = false;
      if (!var.vol_mis) is_surface_scatter = false;
      if (is_in_medium) {
        if (pl.state.medium.type == MEDIUM_ABSORB) {
          throughput *= exp3(-(1.0f - pl.state.medium.color) * pl.state.hit_distance * pl.state.medium.density);
        } else if (pl.state.medium.type == MEDIUM_EMISSIVE) {
          radiance += pl.state.medium.color * pl.state.hit_distance * pl.state.medium.density * throughput;
        } else {
          const float scatter_dist = gmin(-std::log(rand(pl.rng)) / pl.state.medium.density, pl.state.hit_distance);
          is_medium_sampled = scatter_dist < pl.state.hit_distance;
          if (is_medium_sampled) {
            throughput *= pl.state.medium.color;
            pl.ray.origin += pl.ray.direction * scatter_dist;
            pl.state.first_hit_position = pl.ray.origin;
            radiance += direct_lighting(s, g, var, pl, mat, false, any_non_specular_bounce) * throughput;
            const float r1 = rand(pl.rng);
            const float r2 = rand(pl.rng);
            const vec3 scatter_dir = sample_hg(-pl.ray.direction, pl.state.medium.anisotropy, r1, r2);
            pl.state.pdf = phase_hg(dot(-pl.ray.direction, scatter_dir), pl.state.medium.anisotropy);
            pl.ray.direction = scatter_dir;
          }
        }
      }
    }

    if (!is_medium_sampled) {
      bool pass_through = false;
      if (var.transparent) {
        if (mat.opacity == 0.0f || rand(pl.rng) > mat.opacity) {
          // depth is a uint: at depth 0 this wraps to 0xFFFFFFFF and the
          // max_depth test below ends the path (pathtracer.rgen.glsl:199,247).
          pl.state.depth--;
          pass_through = true;
        }
      }
      if (!pass_through) {
        if (surface_flag) is_surface_scatter = true;
        radiance += throughput * direct_lighting(s, g, var, pl, mat, true, any_non_specular_bounce);
        vec3 L;
        float pdf;
        uint32_t flags = 0;
        const vec3 f = bxdf_sample(pl.rng, any_non_specular_bounce, pl.state, mat, -pl.ray.direction,
                                   pl.state.ffnormal, L, pdf, flags);
        pl.ray.direction = L;
        pl.state.pdf = pdf;
        pl.state.flags |= flags;
        if (pl.state.pdf > 0.0f) {
          throughput *= f / pl.state.pdf;
        } else {
          break;
        }
      }
      pl.ray.origin = pl.state.first_hit_position + pl.ray.direction * EPS;
      if (var.medium) {
        if (dot(pl.ray.direction, pl.state.normal) < 0 && mat.medium.type != MEDIUM_NONE) {
          is_in_medium = true;
          pl.state.medium = mat.medium;
        } else {
          is_in_medium = false;
          pl.state.medium.type = MEDIUM_NONE;
        }
      }
    }

    if (pl.state.depth >= g.max_depth) break;

    if (pl.state.depth >= g.rr_depth && g.rr_depth > 0) {
      float q = gmin(gmax(throughput.x, gmax(throughput.y, throughput.z)) + 0.001f, 0.95f);
      if (rand(pl.rng) > q) break;
      throughput /= q;
    }
  }
  return PathOut{radiance, albedo, normal};
}

Variant variant_of(uint32_t features) {
  Variant v;
  v.medium = (features & HPT_FEATURE_MEDIUM) != 0;
  v.vol_mis = v.medium && (features & HPT_FEATURE_VOL_MIS) != 0;
  v.transparent = (features & HPT_FEATURE_TRANSPARENT) != 0;
  return v;
}

vec3 final_color(const Globals& g, vec3 c) {  // pathtracer.rgen.glsl:280-301
  vec3 final_color = c;
  float exposure = std::pow(2.0f, g.exposure_value);
  final_color *= exposure;
  if (g.enable_tonemap) {
    if (g.enable_aces) {
      if (g.use_simple_aces) final_color = aces(final_color);
      else final_color = aces_fitted(final_color);
    } else {
      final_color = tonemap(final_color, 1.5f);
    }
  }
  return linear_2_srgb(final_color);
}

uint8_t to_unorm8(float x) {  // rgba8 image store: clamp, scale, round to nearest
  float c = clampf(x, 0.0f, 1.0f);
  return (uint8_t)std::floor(c * 255.0f + 0.5f);
}

template <typename F>
void parallel_rows(uint32_t y0, uint32_t y1, int threads, F&& fn) {
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::vector<std::thread> pool;
  std::atomic<uint32_t> next{y0};
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&]() {
      for (;;) {
        uint32_t y = next.fetch_add(1);
        if (y >= y1) break;
        fn(y);
      }
    });
  }
  for (auto& th : pool) th.join();
}

void fill_default_blue_noise(BlueNoise& bn) {
  // Procedural stand-in for assets/textures/blue_noise.png (absent from the
  // reference checkout): texel (x,y) = top bytes of pcg4d(x, y, c0, c1).
  bn.width = bn.height = 1024;
  bn.rgba.resize(1024 * 1024 * 4);
  for (uint32_t y = 0; y < 1024; ++y)
    for (uint32_t x = 0; x < 1024; ++x) {
      uvec4 v{x, y, 0x6A09E667u, 0xBB67AE85u};
      pcg4d(v);
      uint8_t* p = &bn.rgba[((size_t)y * 1024 + x) * 4];
      p[0] = (uint8_t)(v.x >> 24);
      p[1] = (uint8_t)(v.y >> 24);
      p[2] = (uint8_t)(v.z >> 24);
      p[3] = (uint8_t)(v.w >> 24);
    }
}

Material to_material(const hpt_material& m) {
  Material o;
  o.medium.color = vec3(m.medium.color[0], m.medium.color[1], m.medium.color[2]);
  o.medium.density = m.medium.density;
  o.medium.anisotropy = m.medium.anisotropy;
  o.medium.type = m.medium.type;
  o.base_color = vec3(m.base_color[0], m.base_color[1], m.base_color[2]);
  o.opacity = m.opacity;
  o.emission = vec3(m.emission[0], m.emission[1], m.emission[2]);
  o.anisotropic = m.anisotropic;
  o.metallic = m.metallic;
  o.roughness = m.roughness;
  o.subsurface = m.subsurface;
  o.specular_tint = m.specular_tint;
  o.sheen = m.sheen;
  o.sheen_tint = m.sheen_tint;
  o.clearcoat = m.clearcoat;
  o.clearcoat_roughness = m.clearcoat_roughness;
  o.clearcoat_tint = vec3(m.clearcoat_tint[0], m.clearcoat_tint[1], m.clearcoat_tint[2]);
  o.specular_transmission = m.specular_transmission;
  o.ior = m.ior;
  o.ax = m.ax;
  o.ay = m.ay;
  o.base_color_map_index = m.base_color_map_index;
  o.normal_map_index = m.normal_map_index;
  o.metallic_roughness_map_index = m.metallic_roughness_map_index;
  o.emission_map_index = m.emission_map_index;
  o.type = m.type;
  return o;
}

Light to_light(const hpt_light& l) {
  Light o;
  o.intensity = vec3(l.intensity[0], l.intensity[1], l.intensity[2]);
  o.position = vec3(l.position[0], l.position[1], l.position[2]);
  o.u = vec3(l.u[0], l.u[1], l.u[2]);
  o.v = vec3(l.v[0], l.v[1], l.v[2]);
  o.radius = l.radius;
  o.area = l.area;
  o.type = l.type;
  return o;
}

State frame_state(const or_frame* fr) {
  State st;
  st.tangent = vec3(fr->tangent[0], fr->tangent[1], fr->tangent[2]);
  st.bitangent = vec3(fr->bitangent[0], fr->bitangent[1], fr->bitangent[2]);
  st.eta = fr->eta;
  st.depth = fr->depth;
  return st;
}

inline vec3 V3(const float* p) { return vec3(p[0], p[1], p[2]); }
inline void S3(float* o, vec3 v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }

}  // namespace

// =================================================================== C API
extern "C" {

or_scene* or_scene_create(const hpt_scene_desc* d) {
  if (!d) return nullptr;
  auto* s = new or_scene();
  for (uint32_t i = 0; i < d->camera_count; ++i) {
    const hpt_camera& c = d->cameras[i];
    s->cameras.push_back(Camera{V3(c.position), V3(c.right), V3(c.up), V3(c.forward), c.yfov,
                                c.focal_distance_or_xmag, c.aperture_or_ymag, c.type});
  }
  s->num_lights = d->light_count;
  s->lights.assign(HPT_MAX_LIGHTS + 1, Light{});
  for (uint32_t i = 0; i < d->light_count; ++i) s->lights[i] = to_light(d->lights[i]);
  for (uint32_t i = 0; i < d->material_count; ++i) s->materials.push_back(to_material(d->materials[i]));
  for (uint32_t p = 0; p < d->primitive_count; ++p) {
    const hpt_primitive_desc& pd = d->primitives[p];
    const float* m = pd.transform;
    // transpose(inverse(mat3(transform))) (triangles.rchit.glsl:21): for
    // columns a, b, c its columns are cross(b,c), cross(c,a), cross(a,b) / det.
    vec3 a(m[0], m[1], m[2]), b(m[4], m[5], m[6]), c(m[8], m[9], m[10]);
    float det = dot(a, cross(b, c));
    vec3 i0 = cross(b, c) / det, i1 = cross(c, a) / det, i2 = cross(a, b) / det;
    Prim pr;
    float it[9] = {i0.x, i0.y, i0.z, i1.x, i1.y, i1.z, i2.x, i2.y, i2.z};
    std::memcpy(pr.it, it, sizeof it);
    pr.material_index = pd.material_index;
    s->prims.push_back(pr);
    uint32_t vbase = (uint32_t)s->verts.size();
    for (uint32_t k = 0; k < pd.vertex_count; ++k) {
      const hpt_vertex& v = pd.vertices[k];
      s->verts.push_back(v);
      const float x = v.position[0], y = v.position[1], z = v.position[2];
      s->wpos.push_back(vec3(((m[0] * x + m[4] * y) + m[8] * z) + m[12],
                             ((m[1] * x + m[5] * y) + m[9] * z) + m[13],
                             ((m[2] * x + m[6] * y) + m[10] * z) + m[14]));
    }
    for (uint32_t t = 0; t < pd.index_count / 3; ++t) {
      for (int k = 0; k < 3; ++k) s->tri_vidx.push_back(vbase + pd.indices[t * 3 + k]);
      s->tri_prim.push_back(p);
      s->tri_local.push_back(t);
    }
  }
  build_bvh2(*s);
  fill_default_blue_noise(s->bn);
  return s;
}

void or_scene_destroy(or_scene* s) { delete s; }

int or_scene_set_env_map(or_scene* s, const float* rgb, uint32_t w, uint32_t h, float rotation) {
  if (!s || !rgb || !w || !h) return 1;
  EnvMap& e = s->env;
  e.width = w;
  e.height = h;
  e.rgb.assign(rgb, rgb + (size_t)w * h * 3);
  e.lum.resize((size_t)w * h);
  for (size_t i = 0; i < (size_t)w * h; ++i) e.lum[i] = luminance(vec3(rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2]));
  e.prob.resize((size_t)w * h);
  e.alias.resize((size_t)w * h);
  double total = build_env_alias(e.lum.data(), w * h, e.prob.data(), e.alias.data());
  s->env_total = (float)total;
  s->env_rotation = rotation;
  s->env_type = ENV_TYPE_MAP;
  return 0;
}

void or_scene_set_env_sky(or_scene* s, const float ground[3], const float sky[3]) {
  s->ground = V3(ground);
  s->sky = V3(sky);
  s->env_type = ENV_TYPE_SKY;
}

void or_scene_set_env_intensity(or_scene* s, float intensity) { s->env_intensity = intensity; }

int or_scene_set_blue_noise(or_scene* s, const uint8_t* rgba8, uint32_t w, uint32_t h) {
  if (!s || !rgba8 || w != 1024 || h != 1024) return 1;
  s->bn.rgba.assign(rgba8, rgba8 + (size_t)w * h * 4);
  return 0;
}

int or_scene_set_textures(or_scene* s, const hpt_texture_desc* t, uint32_t count) {
  if (!s || (count && !t)) return 1;
  s->tex.clear();
  s->tex_w.clear();
  s->tex_h.clear();
  for (uint32_t i = 0; i < count; ++i) {
    if (!t[i].rgba8 || !t[i].width || !t[i].height) return 1;
    s->tex.emplace_back(t[i].rgba8, t[i].rgba8 + (size_t)t[i].width * t[i].height * 4);
    s->tex_w.push_back(t[i].width);
    s->tex_h.push_back(t[i].height);
  }
  return 0;
}

float or_texture_sample(const or_scene* s, uint32_t t, float u, float v, int srgb, float* rgba) {
  const vec4 c = texture_lod(*s, t, vec2(u, v), srgb != 0);
  rgba[0] = c.x;
  rgba[1] = c.y;
  rgba[2] = c.z;
  rgba[3] = c.w;
  return c.x;
}

uint32_t or_scene_triangle_count(const or_scene* s) { return (uint32_t)s->tri_prim.size(); }

void or_trace_closest_brute(const or_scene* s, const hpt_ray* rays, uint32_t n, hpt_hit* hits) {
  for (uint32_t i = 0; i < n; ++i) {
    const hpt_ray& hr = rays[i];
    Ray ray = to_ray(hr);
    RayPre r = prepare_ray(ray.origin, ray.direction);
    HitRec h{0, 0, 0, HPT_HIT_NONE, 0};
    for (uint32_t tri = 0; tri < (uint32_t)s->tri_prim.size(); ++tri) {
      float t, u, v;
      if (!tri_test(r, tri_v(*s, tri, 0), tri_v(*s, tri, 1), tri_v(*s, tri, 2), t, u, v)) continue;
      if (!(t >= hr.tmin && t <= hr.tmax)) continue;
      if (h.kind && !better(t, tri, h.t, h.id)) continue;
      h = HitRec{t, u, v, HPT_HIT_TRIANGLE, tri};
    }
    closest_lights(*s, ray, hr.tmin, hr.tmax, h);
    to_hpt_hit(*s, h, hits[i]);
  }
}

void or_trace_closest(const or_scene* s, const hpt_ray* rays, uint32_t n, hpt_hit* hits) {
  for (uint32_t i = 0; i < n; ++i) {
    HitRec h = closest(*s, to_ray(rays[i]), rays[i].tmin, rays[i].tmax);
    to_hpt_hit(*s, h, hits[i]);
  }
}

void or_trace_any(const or_scene* s, const hpt_ray* rays, uint32_t n, uint32_t* occ) {
  for (uint32_t i = 0; i < n; ++i) occ[i] = occluded(*s, to_ray(rays[i]), rays[i].tmin, rays[i].tmax) ? 1u : 0u;
}

void or_trace_any_brute(const or_scene* s, const hpt_ray* rays, uint32_t n, uint32_t* occ) {
  for (uint32_t i = 0; i < n; ++i) {
    const hpt_ray& hr = rays[i];
    Ray ray = to_ray(hr);
    RayPre r = prepare_ray(ray.origin, ray.direction);
    bool hit = false;
    for (uint32_t tri = 0; tri < (uint32_t)s->tri_prim.size() && !hit; ++tri) {
      float t, u, v;
      if (tri_test(r, tri_v(*s, tri, 0), tri_v(*s, tri, 1), tri_v(*s, tri, 2), t, u, v) &&
          t >= hr.tmin && t <= hr.tmax)
        hit = true;
    }
    if (!hit) hit = any_light(*s, ray, hr.tmin, hr.tmax);
    occ[i] = hit ? 1u : 0u;
  }
}

void or_trace_closest_fp64(const or_scene* s, const hpt_ray* rays, uint32_t n, double* tout,
                           uint32_t* gtri, double* gap) {
  // Plain double-precision Moller-Trumbore over all triangles: an
  // independent check of which triangle is closest and by how much.
  for (uint32_t i = 0; i < n; ++i) {
    const hpt_ray& hr = rays[i];
    double o[3] = {hr.origin[0], hr.origin[1], hr.origin[2]};
    double d[3] = {hr.direction[0], hr.direction[1], hr.direction[2]};
    double best = INFINITY, second = INFINITY;
    uint32_t bid = INVALID_INDEX;
    for (uint32_t tri = 0; tri < (uint32_t)s->tri_prim.size(); ++tri) {
      double p[3][3];
      for (int k = 0; k < 3; ++k) {
        vec3 q = tri_v(*s, tri, k);
        p[k][0] = q.x; p[k][1] = q.y; p[k][2] = q.z;
      }
      double e1[3], e2[3], pv[3], tv[3], qv[3];
      for (int a = 0; a < 3; ++a) { e1[a] = p[1][a] - p[0][a]; e2[a] = p[2][a] - p[0][a]; tv[a] = o[a] - p[0][a]; }
      pv[0] = d[1] * e2[2] - d[2] * e2[1]; pv[1] = d[2] * e2[0] - d[0] * e2[2]; pv[2] = d[0] * e2[1] - d[1] * e2[0];
      double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
      if (det == 0.0) continue;
      double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) / det;
      if (u < 0.0 || u > 1.0) continue;
      qv[0] = tv[1] * e1[2] - tv[2] * e1[1]; qv[1] = tv[2] * e1[0] - tv[0] * e1[2]; qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
      double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) / det;
      if (v < 0.0 || u + v > 1.0) continue;
      double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) / det;
      if (t < hr.tmin || t > hr.tmax) continue;
      if (t < best) { second = best; best = t; bid = tri; }
      else if (t < second) second = t;
    }
    tout[i] = best;
    gtri[i] = bid;
    gap[i] = second - best;
  }
}

// The 8-wide node walk of the product, restated for counting.
struct Node8 {
  float p[3];
  uint8_t e[3];
  uint8_t imask;
  uint32_t child_base, tri_base;
  uint8_t meta[8];
  uint8_t qlo[3][8];
  uint8_t qhi[3][8];
};
static_assert(sizeof(Node8) == 80, "node size");
struct Tri8 {
  float v0[3];
  uint32_t id;
  float v1[3];
  uint32_t pad1;
  float v2[3];
  uint32_t pad2;
};
static_assert(sizeof(Tri8) == 48, "tri size");

void or_trace_bvh8(const or_scene* s, const void* nodes_, uint32_t node_count, const void* tris_,
                   uint32_t tri_count, const hpt_ray* rays, uint32_t n, int any_hit,
                   hpt_hit* hits, uint32_t* node_visits, uint32_t* tri_tests) {
  const Node8* nodes = (const Node8*)nodes_;
  const Tri8* tris = (const Tri8*)tris_;
  for (uint32_t i = 0; i < n; ++i) {
    const hpt_ray& hr = rays[i];
    Ray ray = to_ray(hr);
    RayPre r = prepare_ray(ray.origin, ray.direction);
    const vec3 inv = inv_dir(ray.direction);
    const uint32_t oct = (inv.x < 0.0f ? 1u : 0u) | (inv.y < 0.0f ? 2u : 0u) | (inv.z < 0.0f ? 4u : 0u);
    HitRec h{0, 0, 0, HPT_HIT_NONE, 0};
    float best_t = hr.tmax;
    uint32_t best_id = INVALID_INDEX;
    uint32_t nv = 0, nt = 0;
    bool done = false;
    std::vector<uint64_t> stack;
    // node group: base | imask<<32 | hits<<40
    uint32_t g_base = 0, g_imask = 1, g_hits = node_count ? (1u << (0u ^ oct)) : 0u;
    while (g_hits && !done) {
      const uint32_t b = (uint32_t)__builtin_ctz(g_hits);
      g_hits &= g_hits - 1;
      const uint32_t slot = b ^ oct;
      const uint32_t ni = g_base + (uint32_t)__builtin_popcount(g_imask & ((1u << slot) - 1u));
      if (g_hits) stack.push_back((uint64_t)g_base | ((uint64_t)g_imask << 32) | ((uint64_t)g_hits << 40));
      const Node8& nd = nodes[ni];
      ++nv;
      float sc[3];
      for (int a = 0; a < 3; ++a) {
        uint32_t bits = (uint32_t)nd.e[a] << 23;
        std::memcpy(&sc[a], &bits, 4);
      }
      uint32_t node_hits = 0, tri_hits = 0;
      for (uint32_t c = 0; c < 8; ++c) {
        if (nd.meta[c] == 0) continue;
        float tn = hr.tmin, tf = best_t;
        for (int a = 0; a < 3; ++a) {
          const float lo = std::fma((float)nd.qlo[a][c], sc[a], nd.p[a]);
          const float hi = std::fma((float)nd.qhi[a][c], sc[a], nd.p[a]);
          float t0 = (lo - ray.origin[a]) * inv[a];
          float t1 = (hi - ray.origin[a]) * inv[a];
          if (inv[a] < 0.0f) std::swap(t0, t1);
          tn = std::fmax(tn, t0);
          tf = std::fmin(tf, t1);
        }
        if (!(tn <= tf * BOX_MARGIN)) continue;
        if (nd.imask & (1u << c)) {
          node_hits |= 1u << (c ^ oct);
        } else {
          const uint32_t m = nd.meta[c];
          tri_hits |= (m >> 5) << (m & 31u);
        }
      }
      g_base = nd.child_base;
      g_imask = nd.imask;
      g_hits = node_hits;
      while (tri_hits) {
        const uint32_t tb = (uint32_t)__builtin_ctz(tri_hits);
        tri_hits &= tri_hits - 1;
        const Tri8& tr = tris[nd.tri_base + tb];
        ++nt;
        float t, u, v;
        if (!tri_test(r, V3(tr.v0), V3(tr.v1), V3(tr.v2), t, u, v)) continue;
        if (!(t >= hr.tmin && t <= best_t)) continue;
        if (any_hit) { h = HitRec{t, u, v, HPT_HIT_TRIANGLE, tr.id}; done = true; break; }
        if (h.kind && !better(t, tr.id, best_t, best_id)) continue;
        h = HitRec{t, u, v, HPT_HIT_TRIANGLE, tr.id};
        best_t = t;
        best_id = tr.id;
      }
      if (!g_hits && !done) {
        if (stack.empty()) break;
        uint64_t e = stack.back();
        stack.pop_back();
        g_base = (uint32_t)e;
        g_imask = (uint32_t)(e >> 32) & 0xFFu;
        g_hits = (uint32_t)(e >> 40) & 0xFFu;
      }
    }
    (void)tri_count;
    if (!done) {
      if (any_hit) {
        if (any_light(*s, ray, hr.tmin, hr.tmax)) h = HitRec{0, 0, 0, HPT_HIT_LIGHT, 0};
      } else {
        closest_lights(*s, ray, hr.tmin, hr.tmax, h);
      }
    }
    if (hits) to_hpt_hit(*s, h, hits[i]);
    if (node_visits) node_visits[i] = nv;
    if (tri_tests) tri_tests[i] = nt;
  }
}

void or_camera_ray(const or_scene* s, const or_settings* st, uint32_t frame, uint32_t x, uint32_t y,
                   hpt_ray* out) {
  Globals g = make_globals(*s, *st, frame);
  RNGState rng = rng_initialize(x, y, frame);
  Ray r = camera_ray(rng, g, s->cameras[st->camera_index], s->bn, x, y);
  S3(out->origin, r.origin);
  S3(out->direction, r.direction);
  out->tmin = 0.0f;
  out->tmax = WORLD_SIZE;
}

int or_trace_paths(const or_scene* s, const or_settings* st, uint32_t frame, uint32_t x0,
                   uint32_t y0, uint32_t x1, uint32_t y1, float* radiance, float* albedo,
                   float* normal, int threads) {
  if (!s || !st || st->camera_index >= s->cameras.size() || x1 <= x0 || y1 <= y0) return 1;
  const Globals g = make_globals(*s, *st, frame);
  const Variant var = variant_of(st->features);
  const uint32_t w = x1 - x0;
  parallel_rows(y0, y1, threads, [&](uint32_t y) {
    for (uint32_t x = x0; x < x1; ++x) {
      PathOut p = trace_path(*s, g, var, x, y);
      size_t k = ((size_t)(y - y0) * w + (x - x0)) * 3;
      if (radiance) S3(radiance + k, p.radiance);
      if (albedo) S3(albedo + k, p.albedo);
      if (normal) S3(normal + k, p.normal);
    }
  });
  return 0;
}

int or_render(const or_scene* s, const or_settings* st, uint32_t frame_begin, uint32_t count,
              float* accum, float* albedo, float* normal, uint8_t* final_rgba8, int threads) {
  if (!s || !st || st->camera_index >= s->cameras.size()) return 1;
  const Variant var = variant_of(st->features);
  for (uint32_t f = frame_begin; f < frame_begin + count; ++f) {
    const Globals g = make_globals(*s, *st, f);
    parallel_rows(0, st->height, threads, [&](uint32_t y) {
      for (uint32_t x = 0; x < st->width; ++x) {
        PathOut p = trace_path(*s, g, var, x, y);
        size_t k = ((size_t)y * st->width + x) * 4;
        auto acc = [&](float* img, vec3 v) {
          vec3 old(img[k], img[k + 1], img[k + 2]);
          vec3 nv = g.frame_index == 0 ? v : mix3(old, v, 1.0f / ((float)g.frame_index + 1.0f));
          img[k] = nv.x; img[k + 1] = nv.y; img[k + 2] = nv.z; img[k + 3] = 1.0f;
          return nv;
        };
        vec3 c = acc(accum, p.radiance);
        acc(albedo, p.albedo);
        acc(normal, p.normal);
        if (final_rgba8) {
          vec3 fc = final_color(g, c);
          final_rgba8[k] = to_unorm8(fc.x);
          final_rgba8[k + 1] = to_unorm8(fc.y);
          final_rgba8[k + 2] = to_unorm8(fc.z);
          final_rgba8[k + 3] = 255;
        }
      }
    });
  }
  return 0;
}

void or_final_color(const or_settings* st, const float rgb[3], float out[3]) {
  Globals g;
  g.exposure_value = st->exposure_value;
  g.enable_tonemap = st->enable_tonemap != 0;
  g.enable_aces = st->enable_aces != 0;
  g.use_simple_aces = st->use_simple_aces != 0;
  S3(out, final_color(g, V3(rgb)));
}

void or_pcg4d(uint32_t v[4]) {
  uvec4 u{v[0], v[1], v[2], v[3]};
  pcg4d(u);
  v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
}

float or_u32_to_unit(uint32_t u) { return u32_to_unit(u); }

void or_rng_init(uint32_t x, uint32_t y, uint32_t frame, uint32_t s0[4], uint32_t s1[4]) {
  RNGState r = rng_initialize(x, y, frame);
  s0[0] = r.s0.x; s0[1] = r.s0.y; s0[2] = r.s0.z; s0[3] = r.s0.w;
  s1[0] = r.s1.x; s1[1] = r.s1.y; s1[2] = r.s1.z; s1[3] = r.s1.w;
}

void or_blue_noise_texel(const or_scene* s, uint32_t tx, uint32_t ty, uint8_t out[4]) {
  std::memcpy(out, &s->bn.rgba[((size_t)ty * 1024 + tx) * 4], 4);
}

static RNGState rng_from(const uint32_t r[4]) {
  RNGState s{};
  s.s0 = uvec4{r[0], r[1], r[2], r[3]};
  return s;
}
static void rng_to(const RNGState& s, uint32_t r[4]) {
  r[0] = s.s0.x; r[1] = s.s0.y; r[2] = s.s0.z; r[3] = s.s0.w;
}

void or_bxdf_eval(const hpt_material* m, const or_frame* fr, const float V[3], const float L[3],
                  float f[3], float* pdf) {
  Material mat = to_material(*m);
  State st = frame_state(fr);
  float p = 0.0f;
  vec3 r = bxdf_eval(fr->any_non_specular_bounce != 0, st, mat, V3(V), V3(fr->n), V3(L), p);
  S3(f, r);
  *pdf = p;
}

void or_bxdf_sample(const hpt_material* m, const or_frame* fr, const float V[3], uint32_t rng[4],
                    float L[3], float f[3], float* pdf, uint32_t* flags) {
  Material mat = to_material(*m);
  State st = frame_state(fr);
  RNGState rs = rng_from(rng);
  bool anb = fr->any_non_specular_bounce != 0;
  vec3 Lo;
  float p = 0.0f;
  uint32_t fl = 0;
  vec3 r = bxdf_sample(rs, anb, st, mat, V3(V), V3(fr->n), Lo, p, fl);
  rng_to(rs, rng);
  S3(L, Lo);
  S3(f, r);
  *pdf = p;
  *flags = fl;
}

void or_light_sample(const hpt_light* light, uint32_t num_lights, const float p[3], uint32_t rng[4],
                     float dir[3], float normal[3], float emission[3], float* dist, float* pdf) {
  Globals g;
  g.num_of_lights = num_lights;
  State st;
  st.first_hit_position = V3(p);
  RNGState rs = rng_from(rng);
  LightSample ls;
  light_sample(rs, g, st, to_light(*light), ls);
  rng_to(rs, rng);
  S3(dir, ls.direction);
  S3(normal, ls.normal);
  S3(emission, ls.emission);
  *dist = ls.dist;
  *pdf = ls.pdf;
}

float or_light_intersect(const hpt_light* light, const float org[3], const float dir[3], float* pdf) {
  vec3 nrm;
  float p = 1.0f;
  float t = light_t(to_light(*light), Ray{V3(org), V3(dir)}, nrm, p);
  *pdf = p;
  return t;
}

double or_build_alias(const float* lum, uint32_t n, float* prob, uint32_t* alias) {
  return build_env_alias(lum, n, prob, alias);
}

static Globals env_globals(const or_scene* s) {
  or_settings st{};
  st.width = st.height = 1;
  return make_globals(*s, st, 0);
}

void or_env_sample(const or_scene* s, uint32_t rng[4], float dir[3], float color[3], float* pdf) {
  Globals g = env_globals(s);
  RNGState rs = rng_from(rng);
  vec3 c;
  float p;
  vec3 d = sample_env_map(rs, g, s->env, c, p);
  rng_to(rs, rng);
  S3(dir, d);
  S3(color, c);
  *pdf = p;
}

void or_env_eval(const or_scene* s, const float dir[3], uint32_t depth, float bsdf_pdf, float rgb[3]) {
  Globals g = env_globals(s);
  S3(rgb, env_eval(g, s->env, 0, depth, bsdf_pdf, V3(dir), false, false));
}

float or_env_pdf_dir(const or_scene* s, const float dir[3]) {
  Globals g = env_globals(s);
  vec3 d = V3(dir);
  const float theta = std::acos(d.z);
  const vec2 uv = vec2(-(PI + std::atan2(d.y, d.x)) * INV_TWO_PI - g.env_rotation, theta * INV_PI);
  return env_pdf(g, s->env, uv, theta);
}

void or_env_alias(const or_scene* s, float* prob, uint32_t* alias, float* lum, float* total) {
  size_t n = s->env.prob.size();
  if (prob) std::memcpy(prob, s->env.prob.data(), n * 4);
  if (alias) std::memcpy(alias, s->env.alias.data(), n * 4);
  if (lum) std::memcpy(lum, s->env.lum.data(), n * 4);
  if (total) *total = s->env_total;
}

void or_sample_hg(const float V[3], float g, float r1, float r2, float out[3]) {
  S3(out, sample_hg(V3(V), g, r1, r2));
}
float or_phase_hg(float c, float g) { return phase_hg(c, g); }
void or_sample_ggx_vndf(const float wi[3], float ax, float ay, float r1, float r2, float out[3]) {
  S3(out, sample_ggx_vndf(V3(wi), ax, ay, r1, r2));
}
void or_cosine_sample_hemisphere(float r1, float r2, float out[3]) {
  S3(out, cosine_sample_hemisphere(r1, r2));
}
void or_sample_gtr1(float roughness, float r1, float r2, float out[3]) {
  S3(out, sample_gtr1(roughness, r1, r2));
}
void or_tonemap_aces_fitted(const float in[3], float out[3]) { S3(out, aces_fitted(V3(in))); }
void or_tonemap_aces(const float in[3], float out[3]) { S3(out, aces(V3(in))); }
void or_tonemap_reinhard(const float in[3], float limit, float out[3]) { S3(out, tonemap(V3(in), limit)); }
float or_linear_2_srgb(float x) { return linear_2_srgb(x); }

}  // extern "C"
