// dev_traverse.h — ray queries on gfx950: the watertight ray/triangle test
// and the walk of the 8-wide compressed BVH, plus the analytic light shapes
// that the reference puts in its acceleration structure as procedural AABBs
// (lights.rint.glsl:16-37).
//
// Query semantics (the reference leaves them to the Vulkan driver):
//  * triangles are double-sided (gl_RayFlagsOpaqueEXT, no culling);
//  * closest hit = minimum of (t, global triangle id) over t in [tmin, tmax];
//    an analytic light replaces it only when strictly closer;
//  * any hit = any triangle or analytic light on [tmin, tmax].
#pragma once

#include "common.h"

namespace hpt {

// ----------------------------------------------------------------- triangle
struct RayPre {
  f3 org, dir;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

__device__ __forceinline__ RayPre prepare_ray(f3 org, f3 dir) {
#pragma clang fp contract(off)
  RayPre r;
  r.org = org;
  r.dir = dir;
  const float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
  int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  const float dz = comp(dir, kz);
  if (dz < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.Sx = comp(dir, kx) / dz;
  r.Sy = comp(dir, ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}

// Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1)
// 2013; double-sided, with the paper's double-precision recomputation when an
// edge function is exactly 0.  t = T/det, (u, v) = weights of B and C.
__device__ __forceinline__ bool tri_test(const RayPre& r, f3 A0, f3 B0, f3 C0, float& t, float& u,
                                         float& v) {
#pragma clang fp contract(off)
  const f3 A = A0 - r.org, B = B0 - r.org, C = C0 - r.org;
  const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  const float Ax = comp(A, r.kx) - r.Sx * Akz;
  const float Ay = comp(A, r.ky) - r.Sy * Akz;
  const float Bx = comp(B, r.kx) - r.Sx * Bkz;
  const float By = comp(B, r.ky) - r.Sy * Bkz;
  const float Cx = comp(C, r.kx) - r.Sx * Ckz;
  const float Cy = comp(C, r.ky) - r.Sy * Ckz;
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
  const float Az = r.Sz * Akz;
  const float Bz = r.Sz * Bkz;
  const float Cz = r.Sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  t = T / det;
  u = V / det;
  v = W / det;
  return true;
}

// ------------------------------------------------------------ analytic lights
// Evaluated with every product and sum rounded on its own (no fused
// multiply-add), as in the host restatement: the sphere discriminant
// b^2 - |op|^2 + r^2 cancels for a distant light, where one ulp of
// contraction moves t by ~1e-5 relative.  The vector helpers carry the
// pragma themselves: it covers only the code written inside a function.
namespace nofma {
__device__ __forceinline__ float dot(f3 a, f3 b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
#pragma clang fp contract(off)
  return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ f3 axpy(f3 o, f3 d, float t) {  // o + d * t
#pragma clang fp contract(off)
  return f3{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
}
__device__ __forceinline__ f3 normalize(f3 a) {
  const float l = sqrtf(dot(a, a));
  return f3{a.x / l, a.y / l, a.z / l};
}
}  // namespace nofma

// common/intersection.glsl:6-33
__device__ __forceinline__ float sphere_isect(float rad, f3 pos, f3 o, f3 d) {
#pragma clang fp contract(off)
  const f3 op = pos - o;
  const float b = nofma::dot(op, d);
  float det = b * b - nofma::dot(op, op) + rad * rad;
  if (det < 0.0f) return INF;
  det = sqrtf(det);
  const float t1 = b - det;
  if (t1 > EPS) return t1;
  const float t2 = b + det;
  if (t2 > EPS) return t2;
  return INF;
}

// inc/light/quad.glsl:42-60 with common/intersection.glsl:43-72.
__device__ __forceinline__ float quad_light_isect(const hpt_light& L, f3 o, f3 d, float& pdf) {
#pragma clang fp contract(off)
  const f3 lu = ld3(L.u), lv = ld3(L.v), lp = ld3(L.position);
  const f3 normal = nofma::normalize(nofma::cross(lv, lu));
  const float cos_theta = nofma::dot(-d, normal);
  if (cos_theta < 0.0f) return INF;
  const float pw = nofma::dot(normal, lp);
  const float uu = nofma::dot(lu, lu), vv = nofma::dot(lv, lv);
  const f3 u = f3{lu.x / uu, lu.y / uu, lu.z / uu};
  const f3 v = f3{lv.x / vv, lv.y / vv, lv.z / vv};
  const float dt = nofma::dot(d, normal);
  const float t = (pw - nofma::dot(normal, o)) / dt;
  if (t > EPS) {
    const f3 p = nofma::axpy(o, d, t);
    const f3 vi = p - lp;
    const float a1 = nofma::dot(u, vi);
    if (a1 >= 0.0f && a1 <= 1.0f) {
      const float a2 = nofma::dot(v, vi);
      if (a2 >= 0.0f && a2 <= 1.0f) {
        pdf = (t * t) / (L.area * cos_theta);
        return t;
      }
    }
  }
  return INF;
}

// inc/light/sphere.glsl:50-66.
__device__ __forceinline__ float sphere_light_isect(const hpt_light& L, f3 o, f3 d, float& pdf) {
#pragma clang fp contract(off)
  const f3 lp = ld3(L.position);
  const f3 lo = lp - o;
  const float dist = sqrtf(nofma::dot(lo, lo));
  if (dist < L.radius) return INF;
  const float t = sphere_isect(L.radius, lp, o, d);
  if (t < 0.0f) return INF;
  const f3 hit_pt = nofma::axpy(o, d, t);
  const float cos_theta = nofma::dot(-d, nofma::normalize(hit_pt - lp));
  pdf = (t * t) / (L.area * cos_theta * 0.5f);
  return t;
}

__device__ __forceinline__ float light_isect(const hpt_light& L, f3 o, f3 d, float& pdf) {
  pdf = 1.0f;
  if (L.type == HPT_LIGHT_QUAD) return quad_light_isect(L, o, d, pdf);
  if (L.type == HPT_LIGHT_SPHERE) return sphere_light_isect(L, o, d, pdf);
  return INF;
}

// ------------------------------------------------------------------ BVH walk
struct TraceResult {
  float t, u, v;
  uint32_t id;  // global triangle id, LIGHT_BIT|light, or MISS_ID
};

// Per-lane traversal stack: `lds` points at this lane's slot 0 and entries
// are `lds_stride` uint2 apart; entries past kShortStack go to `spill`
// (stride `spill_stride`).  Needed depth <= tree levels.

// The LDS part is held as an LDS-space pointer: with a generic pointer the
// compiler merges the LDS and spill accesses into one flat access through a
// select of two 64-bit addresses (flat instructions take the vector memory
// path and the addresses two registers each); with distinct address spaces
// the LDS entries use ds_read/ds_write and the spill accesses stay in their
// own, rarely taken, branch.
// Entries are node groups (two words) kept as one 64-bit word in LDS.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
struct LaneStack {
  lds_u64* lds;
  int lds_stride;
  uint2* spill;
  int spill_stride;
  int sp;
  __device__ __forceinline__ LaneStack(uint2* lds_slot, int lds_stride_, uint2* spill_slot, int spill_stride_, int sp_)
      : lds((lds_u64*)lds_slot), lds_stride(lds_stride_), spill(spill_slot), spill_stride(spill_stride_), sp(sp_) {}
  __device__ __forceinline__ void push(uint2 e) {
    if (sp < kShortStack) lds[sp * lds_stride] = (unsigned long long)e.y << 32 | e.x;
    else spill[(sp - kShortStack) * spill_stride] = e;
    sp++;
  }
  __device__ __forceinline__ uint2 pop() {
    sp--;
    if (sp < kShortStack) {
      const unsigned long long v = lds[sp * lds_stride];
      return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    }
    return spill[(sp - kShortStack) * spill_stride];
  }
};

__device__ __forceinline__ float byte_f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }

// Resumable walk of the 8-wide tree for one ray.  Node group =
// (base, hits << 8 | imask); hit bit b stands for slot b ^ oct.  step()
// visits one node and tests its triangles; ANY stops at the first triangle
// on [tmin, tmax].  Every kernel drives the same step(): one ray per lane,
// or persistent lanes that fetch a new ray when theirs ends.
template <bool ANY>
struct Walker {
  RayPre r;
  f3 inv;
  uint32_t oct;
  float tmin, best_t;
  uint32_t best_id;
  float bu, bv;
  bool found;
  uint2 grp;
  uint32_t nv, nt;

  __device__ __forceinline__ void begin(const RayPre& ray, float t0, float t1, LaneStack& st) {
    r = ray;
    const float dx = fabsf(r.dir.x) < 1e-30f ? copysignf(1e-30f, r.dir.x) : r.dir.x;
    const float dy = fabsf(r.dir.y) < 1e-30f ? copysignf(1e-30f, r.dir.y) : r.dir.y;
    const float dz = fabsf(r.dir.z) < 1e-30f ? copysignf(1e-30f, r.dir.z) : r.dir.z;
    inv = mk3(1.0f / dx, 1.0f / dy, 1.0f / dz);
    oct = (inv.x < 0.0f ? 1u : 0u) | (inv.y < 0.0f ? 2u : 0u) | (inv.z < 0.0f ? 4u : 0u);
    tmin = t0;
    best_t = t1;
    best_id = MISS_ID;
    bu = bv = 0.0f;
    found = false;
    nv = nt = 0;
    st.sp = 0;
    grp = make_uint2(0u, ((1u << oct) << 8) | 1u);
  }

  // One node visit; returns true when the walk is over.
  __device__ __forceinline__ bool step(const SceneView& sv, LaneStack& st) {
    uint32_t hits = grp.y >> 8;
    const uint32_t b = __builtin_ctz(hits);
    hits &= hits - 1u;
    const uint32_t slot = b ^ oct;
    const uint32_t imask = grp.y & 0xFFu;
    const uint32_t ni = grp.x + __builtin_popcount(imask & ((1u << slot) - 1u));
    if (hits) st.push(make_uint2(grp.x, (hits << 8) | imask));

    const uint4* np = reinterpret_cast<const uint4*>(sv.nodes + ni);
    const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3], n4 = np[4];
    ++nv;
    const f3 o = r.org;
    const f3 p = mk3(__uint_as_float(n0.x), __uint_as_float(n0.y), __uint_as_float(n0.z));
    const float sx = __uint_as_float((n0.w & 0xFFu) << 23);
    const float sy = __uint_as_float(((n0.w >> 8) & 0xFFu) << 23);
    const float sz = __uint_as_float(((n0.w >> 16) & 0xFFu) << 23);
    const uint32_t nimask = n0.w >> 24;
    // Bytes per child (children 0-3 | 4-7): qlo_x = n2.x|n2.y, qlo_y =
    // n2.z|n2.w, qlo_z = n3.x|n3.y, qhi_x = n3.z|n3.w, qhi_y = n4.x|n4.y,
    // qhi_z = n4.z|n4.w, meta = n1.z|n1.w.  The near plane of an axis is
    // qhi when the ray runs negative along it: choosing the plane words once
    // per node gives the same distances as testing lo/hi and swapping.
    const bool fx = oct & 1u, fy = oct & 2u, fz = oct & 4u;
    const uint32_t nx0 = fx ? n3.z : n2.x, nx1 = fx ? n3.w : n2.y, ex0 = fx ? n2.x : n3.z, ex1 = fx ? n2.y : n3.w;
    const uint32_t ny0 = fy ? n4.x : n2.z, ny1 = fy ? n4.y : n2.w, ey0 = fy ? n2.z : n4.x, ey1 = fy ? n2.w : n4.y;
    const uint32_t nz0 = fz ? n4.z : n3.x, nz1 = fz ? n4.w : n3.y, ez0 = fz ? n3.x : n4.z, ez1 = fz ? n3.y : n4.w;
    uint32_t node_hits = 0, tri_hits = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int k = c & 3;
      const bool hi4 = c >= 4;
      const uint32_t meta = (hi4 ? n1.w : n1.z) >> (8 * k) & 0xFFu;
      const float qnx = byte_f(hi4 ? nx1 : nx0, k), qfx = byte_f(hi4 ? ex1 : ex0, k);
      const float qny = byte_f(hi4 ? ny1 : ny0, k), qfy = byte_f(hi4 ? ey1 : ey0, k);
      const float qnz = byte_f(hi4 ? nz1 : nz0, k), qfz = byte_f(hi4 ? ez1 : ez0, k);
      float tn, tf;
      {
#pragma clang fp contract(off)
        const float tnx = (fmaf(qnx, sx, p.x) - o.x) * inv.x, tfx = (fmaf(qfx, sx, p.x) - o.x) * inv.x;
        const float tny = (fmaf(qny, sy, p.y) - o.y) * inv.y, tfy = (fmaf(qfy, sy, p.y) - o.y) * inv.y;
        const float tnz = (fmaf(qnz, sz, p.z) - o.z) * inv.z, tfz = (fmaf(qfz, sz, p.z) - o.z) * inv.z;
        tn = fmaxf(fmaxf(tmin, tnx), fmaxf(tny, tnz));
        tf = fminf(fminf(best_t, tfx), fminf(tfy, tfz));
        tf = tf * BOX_MARGIN;
      }
      // Inner slots carry meta 0x1F, whose triangle mask (meta >> 5) is 0.
      const uint32_t hit = (meta != 0u && tn <= tf) ? ~0u : 0u;
      node_hits |= hit & (((nimask >> c) & 1u) << (c ^ oct));
      tri_hits |= hit & ((meta >> 5) << (meta & 31u));
    }
    grp = make_uint2(n1.x, (node_hits << 8) | nimask);
    const uint32_t tri_base = n1.y;
    while (tri_hits) {
      const uint32_t tb = __builtin_ctz(tri_hits);
      tri_hits &= tri_hits - 1u;
      const float4* tp = reinterpret_cast<const float4*>(sv.ltris + tri_base + tb);
      const float4 a = tp[0], bq = tp[1], cq = tp[2];
      ++nt;
      float t, u, v;
      if (!tri_test(r, mk3(a.x, a.y, a.z), mk3(bq.x, bq.y, bq.z), mk3(cq.x, cq.y, cq.z), t, u, v)) continue;
      if (!(t >= tmin && t <= best_t)) continue;
      const uint32_t id = __float_as_uint(a.w);
      if (!ANY && found && !(t < best_t || id < best_id)) continue;
      found = true;
      best_t = t;
      best_id = id;
      bu = u;
      bv = v;
      if (ANY) return true;
    }
    if ((grp.y >> 8) == 0u) {
      if (st.sp == 0) return true;
      grp = st.pop();
    }
    return false;
  }
};

template <bool ANY>
__device__ __forceinline__ bool trace_triangles(const SceneView& sv, const RayPre& r, float tmin,
                                                float tmax, TraceResult& res, LaneStack& st,
                                                uint32_t* nodes_visited, uint32_t* tris_tested) {
  if (sv.node_count == 0) return false;
  Walker<ANY> w;
  w.begin(r, tmin, tmax, st);
  while (!w.step(sv, st)) {
  }
  if (nodes_visited) *nodes_visited = w.nv;
  if (tris_tested) *tris_tested = w.nt;
  if (w.found) {
    res.t = w.best_t;
    res.u = w.bu;
    res.v = w.bv;
    res.id = w.best_id;
  }
  return w.found;
}

// Analytic quad/sphere lights after the triangle walk: a light replaces the
// triangle hit only when strictly closer; among lights the lowest index wins.
__device__ __forceinline__ void closest_lights(const SceneView& sv, f3 o, f3 d, float tmin, float tmax, bool found,
                                               TraceResult& res) {
  for (uint32_t i = 0; i < sv.num_lights; ++i) {
    const hpt_light& L = sv.lights[i];
    if (L.type != HPT_LIGHT_QUAD && L.type != HPT_LIGHT_SPHERE) continue;
    float pdf;
    const float t = light_isect(L, o, d, pdf);
    if (!(t >= tmin && t <= tmax)) continue;
    if (found && !(t < res.t)) continue;
    found = true;
    res = TraceResult{t, 0.0f, 0.0f, LIGHT_BIT | i};
  }
}

__device__ __forceinline__ bool any_light(const SceneView& sv, f3 o, f3 d, float tmin, float tmax) {
  for (uint32_t i = 0; i < sv.num_lights; ++i) {
    const hpt_light& L = sv.lights[i];
    if (L.type != HPT_LIGHT_QUAD && L.type != HPT_LIGHT_SPHERE) continue;
    float pdf;
    const float t = light_isect(L, o, d, pdf);
    if (t >= tmin && t <= tmax) return true;
  }
  return false;
}

// Closest hit over triangles and analytic lights.
__device__ __forceinline__ TraceResult trace_closest(const SceneView& sv, f3 o, f3 d, float tmin,
                                                     float tmax, LaneStack& st,
                                                     uint32_t* nv = nullptr, uint32_t* nt = nullptr) {
  TraceResult res{0.0f, 0.0f, 0.0f, MISS_ID};
  const RayPre r = prepare_ray(o, d);
  const bool found = trace_triangles<false>(sv, r, tmin, tmax, res, st, nv, nt);
  closest_lights(sv, o, d, tmin, tmax, found, res);
  return res;
}

__device__ __forceinline__ bool trace_any(const SceneView& sv, f3 o, f3 d, float tmin, float tmax,
                                          LaneStack& st, uint32_t* nv = nullptr, uint32_t* nt = nullptr) {
  TraceResult res;
  const RayPre r = prepare_ray(o, d);
  if (trace_triangles<true>(sv, r, tmin, tmax, res, st, nv, nt)) return true;
  return any_light(sv, o, d, tmin, tmax);
}

}  // namespace hpt
