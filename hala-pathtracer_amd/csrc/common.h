// common.h — types shared by the host library and the gfx950 kernels.
//
// HBM layout of one committed scene (all arrays 16-byte aligned):
//   nodes    Node8[node_count]      80 B  8-wide compressed BVH nodes
//   ltris    LeafTri[tri_count]     48 B  world-space vertices in leaf order
//   tri_shade TriShade[tri_count]  128 B  vertex normals/tangents/uvs + primitive, by global triangle id
//   prims    PrimInfo[nprims]       48 B  inverse-transpose 3x3 + material
//   materials hpt_material[nmat]   144 B  the std140 struct as given
//   lights   hpt_light[17]          80 B  std140; slot num_lights.. zero-filled
//   cameras  hpt_camera[8]          80 B
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hpt.h"

#define HD __host__ __device__ __forceinline__

namespace hpt {

// ------------------------------------------------------------------ vectors
struct f3 {
  float x, y, z;
};
HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
HD f3 mk3(float s) { return f3{s, s, s}; }
HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
HD f3 operator+(f3 a, float s) { return f3{a.x + s, a.y + s, a.z + s}; }
HD f3 operator-(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
HD f3 operator-(float s, f3 a) { return f3{s - a.x, s - a.y, s - a.z}; }
HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
HD f3& operator+=(f3& a, f3 b) { a = a + b; return a; }
HD f3& operator*=(f3& a, f3 b) { a = a * b; return a; }
HD f3& operator*=(f3& a, float s) { a = a * s; return a; }
HD f3& operator/=(f3& a, float s) { a = a / s; return a; }
HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
HD float length(f3 a) { return sqrtf(dot(a, a)); }
HD f3 normalize(f3 a) { return a / length(a); }
HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
HD f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

// ---------------------------------------------------------------- constants
// common/common.glsl:6-15 and inc/constants.glsl:1-44.
constexpr float PI = 3.14159265358979323846264338327950288f;
constexpr float INV_PI = 1.0f / PI;
constexpr float TWO_PI = 2.0f * PI;
constexpr float INV_TWO_PI = 1.0f / TWO_PI;
constexpr float INV_FOUR_PI = 1.0f / (4.0f * PI);
constexpr float EPS = 0.0003f;
constexpr float INF = 1000000.0f;
constexpr float WORLD_SIZE = 10000.0f;
constexpr uint32_t INVALID = 0xFFFFFFFFu;
constexpr uint32_t RAY_FLAGS_REFLECTION = 0x8;
constexpr uint32_t RAY_FLAGS_TRANSMISSION = 0x10;

// Far-distance slack of the node slab test, 1 + 2^-18: keeps the quantised
// box test conservative against the rounding of (lo - o) * inv.
constexpr float BOX_MARGIN = 1.0f + 1.0f / 262144.0f;

// Traversal stack entries each lane keeps in LDS: 8 entries x 8 B x 256 lanes
// = 16 KB per block, small enough for 7 resident waves per SIMD (12 entries
// capped the traversal kernels at 6 and measured 6% slower).  Deeper trees
// spill the rest to a global buffer.
#ifndef HPT_SHORT_STACK
#define HPT_SHORT_STACK 8
#endif
constexpr int kShortStack = HPT_SHORT_STACK;

// Hit record id encoding: triangle -> global triangle id (< 2^31);
// light -> LIGHT_BIT | light index; miss -> MISS_ID.
constexpr uint32_t LIGHT_BIT = 0x80000000u;
constexpr uint32_t MISS_ID = 0xFFFFFFFFu;

// ---------------------------------------------------------------- BVH layout
// One 8-wide node (80 B, five 16-B loads).  Children are quantised to 8 bits
// per plane on the grid p + q * 2^(e-127).  Slot s holds the child lying in
// octant s of the node (bit0 +x, bit1 +y, bit2 +z), so walking hit slots in
// the order (bit ^ ray_octant) visits near children first.
//  imask bit s : slot s is an inner node; inner children are stored
//                contiguously from child_base in slot order.
//  meta[s]     : 0 = empty slot; inner = 0x1F; leaf = (unary count << 5) |
//                offset of its first triangle from tri_base (count 1..3).
struct Node8 {
  float p[3];
  uint8_t e[3];
  uint8_t imask;
  uint32_t child_base;
  uint32_t tri_base;
  uint8_t meta[8];
  uint8_t qlo[3][8];
  uint8_t qhi[3][8];
};
static_assert(sizeof(Node8) == 80, "Node8 is 80 bytes");

// Leaf triangle (48 B): world-space vertices; id = global triangle id
// (primitive-major, index order), which also breaks distance ties.
struct LeafTri {
  float v0[3];
  uint32_t id;
  float v1[3];
  uint32_t pad1;
  float v2[3];
  uint32_t pad2;
};
static_assert(sizeof(LeafTri) == 48, "LeafTri is 48 bytes");

// Shading record of one triangle (128 B, one cache line): its vertices'
// normals and tangents copied from the vertex buffer, its primitive, its
// texture coordinates and its primitive's material.  A hit reads one line
// here instead of an index triple and three scattered vertices, and the
// material and primitive records can then be fetched side by side.
struct TriShade {
  float n[3][3];   // vertex normals (object space)
  float t[3][3];   // vertex tangents (object space)
  uint32_t prim;
  float uv[3][2];
  uint32_t material;  // the primitive's material index (PrimInfo::material_index)
  uint32_t pad[6];
};
static_assert(sizeof(TriShade) == 128, "TriShade is 128 bytes");

struct PrimInfo {  // 48 B
  float it[9];  // transpose(inverse(mat3(transform))), column-major
  uint32_t material_index;
  uint32_t local_base;  // global id of the primitive's first triangle
  uint32_t pad;
};
static_assert(sizeof(PrimInfo) == 48, "PrimInfo is 48 bytes");

// One environment texel as the shading kernels read it (128 B, one cache
// line): its alias-table entry and the RGB of the 3x3 texels around it, so a
// bilinear look-up inside the texel and the alias draw touch one line
// instead of the table and two texture rows.
struct EnvTexel {
  float prob;       // acceptance probability of the texel
  uint32_t alias;   // texel drawn otherwise
  float lum;        // luminance of the texel
  float lum_alias;  // luminance of the alias
  float rgb[3][3][3];  // [row j-1..j+1, clamped][column i-1..i+1, wrapped][rgb]
  float pad;
};
static_assert(sizeof(EnvTexel) == 128, "EnvTexel is 128 bytes");

// One material texture in the texel pool (RGBA8, row 0 at v = 0).
struct TexInfo {
  uint32_t offset;  // first texel in the pool
  uint32_t width, height;
  uint32_t pad;
};

// Device view of a committed scene, passed to kernels by value.
struct SceneView {
  const Node8* nodes;
  const LeafTri* ltris;
  const TriShade* tri_shade;
  const uint8_t* tri_key;  // shading sort key (material bucket) of each global triangle
  const PrimInfo* prims;
  const hpt_material* materials;
  const hpt_light* lights;
  const hpt_camera* cameras;
  const float4* env_rgb;       // width*height texels, .w unused
  const EnvTexel* env_texels;  // width*height records built from env_rgb and the alias table
  const uchar4* blue_noise;  // 1024x1024 RGBA8
  const uchar4* tex_texels;  // all material textures, RGBA8
  const TexInfo* tex_info;
  const float* srgb_lut;     // 256 entries: 8-bit sRGB code -> linear
  uint32_t node_count;
  uint32_t tri_count;
  uint32_t num_lights;
  uint32_t material_count;
};

// The MainUBO values (inc/global.glsl:11-30) one frame reads.
struct FrameParams {
  float ground[3];
  float sky[3];
  float res_x, res_y;
  uint32_t width, height;
  uint32_t max_depth, rr_depth;
  uint32_t frame_index;   // frame of the first path slice of the batch
  uint32_t batch_frames;  // frames traced together: path p is pixel p % pixels of frame frame_index + p / pixels
  uint32_t pixels;        // width * height
  uint32_t camera_index;
  uint32_t env_type;
  uint32_t env_w, env_h;
  float env_total_sum;
  float env_rotation;
  float env_intensity;
  float exposure_value;
  uint32_t enable_tonemap, enable_aces, use_simple_aces;
  uint32_t num_lights;
  float tan_half_fov;  // tan(camera.yfov * 0.5), evaluated once per frame on the host
  // When non-null the batch's first frame is read from this device counter
  // instead of frame_index (a captured batch replayed as a graph advances it).
  const uint32_t* frame_ptr;
  // Shading in a wavefront of several jobs: (max_depth, rr_depth) of job j at
  // job_limits[j], the job of path p in the path state; null for one job.
  const uint2* job_limits;
};

}  // namespace hpt
