// bvh_gpu.hip — on-GPU build of the 8-wide compressed BVH (Node8/LeafTri).
//
// The same tree as the host build (bvh_build.cpp), made on the device:
//  * top phase, one level of the binary tree per round, for nodes of more than
//    kSmall triangles: blocks of kChunk triangles bin their centroids into 32
//    bins per axis in LDS (counts, triangle and centroid bounds as
//    order-preserving integers, so atomics give exact min/max) and flush them
//    to the node's bins; one block per node evaluates the SAH over the 93
//    splits; a stable three-pass partition (count, offsets, scatter with wave
//    ballots) reorders the node's triangle indices;
//  * bottom phase: one block per subtree of at most kSmall triangles builds it
//    depth-first in LDS (bounds reduction, binning, SAH, stable partition);
//  * collapse: one thread per 8-wide node per level opens the largest binary
//    child until 8 slots are used, assigns slots by octant, quantises, and
//    writes nodes and leaf triangles at offsets from device prefix sums, in
//    the breadth-first order of the host build.
// Every floating-point step is the host build's (same operations, no
// contraction), so the bytes match.
#include "bvh_gpu.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#pragma clang fp contract(off)

namespace hpt {
namespace {

constexpr int kBins = 32;
constexpr uint32_t kMaxLeaf = 3;
constexpr uint32_t kSmall = 1024;  // subtrees up to this many triangles: one block each
constexpr uint32_t kChunk = 4096;  // triangles per block in the top phase
constexpr int kT = 256;
constexpr int kTopWords = 13;      // per bin: count, triangle lo/hi, centroid lo/hi
constexpr int kSmallWords = 7;     // per bin: count, triangle lo/hi
constexpr uint32_t kLeafBit = 0x80000000u;

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Fail(std::string(what) + ": " + hipGetErrorString(e));
}

// Order-preserving float <-> uint32 (for atomic min/max).
__device__ __forceinline__ uint32_t f2o(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
float o2f_host(uint32_t u) {
  const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
constexpr uint32_t kOrdPosInf = 0xFF800000u;  // f2o(+inf)
constexpr uint32_t kOrdNegInf = 0x007FFFFFu;  // f2o(-inf)

__host__ __device__ inline double box_area(const float lo[3], const float hi[3]) {
  double d[3];
  for (int a = 0; a < 3; ++a) {
    const double x = (double)hi[a] - (double)lo[a];
    d[a] = x > 0.0 ? x : 0.0;
  }
  return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}

__device__ __forceinline__ int bin_of(float c, float lo, float scale) {
  int b = (int)((c - lo) * scale);
  return b < 0 ? 0 : (b > kBins - 1 ? kBins - 1 : b);
}

// Binary node: lo.w = left child or first triangle, hi.w = right child or
// (count | kLeafBit).
struct BNode {
  float4 lo, hi;
};

__device__ __forceinline__ bool bn_leaf(const BNode& n) { return (__float_as_uint(n.hi.w) & kLeafBit) != 0; }

// ------------------------------------------------------------------ prep
__global__ void k_prep(const float* __restrict__ pos, uint32_t n, float4* __restrict__ tlo, float4* __restrict__ thi,
                       float4* __restrict__ cen, uint32_t* __restrict__ flags, uint32_t* __restrict__ idx) {
  const uint32_t t = blockIdx.x * kT + threadIdx.x;
  if (t >= n) return;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool finite = true;
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) {
      const float x = pos[(size_t)t * 9 + k * 3 + a];
      finite = finite && isfinite(x);
      lo[a] = fminf(lo[a], x);
      hi[a] = fmaxf(hi[a], x);
    }
  if (!finite) atomicOr(flags, 1u);
  tlo[t] = make_float4(lo[0], lo[1], lo[2], 0.f);
  thi[t] = make_float4(hi[0], hi[1], hi[2], 0.f);
  cen[t] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.f);
  idx[t] = t;
}

// Bounds of all triangles and of their centroids (ordered uints, 12 words).
__global__ void k_root_bounds(const float4* __restrict__ tlo, const float4* __restrict__ thi,
                              const float4* __restrict__ cen, uint32_t n, uint32_t* __restrict__ out) {
  __shared__ uint32_t red[12];
  if (threadIdx.x < 12) red[threadIdx.x] = threadIdx.x % 6 < 3 ? kOrdPosInf : kOrdNegInf;
  __syncthreads();
  float b[12] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY,
                 INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t t = blockIdx.x * kT + threadIdx.x; t < n; t += gridDim.x * kT) {
    const float4 l = tlo[t], h = thi[t], c = cen[t];
    b[0] = fminf(b[0], l.x), b[1] = fminf(b[1], l.y), b[2] = fminf(b[2], l.z);
    b[3] = fmaxf(b[3], h.x), b[4] = fmaxf(b[4], h.y), b[5] = fmaxf(b[5], h.z);
    b[6] = fminf(b[6], c.x), b[7] = fminf(b[7], c.y), b[8] = fminf(b[8], c.z);
    b[9] = fmaxf(b[9], c.x), b[10] = fmaxf(b[10], c.y), b[11] = fmaxf(b[11], c.z);
  }
  for (int k = 0; k < 12; ++k) {
    if (k % 6 < 3)
      atomicMin(&red[k], f2o(b[k]));
    else
      atomicMax(&red[k], f2o(b[k]));
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    if (threadIdx.x % 6 < 3)
      atomicMin(&out[threadIdx.x], red[threadIdx.x]);
    else
      atomicMax(&out[threadIdx.x], red[threadIdx.x]);
  }
}

// ------------------------------------------------------------------ top phase
struct TopTask {
  uint32_t node, first, count, block0;  // block0: first block of this task in the level's grid
  float lo[3], hi[3], clo[3], chi[3];   // triangle bounds, centroid bounds
};

struct TopSplit {
  int32_t axis;
  uint32_t split, nleft, nright;
  uint32_t left[12], right[12];  // ordered bounds: triangle lo/hi, centroid lo/hi
};

__device__ __forceinline__ void task_scales(const TopTask& tk, bool on[3], float scale[3]) {
  for (int a = 0; a < 3; ++a) {
    const float ext = tk.chi[a] - tk.clo[a];
    on[a] = ext > 0.0f;
    scale[a] = on[a] ? (float)kBins / ext : 0.0f;
  }
}

__global__ __launch_bounds__(kT) void k_top_bin(const TopTask* __restrict__ tasks, const uint32_t* __restrict__ block_task,
                                                const uint32_t* __restrict__ idx, const float4* __restrict__ tlo,
                                                const float4* __restrict__ thi, const float4* __restrict__ cen,
                                                uint32_t* __restrict__ bins) {
  __shared__ uint32_t sb[3 * kBins * kTopWords];
  const uint32_t ti = block_task[blockIdx.x];
  const TopTask tk = tasks[ti];
  const uint32_t chunk = blockIdx.x - tk.block0;
  for (int w = threadIdx.x; w < 3 * kBins * kTopWords; w += kT) {
    const int f = w % kTopWords;
    sb[w] = f == 0 ? 0u : ((f - 1) % 6 < 3 ? kOrdPosInf : kOrdNegInf);
  }
  __syncthreads();
  bool on[3];
  float scale[3];
  task_scales(tk, on, scale);
  const uint32_t beg = chunk * kChunk, end = min(tk.count, beg + kChunk);
  for (uint32_t j = beg + threadIdx.x; j < end; j += kT) {
    const uint32_t t = idx[tk.first + j];
    const float4 c4 = cen[t], l4 = tlo[t], h4 = thi[t];
    const float c[3] = {c4.x, c4.y, c4.z}, l[3] = {l4.x, l4.y, l4.z}, h[3] = {h4.x, h4.y, h4.z};
    for (int a = 0; a < 3; ++a) {
      if (!on[a]) continue;
      uint32_t* r = sb + (a * kBins + bin_of(c[a], tk.clo[a], scale[a])) * kTopWords;
      atomicAdd(&r[0], 1u);
      for (int k = 0; k < 3; ++k) {
        atomicMin(&r[1 + k], f2o(l[k]));
        atomicMax(&r[4 + k], f2o(h[k]));
        atomicMin(&r[7 + k], f2o(c[k]));
        atomicMax(&r[10 + k], f2o(c[k]));
      }
    }
  }
  __syncthreads();
  uint32_t* gb = bins + (size_t)ti * 3 * kBins * kTopWords;
  for (int w = threadIdx.x; w < 3 * kBins * kTopWords; w += kT) {
    const int f = w % kTopWords;
    if (sb[w - f] == 0) continue;  // empty bin
    if (f == 0)
      atomicAdd(&gb[w], sb[w]);
    else if ((f - 1) % 6 < 3)
      atomicMin(&gb[w], sb[w]);
    else
      atomicMax(&gb[w], sb[w]);
  }
}

__global__ void k_init_bins(uint32_t* __restrict__ bins, size_t words) {
  const size_t w = (size_t)blockIdx.x * kT + threadIdx.x;
  if (w >= words) return;
  const int f = (int)(w % kTopWords);
  bins[w] = f == 0 ? 0u : ((f - 1) % 6 < 3 ? kOrdPosInf : kOrdNegInf);
}

// SAH over the 3 x 31 splits of one node, in the host build's order (axis,
// then split position; the first strictly smaller cost wins).
__global__ __launch_bounds__(128) void k_top_split(const TopTask* __restrict__ tasks, const uint32_t* __restrict__ bins,
                                                   TopSplit* __restrict__ splits) {
  __shared__ double cost[3 * kBins];
  const TopTask tk = tasks[blockIdx.x];
  const uint32_t* gb = bins + (size_t)blockIdx.x * 3 * kBins * kTopWords;
  const int t = threadIdx.x;
  if (t < 3 * kBins) {
    const int a = t / kBins, s = t % kBins;
    double c = INFINITY;
    if (s >= 1 && tk.chi[a] - tk.clo[a] > 0.0f) {
      float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t nl = 0, nr = 0;
      for (int b = 0; b < kBins; ++b) {
        const uint32_t* r = gb + (a * kBins + b) * kTopWords;
        if (r[0] == 0) continue;
        float* lo = b < s ? llo : rlo;
        float* hi = b < s ? lhi : rhi;
        (b < s ? nl : nr) += r[0];
        for (int k = 0; k < 3; ++k) {
          lo[k] = fminf(lo[k], o2f(r[1 + k]));
          hi[k] = fmaxf(hi[k], o2f(r[4 + k]));
        }
      }
      if (nl && nr) {
        const double parent = fmax(box_area(tk.lo, tk.hi), 1e-30);
        c = 1.0 + (box_area(llo, lhi) * nl + box_area(rlo, rhi) * nr) / parent;
      }
    }
    cost[t] = c;
  }
  __syncthreads();
  if (t == 0) {
    double best = INFINITY;
    int ba = -1, bs = -1;
    for (int a = 0; a < 3; ++a)
      for (int s = 1; s < kBins; ++s)
        if (cost[a * kBins + s] < best) {
          best = cost[a * kBins + s];
          ba = a;
          bs = s;
        }
    TopSplit sp{};
    sp.axis = ba;
    sp.split = (uint32_t)bs;
    for (int k = 0; k < 12; ++k) {
      sp.left[k] = k % 6 < 3 ? kOrdPosInf : kOrdNegInf;
      sp.right[k] = sp.left[k];
    }
    if (ba >= 0) {
      for (int b = 0; b < kBins; ++b) {
        const uint32_t* r = gb + (ba * kBins + b) * kTopWords;
        if (r[0] == 0) continue;
        uint32_t* dst = b < bs ? sp.left : sp.right;
        (b < bs ? sp.nleft : sp.nright) += r[0];
        for (int k = 0; k < 12; ++k) dst[k] = k % 6 < 3 ? min(dst[k], r[1 + k]) : max(dst[k], r[1 + k]);
      }
    }
    splits[blockIdx.x] = sp;
  }
}

__device__ __forceinline__ bool top_left(const TopTask& tk, const TopSplit& sp, const float4& c4, const bool on[3],
                                         const float scale[3]) {
  const float c = sp.axis == 0 ? c4.x : (sp.axis == 1 ? c4.y : c4.z);
  (void)on;
  return bin_of(c, tk.clo[sp.axis], scale[sp.axis]) < (int)sp.split;
}

__global__ __launch_bounds__(kT) void k_top_count(const TopTask* __restrict__ tasks, const uint32_t* __restrict__ block_task,
                                                  const TopSplit* __restrict__ splits, const uint32_t* __restrict__ idx,
                                                  const float4* __restrict__ cen, uint32_t* __restrict__ chunk_left) {
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const uint32_t ti = block_task[blockIdx.x];
  const TopTask tk = tasks[ti];
  const TopSplit sp = splits[ti];
  bool on[3];
  float scale[3];
  task_scales(tk, on, scale);
  const uint32_t chunk = blockIdx.x - tk.block0;
  const uint32_t beg = chunk * kChunk, end = min(tk.count, beg + kChunk);
  uint32_t c = 0;
  for (uint32_t j = beg + threadIdx.x; j < end; j += kT) c += top_left(tk, sp, cen[idx[tk.first + j]], on, scale);
  atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) chunk_left[blockIdx.x] = cnt;
}

// Exclusive offsets of each block's left count within its task; checks the
// sum against the bins' count (a mismatch sets flags bit 1).
__global__ void k_top_offsets(const TopTask* __restrict__ tasks, uint32_t ntasks, const TopSplit* __restrict__ splits,
                              const uint32_t* __restrict__ chunk_left, uint32_t* __restrict__ chunk_off,
                              uint32_t* __restrict__ flags) {
  const uint32_t ti = blockIdx.x * kT + threadIdx.x;
  if (ti >= ntasks) return;
  const TopTask tk = tasks[ti];
  const uint32_t nb = (tk.count + kChunk - 1) / kChunk;
  uint32_t run = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    chunk_off[tk.block0 + b] = run;
    run += chunk_left[tk.block0 + b];
  }
  if (run != splits[ti].nleft) atomicOr(flags, 2u);
}

__global__ __launch_bounds__(kT) void k_top_scatter(const TopTask* __restrict__ tasks, const uint32_t* __restrict__ block_task,
                                                    const TopSplit* __restrict__ splits, const uint32_t* __restrict__ idx,
                                                    const float4* __restrict__ cen, const uint32_t* __restrict__ chunk_off,
                                                    uint32_t* __restrict__ tmp, uint32_t* __restrict__ flags) {
  __shared__ uint32_t wtot[kT / 64];
  const uint32_t ti = block_task[blockIdx.x];
  const TopTask tk = tasks[ti];
  const TopSplit sp = splits[ti];
  bool on[3];
  float scale[3];
  task_scales(tk, on, scale);
  const uint32_t chunk = blockIdx.x - tk.block0;
  const uint32_t beg = chunk * kChunk, end = min(tk.count, beg + kChunk);
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t run = chunk_off[blockIdx.x];
  for (uint32_t r0 = beg; r0 < end; r0 += kT) {
    const uint32_t j = r0 + threadIdx.x;
    const bool valid = j < end;
    uint32_t t = 0;
    bool left = false;
    if (valid) {
      t = idx[tk.first + j];
      left = top_left(tk, sp, cen[t], on, scale);
    }
    const unsigned long long m = __ballot(left);
    const uint32_t before_lane = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wtot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before_wave = 0, total = 0;
    for (uint32_t w = 0; w < kT / 64; ++w) {
      before_wave += w < wave ? wtot[w] : 0u;
      total += wtot[w];
    }
    if (valid) {
      const uint32_t left_before = run + before_wave + before_lane;
      const uint32_t pos = left ? left_before : sp.nleft + (j - left_before);
      if (pos < tk.count)
        tmp[tk.first + pos] = t;
      else
        atomicOr(flags, 2u);  // counts disagree with the bins (reported, never written out of range)
    }
    run += total;
    __syncthreads();
  }
}

__global__ void k_top_copy(const TopTask* __restrict__ tasks, const uint32_t* __restrict__ block_task,
                           const uint32_t* __restrict__ tmp, uint32_t* __restrict__ idx) {
  const TopTask tk = tasks[block_task[blockIdx.x]];
  const uint32_t chunk = blockIdx.x - tk.block0;
  const uint32_t beg = chunk * kChunk, end = min(tk.count, beg + kChunk);
  for (uint32_t j = beg + threadIdx.x; j < end; j += kT) idx[tk.first + j] = tmp[tk.first + j];
}

__global__ void k_write_nodes(const uint32_t* __restrict__ ids, const BNode* __restrict__ recs, uint32_t n,
                              BNode* __restrict__ bn) {
  const uint32_t i = blockIdx.x * kT + threadIdx.x;
  if (i < n) bn[ids[i]] = recs[i];
}

// ------------------------------------------------------------------ bottom phase
struct SmallTask {
  uint32_t node, first, count, base;  // base: first id of the subtree's other nodes
};

__device__ __forceinline__ void write_node(BNode* bn, uint32_t id, const float lo[3], const float hi[3], uint32_t a,
                                           uint32_t b) {
  BNode n;
  n.lo = make_float4(lo[0], lo[1], lo[2], __uint_as_float(a));
  n.hi = make_float4(hi[0], hi[1], hi[2], __uint_as_float(b));
  bn[id] = n;
}

__global__ __launch_bounds__(kT) void k_small(const SmallTask* __restrict__ tasks, uint32_t* __restrict__ idx,
                                              const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                              const float4* __restrict__ cen, BNode* __restrict__ bn) {
  __shared__ uint32_t ids[kSmall], tmp[kSmall];
  __shared__ uint32_t stack[kSmall][3];
  __shared__ uint32_t sb[3 * kBins * kSmallWords];
  __shared__ uint32_t red[12];
  __shared__ double cost[3 * kBins];
  __shared__ uint32_t scan[kT];
  __shared__ uint32_t cur[3], sp_top, next_local, done;
  __shared__ int best_axis, best_split, make_leaf;
  const SmallTask tk = tasks[blockIdx.x];
  for (uint32_t j = threadIdx.x; j < tk.count; j += kT) ids[j] = idx[tk.first + j];
  if (threadIdx.x == 0) {
    stack[0][0] = tk.node;
    stack[0][1] = 0;
    stack[0][2] = tk.count;
    sp_top = 1;
    next_local = 0;
  }
  __syncthreads();
  for (;;) {
    if (threadIdx.x == 0) {
      done = sp_top == 0;
      if (!done) {
        --sp_top;
        cur[0] = stack[sp_top][0];
        cur[1] = stack[sp_top][1];
        cur[2] = stack[sp_top][2];
      }
    }
    if (threadIdx.x < 12) red[threadIdx.x] = threadIdx.x % 6 < 3 ? kOrdPosInf : kOrdNegInf;
    __syncthreads();
    if (done) break;
    const uint32_t node = cur[0], f = cur[1], c = cur[2];
    // Bounds of the node's triangles and centroids.
    {
      float b[12] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY,
                     INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      bool any = false;
      for (uint32_t j = threadIdx.x; j < c; j += kT) {
        const uint32_t t = ids[f + j];
        const float4 l = tlo[t], h = thi[t], q = cen[t];
        b[0] = fminf(b[0], l.x), b[1] = fminf(b[1], l.y), b[2] = fminf(b[2], l.z);
        b[3] = fmaxf(b[3], h.x), b[4] = fmaxf(b[4], h.y), b[5] = fmaxf(b[5], h.z);
        b[6] = fminf(b[6], q.x), b[7] = fminf(b[7], q.y), b[8] = fminf(b[8], q.z);
        b[9] = fmaxf(b[9], q.x), b[10] = fmaxf(b[10], q.y), b[11] = fmaxf(b[11], q.z);
        any = true;
      }
      if (any)
        for (int k = 0; k < 12; ++k) {
          if (k % 6 < 3)
            atomicMin(&red[k], f2o(b[k]));
          else
            atomicMax(&red[k], f2o(b[k]));
        }
    }
    for (int w = threadIdx.x; w < 3 * kBins * kSmallWords; w += kT) {
      const int fw = w % kSmallWords;
      sb[w] = fw == 0 ? 0u : (fw <= 3 ? kOrdPosInf : kOrdNegInf);
    }
    __syncthreads();
    float lo[3], hi[3], clo[3], chi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = o2f(red[k]);
      hi[k] = o2f(red[3 + k]);
      clo[k] = o2f(red[6 + k]);
      chi[k] = o2f(red[9 + k]);
    }
    if (c <= 1) {
      if (threadIdx.x == 0) write_node(bn, node, lo, hi, tk.first + f, c | kLeafBit);
      __syncthreads();
      continue;
    }
    bool on[3];
    float scale[3];
    for (int a = 0; a < 3; ++a) {
      const float ext = chi[a] - clo[a];
      on[a] = ext > 0.0f;
      scale[a] = on[a] ? (float)kBins / ext : 0.0f;
    }
    for (uint32_t j = threadIdx.x; j < c; j += kT) {
      const uint32_t t = ids[f + j];
      const float4 l4 = tlo[t], h4 = thi[t], q4 = cen[t];
      const float q[3] = {q4.x, q4.y, q4.z}, l[3] = {l4.x, l4.y, l4.z}, h[3] = {h4.x, h4.y, h4.z};
      for (int a = 0; a < 3; ++a) {
        if (!on[a]) continue;
        uint32_t* r = sb + (a * kBins + bin_of(q[a], clo[a], scale[a])) * kSmallWords;
        atomicAdd(&r[0], 1u);
        for (int k = 0; k < 3; ++k) {
          atomicMin(&r[1 + k], f2o(l[k]));
          atomicMax(&r[4 + k], f2o(h[k]));
        }
      }
    }
    __syncthreads();
    if (threadIdx.x < 3 * kBins) {
      const int a = threadIdx.x / kBins, s = threadIdx.x % kBins;
      double cc = INFINITY;
      if (s >= 1 && on[a]) {
        float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t nl = 0, nr = 0;
        for (int b = 0; b < kBins; ++b) {
          const uint32_t* r = sb + (a * kBins + b) * kSmallWords;
          if (r[0] == 0) continue;
          float* blo = b < s ? llo : rlo;
          float* bhi = b < s ? lhi : rhi;
          (b < s ? nl : nr) += r[0];
          for (int k = 0; k < 3; ++k) {
            blo[k] = fminf(blo[k], o2f(r[1 + k]));
            bhi[k] = fmaxf(bhi[k], o2f(r[4 + k]));
          }
        }
        if (nl && nr) {
          const double parent = fmax(box_area(lo, hi), 1e-30);
          cc = 1.0 + (box_area(llo, lhi) * nl + box_area(rlo, rhi) * nr) / parent;
        }
      }
      cost[threadIdx.x] = cc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double best = INFINITY;
      int ba = -1, bs = -1;
      for (int a = 0; a < 3; ++a)
        for (int s = 1; s < kBins; ++s)
          if (cost[a * kBins + s] < best) {
            best = cost[a * kBins + s];
            ba = a;
            bs = s;
          }
      best_axis = ba;
      best_split = bs;
      make_leaf = c <= kMaxLeaf && (double)c <= best;
    }
    __syncthreads();
    if (make_leaf) {
      if (threadIdx.x == 0) write_node(bn, node, lo, hi, tk.first + f, c | kLeafBit);
      __syncthreads();
      continue;
    }
    // Stable partition: thread t owns elements [4t, 4t + 4) of the node.
    const int ax = best_axis;
    uint32_t flags = 0, nleft_mine = 0;
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t j = 4 * threadIdx.x + k;
      if (j >= c) break;
      bool left;
      if (ax >= 0) {
        const float4 q4 = cen[ids[f + j]];
        const float q = ax == 0 ? q4.x : (ax == 1 ? q4.y : q4.z);
        left = bin_of(q, clo[ax], scale[ax]) < best_split;
      } else {
        left = j < c / 2;  // all centroids coincide: halves in the current order
      }
      flags |= (uint32_t)left << k;
      nleft_mine += left;
    }
    scan[threadIdx.x] = nleft_mine;
    __syncthreads();
    for (int off = 1; off < kT; off <<= 1) {  // inclusive scan
      const uint32_t v = threadIdx.x >= (uint32_t)off ? scan[threadIdx.x - off] : 0u;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    const uint32_t nleft = scan[kT - 1];
    uint32_t lpos = scan[threadIdx.x] - nleft_mine;
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t j = 4 * threadIdx.x + k;
      if (j >= c) break;
      const uint32_t pos = (flags >> k & 1u) ? lpos++ : nleft + (j - lpos);
      tmp[pos] = ids[f + j];
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < c; j += kT) ids[f + j] = tmp[j];
    if (threadIdx.x == 0) {
      const uint32_t l = tk.base + next_local, r = l + 1;
      next_local += 2;
      write_node(bn, node, lo, hi, l, r);
      stack[sp_top][0] = l;
      stack[sp_top][1] = f;
      stack[sp_top][2] = nleft;
      stack[sp_top + 1][0] = r;
      stack[sp_top + 1][1] = f + nleft;
      stack[sp_top + 1][2] = c - nleft;
      sp_top += 2;
    }
    __syncthreads();
  }
  for (uint32_t j = threadIdx.x; j < tk.count; j += kT) idx[tk.first + j] = ids[j];
}

// ------------------------------------------------------------------ collapse
__device__ int gather(const BNode* __restrict__ bn, uint32_t root, uint32_t ch[8]) {
  const BNode r = bn[root];
  if (bn_leaf(r)) {
    ch[0] = root;
    return 1;
  }
  int n = 2;
  ch[0] = __float_as_uint(r.lo.w);
  ch[1] = __float_as_uint(r.hi.w);
  while (n < 8) {
    int pick = -1;
    double best = -1.0;
    for (int i = 0; i < n; ++i) {
      const BNode c = bn[ch[i]];
      if (bn_leaf(c)) continue;
      const float lo[3] = {c.lo.x, c.lo.y, c.lo.z}, hi[3] = {c.hi.x, c.hi.y, c.hi.z};
      const double a = box_area(lo, hi);
      if (a > best) {
        best = a;
        pick = i;
      }
    }
    if (pick < 0) break;
    const BNode o = bn[ch[pick]];
    ch[pick] = __float_as_uint(o.lo.w);
    ch[n++] = __float_as_uint(o.hi.w);
  }
  return n;
}

__global__ void k_collapse_count(const uint32_t* __restrict__ entries, uint32_t n, const BNode* __restrict__ bn,
                                 uint32_t* __restrict__ inner, uint32_t* __restrict__ tris) {
  const uint32_t e = blockIdx.x * kT + threadIdx.x;
  if (e >= n) return;
  uint32_t ch[8];
  const int nc = gather(bn, entries[e], ch);
  uint32_t ni = 0, nt = 0;
  for (int i = 0; i < nc; ++i) {
    const BNode c = bn[ch[i]];
    if (bn_leaf(c))
      nt += __float_as_uint(c.hi.w) & ~kLeafBit;
    else
      ++ni;
  }
  inner[e] = ni;
  tris[e] = nt;
}

__device__ int exponent_for(double ext) {
  if (!(ext > 0.0)) return 1;
  int x;
  const double m = frexp(ext / 255.0, &x);
  const int e = (m == 0.5) ? x - 1 : x;
  return min(max(e + 127, 1), 254);
}

__global__ void k_collapse_write(const uint32_t* __restrict__ entries, uint32_t n, const BNode* __restrict__ bn,
                                 const uint32_t* __restrict__ inner_off, const uint32_t* __restrict__ tri_off,
                                 uint32_t level_start, uint32_t child_start, uint32_t tri_start,
                                 const uint32_t* __restrict__ idx, const float* __restrict__ pos,
                                 Node8* __restrict__ nodes, uint32_t* __restrict__ next, LeafTri* __restrict__ ltris,
                                 uint32_t* __restrict__ flags) {
  const uint32_t e = blockIdx.x * kT + threadIdx.x;
  if (e >= n) return;
  uint32_t ch[8];
  const int nc = gather(bn, entries[e], ch);
  BNode cb[8];
  float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < nc; ++i) {
    cb[i] = bn[ch[i]];
    nlo[0] = fminf(nlo[0], cb[i].lo.x), nlo[1] = fminf(nlo[1], cb[i].lo.y), nlo[2] = fminf(nlo[2], cb[i].lo.z);
    nhi[0] = fmaxf(nhi[0], cb[i].hi.x), nhi[1] = fmaxf(nhi[1], cb[i].hi.y), nhi[2] = fmaxf(nhi[2], cb[i].hi.z);
  }
  const double cx = 0.5 * ((double)nlo[0] + nhi[0]), cy = 0.5 * ((double)nlo[1] + nhi[1]),
               cz = 0.5 * ((double)nlo[2] + nhi[2]);
  int slot_of[8];
  bool used[8] = {false, false, false, false, false, false, false, false};
  bool done[8] = {false, false, false, false, false, false, false, false};
  for (int r = 0; r < nc; ++r) {
    double best = -INFINITY;
    int bi = -1, bs = -1;
    for (int i = 0; i < nc; ++i) {
      if (done[i]) continue;
      const double dx = 0.5 * ((double)cb[i].lo.x + cb[i].hi.x) - cx;
      const double dy = 0.5 * ((double)cb[i].lo.y + cb[i].hi.y) - cy;
      const double dz = 0.5 * ((double)cb[i].lo.z + cb[i].hi.z) - cz;
      for (int s = 0; s < 8; ++s) {
        if (used[s]) continue;
        const double c = ((s & 1) ? dx : -dx) + ((s & 2) ? dy : -dy) + ((s & 4) ? dz : -dz);
        if (c > best) {
          best = c;
          bi = i;
          bs = s;
        }
      }
    }
    done[bi] = true;
    used[bs] = true;
    slot_of[bi] = bs;
  }
  int child_at[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  for (int i = 0; i < nc; ++i) child_at[slot_of[i]] = i;

  Node8 nd;
  memset(&nd, 0, sizeof nd);
  nd.p[0] = nlo[0];
  nd.p[1] = nlo[1];
  nd.p[2] = nlo[2];
  nd.child_base = child_start + inner_off[e];
  nd.tri_base = tri_start + tri_off[e];
  uint32_t rank = 0, toff = 0;
  for (int s = 0; s < 8; ++s) {
    if (child_at[s] < 0) continue;
    const BNode& c = cb[child_at[s]];
    if (!bn_leaf(c)) {
      nd.imask |= (uint8_t)(1u << s);
      nd.meta[s] = 0x1F;
      next[nd.child_base - child_start + rank] = ch[child_at[s]];
      ++rank;
    } else {
      const uint32_t cnt = __float_as_uint(c.hi.w) & ~kLeafBit, first = __float_as_uint(c.lo.w);
      nd.meta[s] = (uint8_t)((((1u << cnt) - 1u) << 5) | toff);
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t t = idx[first + k];
        LeafTri lt;
        const float* v = pos + (size_t)t * 9;
        lt.v0[0] = v[0], lt.v0[1] = v[1], lt.v0[2] = v[2];
        lt.v1[0] = v[3], lt.v1[1] = v[4], lt.v1[2] = v[5];
        lt.v2[0] = v[6], lt.v2[1] = v[7], lt.v2[2] = v[8];
        lt.id = t;
        lt.pad1 = lt.pad2 = 0;
        ltris[nd.tri_base + toff + k] = lt;
      }
      toff += cnt;
    }
  }
  // Quantise: fma(qlo, scale, p) <= child lo and fma(qhi, scale, p) >= child hi.
  for (int a = 0; a < 3; ++a) {
    int eb = exponent_for((double)nhi[a] - (double)nlo[a]);
    for (;;) {
      const float sc = __uint_as_float((uint32_t)eb << 23);
      bool ok = true;
      for (int s = 0; s < 8 && ok; ++s) {
        if (child_at[s] < 0) {
          nd.qlo[a][s] = 0;
          nd.qhi[a][s] = 0;
          continue;
        }
        const BNode& c = cb[child_at[s]];
        const float clo = a == 0 ? c.lo.x : (a == 1 ? c.lo.y : c.lo.z);
        const float chi = a == 0 ? c.hi.x : (a == 1 ? c.hi.y : c.hi.z);
        const double ql = floor(((double)clo - (double)nd.p[a]) / (double)sc);
        const double qh = ceil(((double)chi - (double)nd.p[a]) / (double)sc);
        int il = (int)fmin(fmax(ql, 0.0), 255.0);
        int ih = (int)fmin(fmax(qh, 0.0), 255.0);
        while (il > 0 && fmaf((float)il, sc, nd.p[a]) > clo) il--;
        while (ih < 255 && fmaf((float)ih, sc, nd.p[a]) < chi) ih++;
        if (fmaf((float)il, sc, nd.p[a]) > clo || fmaf((float)ih, sc, nd.p[a]) < chi) {
          ok = false;
          break;
        }
        nd.qlo[a][s] = (uint8_t)il;
        nd.qhi[a][s] = (uint8_t)ih;
      }
      if (ok) {
        nd.e[a] = (uint8_t)eb;
        break;
      }
      if (eb >= 254) {
        atomicOr(flags, 4u);
        break;
      }
      eb++;
    }
  }
  nodes[level_start + e] = nd;
}

// Device buffer freed on scope exit.
struct Buf {
  void* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(size_t n) {
    if (p) (void)hipFree(p);
    p = nullptr;
    hip(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

unsigned grid(size_t n) { return (unsigned)((n + kT - 1) / kT); }

void build(const float* d_pos, uint32_t n, hipStream_t s, GpuBvh& out) {
  Buf b_tlo, b_thi, b_cen, b_idx, b_tmp, b_flags, b_bn, b_root;
  float4* tlo = b_tlo.alloc<float4>(n);
  float4* thi = b_thi.alloc<float4>(n);
  float4* cen = b_cen.alloc<float4>(n);
  uint32_t* idx = b_idx.alloc<uint32_t>(n);
  uint32_t* tmp = b_tmp.alloc<uint32_t>(n);
  uint32_t* flags = b_flags.alloc<uint32_t>(1);
  BNode* bn = b_bn.alloc<BNode>(2 * (size_t)n);
  uint32_t* root = b_root.alloc<uint32_t>(12);
  hip(hipMemsetAsync(flags, 0, 4, s), "hipMemsetAsync");
  hipLaunchKernelGGL(k_prep, dim3(grid(n)), dim3(kT), 0, s, d_pos, n, tlo, thi, cen, flags, idx);
  const uint32_t init[12] = {kOrdPosInf, kOrdPosInf, kOrdPosInf, kOrdNegInf, kOrdNegInf, kOrdNegInf,
                             kOrdPosInf, kOrdPosInf, kOrdPosInf, kOrdNegInf, kOrdNegInf, kOrdNegInf};
  hip(hipMemcpyAsync(root, init, sizeof init, hipMemcpyHostToDevice, s), "upload");
  hipLaunchKernelGGL(k_root_bounds, dim3(std::min<unsigned>(grid(n), 1024)), dim3(kT), 0, s, tlo, thi, cen, n, root);
  hip(hipGetLastError(), "prep launch");
  uint32_t hroot[12], hflags = 0;
  hip(hipMemcpyAsync(hroot, root, sizeof hroot, hipMemcpyDeviceToHost, s), "download");
  hip(hipMemcpyAsync(&hflags, flags, 4, hipMemcpyDeviceToHost, s), "download");
  hip(hipStreamSynchronize(s), "sync");
  if (hflags & 1u) throw Fail("triangle vertex is not finite");

  // ---- top phase
  std::vector<SmallTask> small;
  std::vector<uint32_t> top_ids;
  std::vector<BNode> top_recs;
  uint32_t next_id = 1;
  std::vector<TopTask> tasks;
  if (n > kSmall) {
    TopTask t{};
    t.node = 0;
    t.first = 0;
    t.count = n;
    for (int a = 0; a < 3; ++a) {
      t.lo[a] = o2f_host(hroot[a]);
      t.hi[a] = o2f_host(hroot[3 + a]);
      t.clo[a] = o2f_host(hroot[6 + a]);
      t.chi[a] = o2f_host(hroot[9 + a]);
    }
    tasks.push_back(t);
  } else {
    small.push_back({0, 0, n, 0});
  }
  Buf b_tasks, b_btask, b_bins, b_splits, b_cleft, b_coff, b_ids, b_recs;
  while (!tasks.empty()) {
    std::vector<uint32_t> block_task;
    for (uint32_t i = 0; i < tasks.size(); ++i) {
      tasks[i].block0 = (uint32_t)block_task.size();
      const uint32_t nb = (tasks[i].count + kChunk - 1) / kChunk;
      block_task.insert(block_task.end(), nb, i);
    }
    const uint32_t nt = (uint32_t)tasks.size(), nb = (uint32_t)block_task.size();
    TopTask* d_tasks = b_tasks.alloc<TopTask>(nt);
    uint32_t* d_btask = b_btask.alloc<uint32_t>(nb);
    const size_t words = (size_t)nt * 3 * kBins * kTopWords;
    uint32_t* d_bins = b_bins.alloc<uint32_t>(words);
    TopSplit* d_splits = b_splits.alloc<TopSplit>(nt);
    uint32_t* d_cleft = b_cleft.alloc<uint32_t>(nb);
    uint32_t* d_coff = b_coff.alloc<uint32_t>(nb);
    hip(hipMemcpyAsync(d_tasks, tasks.data(), nt * sizeof(TopTask), hipMemcpyHostToDevice, s), "upload");
    hip(hipMemcpyAsync(d_btask, block_task.data(), nb * 4, hipMemcpyHostToDevice, s), "upload");
    hipLaunchKernelGGL(k_init_bins, dim3(grid(words)), dim3(kT), 0, s, d_bins, words);
    hipLaunchKernelGGL(k_top_bin, dim3(nb), dim3(kT), 0, s, d_tasks, d_btask, idx, tlo, thi, cen, d_bins);
    hipLaunchKernelGGL(k_top_split, dim3(nt), dim3(128), 0, s, d_tasks, d_bins, d_splits);
    hipLaunchKernelGGL(k_top_count, dim3(nb), dim3(kT), 0, s, d_tasks, d_btask, d_splits, idx, cen, d_cleft);
    hipLaunchKernelGGL(k_top_offsets, dim3(grid(nt)), dim3(kT), 0, s, d_tasks, nt, d_splits, d_cleft, d_coff, flags);
    hipLaunchKernelGGL(k_top_scatter, dim3(nb), dim3(kT), 0, s, d_tasks, d_btask, d_splits, idx, cen, d_coff, tmp,
                       flags);
    hipLaunchKernelGGL(k_top_copy, dim3(nb), dim3(kT), 0, s, d_tasks, d_btask, tmp, idx);
    hip(hipGetLastError(), "top phase launch");
    std::vector<TopSplit> sp(nt);
    hip(hipMemcpyAsync(sp.data(), d_splits, nt * sizeof(TopSplit), hipMemcpyDeviceToHost, s), "download");
    hip(hipMemcpyAsync(&hflags, flags, 4, hipMemcpyDeviceToHost, s), "download");
    hip(hipStreamSynchronize(s), "sync");
    if (hflags & 2u) throw Fail("GPU BVH partition count mismatch");
    std::vector<TopTask> next;
    for (uint32_t i = 0; i < nt; ++i) {
      const TopTask& t = tasks[i];
      if (sp[i].axis < 0) throw Fail("all centroids of a large node coincide");
      const uint32_t l = next_id++, r = next_id++;
      BNode rec;
      rec.lo = make_float4(t.lo[0], t.lo[1], t.lo[2], 0.f);
      rec.hi = make_float4(t.hi[0], t.hi[1], t.hi[2], 0.f);
      std::memcpy(&rec.lo.w, &l, 4);
      std::memcpy(&rec.hi.w, &r, 4);
      top_ids.push_back(t.node);
      top_recs.push_back(rec);
      const uint32_t cnt[2] = {sp[i].nleft, sp[i].nright}, firsts[2] = {t.first, t.first + sp[i].nleft};
      const uint32_t* bounds[2] = {sp[i].left, sp[i].right};
      const uint32_t ids[2] = {l, r};
      for (int k = 0; k < 2; ++k) {
        if (cnt[k] > kSmall) {
          TopTask c{};
          c.node = ids[k];
          c.first = firsts[k];
          c.count = cnt[k];
          for (int a = 0; a < 3; ++a) {
            c.lo[a] = o2f_host(bounds[k][a]);
            c.hi[a] = o2f_host(bounds[k][3 + a]);
            c.clo[a] = o2f_host(bounds[k][6 + a]);
            c.chi[a] = o2f_host(bounds[k][9 + a]);
          }
          next.push_back(c);
        } else {
          small.push_back({ids[k], firsts[k], cnt[k], 0});
        }
      }
    }
    tasks.swap(next);
  }
  uint32_t base = next_id;
  for (SmallTask& t : small) {
    t.base = base;
    base += 2 * t.count - 2;
  }
  const uint32_t bnode_count = base;
  if (!top_ids.empty()) {
    uint32_t* d_ids = b_ids.alloc<uint32_t>(top_ids.size());
    BNode* d_recs = b_recs.alloc<BNode>(top_recs.size());
    hip(hipMemcpyAsync(d_ids, top_ids.data(), top_ids.size() * 4, hipMemcpyHostToDevice, s), "upload");
    hip(hipMemcpyAsync(d_recs, top_recs.data(), top_recs.size() * sizeof(BNode), hipMemcpyHostToDevice, s), "upload");
    hipLaunchKernelGGL(k_write_nodes, dim3(grid(top_ids.size())), dim3(kT), 0, s, d_ids, d_recs,
                       (uint32_t)top_ids.size(), bn);
  }
  Buf b_small;
  SmallTask* d_small = b_small.alloc<SmallTask>(small.size());
  hip(hipMemcpyAsync(d_small, small.data(), small.size() * sizeof(SmallTask), hipMemcpyHostToDevice, s), "upload");
  hipLaunchKernelGGL(k_small, dim3((unsigned)small.size()), dim3(kT), 0, s, d_small, idx, tlo, thi, cen, bn);
  hip(hipGetLastError(), "bottom phase launch");
  (void)bnode_count;

  // ---- collapse, level by level in breadth-first order
  Buf b_nodes, b_ltris, b_ent0, b_ent1, b_inner, b_tris, b_ioff, b_toff, b_scan;
  Node8* nodes = b_nodes.alloc<Node8>(n + 1);
  LeafTri* ltris = b_ltris.alloc<LeafTri>(n);
  uint32_t* ent[2] = {b_ent0.alloc<uint32_t>(n + 1), b_ent1.alloc<uint32_t>(n + 1)};
  uint32_t* inner = b_inner.alloc<uint32_t>(n + 1);
  uint32_t* trisc = b_tris.alloc<uint32_t>(n + 1);
  uint32_t* ioff = b_ioff.alloc<uint32_t>(n + 1);
  uint32_t* toff = b_toff.alloc<uint32_t>(n + 1);
  size_t scan_bytes = 0;
  hip(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, inner, ioff, (int)(n + 1), s), "scan size");
  void* scan_tmp = b_scan.alloc<uint8_t>(scan_bytes);
  const uint32_t zero = 0;
  hip(hipMemcpyAsync(ent[0], &zero, 4, hipMemcpyHostToDevice, s), "upload");
  uint32_t level_n = 1, level_start = 0, tri_start = 0, depth = 0, cur = 0;
  while (level_n) {
    hipLaunchKernelGGL(k_collapse_count, dim3(grid(level_n)), dim3(kT), 0, s, ent[cur], level_n, bn, inner, trisc);
    hip(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, inner, ioff, (int)level_n, s), "scan");
    hip(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, trisc, toff, (int)level_n, s), "scan");
    const uint32_t child_start = level_start + level_n;
    hipLaunchKernelGGL(k_collapse_write, dim3(grid(level_n)), dim3(kT), 0, s, ent[cur], level_n, bn, ioff, toff,
                       level_start, child_start, tri_start, idx, d_pos, nodes, ent[cur ^ 1], ltris, flags);
    hip(hipGetLastError(), "collapse launch");
    uint32_t last[4];
    hip(hipMemcpyAsync(&last[0], ioff + level_n - 1, 4, hipMemcpyDeviceToHost, s), "download");
    hip(hipMemcpyAsync(&last[1], inner + level_n - 1, 4, hipMemcpyDeviceToHost, s), "download");
    hip(hipMemcpyAsync(&last[2], toff + level_n - 1, 4, hipMemcpyDeviceToHost, s), "download");
    hip(hipMemcpyAsync(&last[3], trisc + level_n - 1, 4, hipMemcpyDeviceToHost, s), "download");
    hip(hipStreamSynchronize(s), "sync");
    ++depth;
    level_start = child_start;
    tri_start += last[2] + last[3];
    level_n = last[0] + last[1];
    cur ^= 1;
  }
  hip(hipMemcpyAsync(&hflags, flags, 4, hipMemcpyDeviceToHost, s), "download");
  hip(hipStreamSynchronize(s), "sync");
  if (hflags & 4u) throw Fail("scene extent too large to quantise");
  if (tri_start != n) throw Fail("GPU BVH lost triangles");
  // Exact-size results owned by the caller.
  hip(hipMalloc(&out.nodes, (size_t)level_start * sizeof(Node8)), "hipMalloc");
  hip(hipMalloc(&out.tris, (size_t)n * sizeof(LeafTri)), "hipMalloc");
  hip(hipMemcpyAsync(out.nodes, nodes, (size_t)level_start * sizeof(Node8), hipMemcpyDeviceToDevice, s), "copy");
  hip(hipMemcpyAsync(out.tris, ltris, (size_t)n * sizeof(LeafTri), hipMemcpyDeviceToDevice, s), "copy");
  hip(hipStreamSynchronize(s), "sync");
  out.node_count = level_start;
  out.depth = depth;
}

}  // namespace

bool build_bvh8_gpu(const float* d_pos, uint32_t tri_count, hipStream_t stream, GpuBvh& out, std::string& err) {
  out = GpuBvh{};
  if (tri_count == 0) return true;
  try {
    build(d_pos, tri_count, stream, out);
    return true;
  } catch (const Fail& f) {
    if (out.nodes) (void)hipFree(out.nodes);
    if (out.tris) (void)hipFree(out.tris);
    out = GpuBvh{};
    (void)hipStreamSynchronize(stream);
    err = f.what();
    return false;
  }
}

}  // namespace hpt
