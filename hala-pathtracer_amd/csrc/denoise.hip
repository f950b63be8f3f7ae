// Kernels of the learned denoiser (C ABI in include/hpt_nn.h): BF16 implicit-GEMM
// convolutions on MFMA with float32 accumulation (forward, data gradient, weight gradient) and
// the U-Net's elementwise steps.  The reference has no counterpart: its denoise.py shells out to
// Intel OIDN (denoise.py:7-30).
//
// Convolution as GEMM.  For a k x k convolution (zero padding k/2) over NHWC tensors,
//     out[p][n] = sum_t sum_c in[p + off(t)][c] * W[n][t][c],
// with p a pixel, t a tap and off(t) its neighbour inside the same image (zero outside).  One
// wave owns 16*MT pixels x 16*NT output channels and runs v_mfma_f32_16x16x32_bf16: A = 16
// pixels x 32 K-values, B = 32 K-values x 16 channels.  The K axis is the sequence of
// (tap, 8-channel segment) pairs; each of the four lane groups of a wave loads one pair per
// step (16 contiguous bytes of the pixel's channels, 16 contiguous bytes of the weight row), so
// A and B fill their K slots with the same (lane group, element) -> K assignment and the product
// does not depend on how the instruction orders K internally.  Any channel count that is a
// multiple of 8 fits; the last step of a K loop zero-fills the lane groups past the end.
// The data gradient is the same kernel over dy (float32, rounded to bf16 on load) with the
// transposed, spatially flipped weights.  The weight gradient contracts over pixels: one wave
// owns 16*MT output x 16*NT input channels of one tap over a chunk of pixels, writes a partial
// sum, and a second pass adds the chunks in order (no atomics: the result is deterministic).
#include <hip/hip_runtime.h>

#include <mutex>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "../../include/hpt_nn.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / 64;
constexpr int kWgradChunk = 1024;  // pixels per weight-gradient partial sum (a multiple of 32)
constexpr unsigned kXcds = 8;      // MI355X: workgroups are dealt round-robin over 8 XCDs, each with its own L2
#ifndef HPT_CONV_TILE
#define HPT_CONV_TILE 1
#endif
constexpr bool kConvTile = HPT_CONV_TILE;  // forward 3x3 convolutions stage their input tile in LDS
#ifndef HPT_CONV_FAST
#define HPT_CONV_FAST 1
#endif
constexpr bool kConvFast = HPT_CONV_FAST;  // ... in the kernel specialised on the channel count

// Block b runs on XCD b % 8; this maps it to tile (b % 8) * (grid / 8) + b / 8
// so that each XCD walks a contiguous range of pixel tiles and a tile's
// neighbouring image rows (the 3x3 halo) are in the same L2.  The grid is a
// multiple of 8 (xcd_grid), so the map is a bijection.
__device__ __forceinline__ unsigned xcd_block() {
  return (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
}
inline unsigned xcd_grid(unsigned blocks) { return (blocks + kXcds - 1) / kXcds * kXcds; }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bit pattern, round to nearest even (finite inputs).
__device__ __forceinline__ uint32_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_value(uint32_t b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) { return bf16_bits(lo) | (bf16_bits(hi) << 16); }

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }
__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ bf16x8 load8(const float* p) {
  const f32x4 a = *reinterpret_cast<const f32x4*>(p);
  const f32x4 b = *reinterpret_cast<const f32x4*>(p + 4);
  return __builtin_bit_cast(
      bf16x8, u32x4{pack_bf16(a.x, a.y), pack_bf16(a.z, a.w), pack_bf16(b.x, b.y), pack_bf16(b.z, b.w)});
}

// A convolution input given as concat(upsample2x(low), skip): channels
// [0, c_low) are those of the half-resolution tensor `low` at (y / 2, x / 2)
// (nearest upsampling), channels [c_low, cin) those of the full-resolution
// tensor passed as the input.  The operands are the same values as in the
// materialised concatenation, so the results are the same bits.
struct UpCat {
  const uint16_t* low;  // null: a plain input
  int c_low;
};

// ------------------------------------------------------------------ convolution
// FORWARD: y = act(out + bias) as bf16 and/or float32, `nstore` channels per pixel.
// Otherwise (data gradient): yf = out as float32, `nstore` (= the GEMM's N) channels.
template <typename InT, int KS, int MT, int NT, bool FORWARD, bool UPCAT = false>
__global__ __launch_bounds__(kThreads) void k_conv(const InT* __restrict__ in, const uint16_t* __restrict__ wgt,
                                                   const float* __restrict__ bias, uint16_t* __restrict__ y,
                                                   float* __restrict__ yf, int P, int h, int w, int cin, int nstore,
                                                   int relu, UpCat uc) {
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  const long long p0l = ((long long)xcd_block() * kWavesPerBlock + (threadIdx.x >> 6)) * (16 * MT);
  if (p0l >= P) return;  // uniform over the wave; no barriers below
  const int p0 = (int)p0l;
  const int n0 = blockIdx.y * (16 * NT);
  const int C8 = cin >> 3, S = KS * KS * C8, K = KS * KS * cin;

  int px[MT], py[MT], pimg[MT];
  bool pv[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) {
    const int p = p0 + 16 * i + m;
    pv[i] = p < P;
    const int pc = pv[i] ? p : 0;
    const int row = pc / w;
    px[i] = pc - row * w;
    py[i] = row % h;
    pimg[i] = row / h;
  }
  const int C8low = UPCAT ? uc.c_low >> 3 : 0, cskip = UPCAT ? cin - uc.c_low : 0;
  const uint16_t* wrow[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) wrow[j] = wgt + (size_t)(n0 + 16 * j + m) * K;

  f32x4 acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  int s = g, tap = g / C8, c8 = g - (g / C8) * C8;
  const int steps = (S + 3) >> 2;
  for (int step = 0; step < steps; ++step) {
    const bool sv = s < S;
    const int ty = tap / KS, tx = tap - ty * KS;
    const int oy = ty - KS / 2, ox = tx - KS / 2;
    bf16x8 a[MT], b[NT];
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const bool ok = sv && pv[i] && (unsigned)(py[i] + oy) < (unsigned)h && (unsigned)(px[i] + ox) < (unsigned)w;
      if (!ok)
        a[i] = zero8();
      else if (!UPCAT)
        a[i] = load8(in + (size_t)(p0 + 16 * i + m + oy * w + ox) * cin + c8 * 8);
      else if (c8 < C8low)
        a[i] = load8(uc.low + ((size_t)(pimg[i] * (h >> 1) + ((py[i] + oy) >> 1)) * (w >> 1) + ((px[i] + ox) >> 1)) *
                                  uc.c_low + c8 * 8);
      else
        a[i] = load8(in + (size_t)(p0 + 16 * i + m + oy * w + ox) * cskip + (c8 - C8low) * 8);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      if (sv)
        b[j] = load8(wrow[j] + s * 8);
      else
        b[j] = zero8();
    }
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    s += 4;
    c8 += 4;
    while (c8 >= C8) {
      c8 -= C8;
      ++tap;
    }
  }

  // Lane (m, g) holds D[4g + r][m]: pixel p0 + 16i + 4g + r, channel n0 + 16j + m.
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int co = n0 + 16 * j + m;
    if (co >= nstore) continue;
    const float bv = FORWARD ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int p = p0 + 16 * i + 4 * g + r;
        if (p >= P) continue;
        const size_t o = (size_t)p * nstore + co;
        if (FORWARD) {
          float v = acc[i][j][r] + bv;
          if (relu) v = fmaxf(v, 0.f);
          if (y) y[o] = (uint16_t)bf16_bits(v);
          if (yf) yf[o] = v;
        } else {
          yf[o] = acc[i][j][r];
        }
      }
  }
}

// Forward 3x3 convolution over bf16 NHWC with the block's input staged in
// LDS: a block owns a 16 x 16 pixel tile of one image and first copies the
// tile's 18 x 18 pixels (a one-pixel halo, zeros outside the image) of all
// input channels into LDS; wave q then computes tile rows 4q .. 4q + 3 (16
// pixels each) for 16*NT output channels, reading every tap's A fragment
// from LDS instead of through L1.  The K order per output is that of k_conv,
// so the results are the same bits.
constexpr int kTile = 16;
constexpr int kHalo = kTile + 2;

template <int NT, bool UPCAT = false>
__global__ __launch_bounds__(kThreads) void k_conv3_tile(const uint16_t* __restrict__ in,
                                                         const uint16_t* __restrict__ wgt,
                                                         const float* __restrict__ bias, uint16_t* __restrict__ y,
                                                         float* __restrict__ yf, int n_img, int h, int w, int cin,
                                                         int nstore, int relu, UpCat uc) {
  extern __shared__ __align__(16) uint16_t tile[];  // kHalo * kHalo pixels x cin
  const int tiles_x = (w + kTile - 1) / kTile, tiles_y = (h + kTile - 1) / kTile;
  const unsigned b = xcd_block();
  if (b >= (unsigned)(n_img * tiles_x * tiles_y)) return;  // the whole block, before any barrier
  const int img = (int)b / (tiles_x * tiles_y);
  const int rem = (int)b - img * tiles_x * tiles_y;
  const int ty0 = (rem / tiles_x) * kTile, tx0 = (rem - (rem / tiles_x) * tiles_x) * kTile;
  const int C8 = cin >> 3;
  for (int idx = threadIdx.x; idx < kHalo * kHalo * C8; idx += kThreads) {
    const int pix = idx / C8, c8 = idx - pix * C8;
    const int yy = ty0 - 1 + pix / kHalo, xx = tx0 - 1 + (pix - (pix / kHalo) * kHalo);
    u32x4 v = u32x4{0u, 0u, 0u, 0u};
    if ((unsigned)yy < (unsigned)h && (unsigned)xx < (unsigned)w) {
      const uint16_t* src;
      if (!UPCAT)
        src = in + ((size_t)(img * h + yy) * w + xx) * cin + c8 * 8;
      else if (c8 < (uc.c_low >> 3))
        src = uc.low + ((size_t)(img * (h >> 1) + (yy >> 1)) * (w >> 1) + (xx >> 1)) * uc.c_low + c8 * 8;
      else
        src = in + ((size_t)(img * h + yy) * w + xx) * (cin - uc.c_low) + (c8 - (uc.c_low >> 3)) * 8;
      v = *reinterpret_cast<const u32x4*>(src);
    }
    *reinterpret_cast<u32x4*>(tile + (size_t)pix * cin + c8 * 8) = v;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4, q = threadIdx.x >> 6;
  const int n0 = blockIdx.y * (16 * NT);
  const int S = 9 * C8, K = 9 * cin;
  const uint16_t* wrow[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) wrow[j] = wgt + (size_t)(n0 + 16 * j + m) * K;
  f32x4 acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  int s = g, tap = g / C8, c8 = g - (g / C8) * C8;
  const int steps = (S + 3) >> 2;
  for (int step = 0; step < steps; ++step) {
    const bool sv = s < S;
    const int ty = tap / 3, tx = tap - ty * 3;
    bf16x8 a[4], bb[NT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      a[i] = sv ? load8(tile + ((size_t)((4 * q + i + ty) * kHalo + m + tx)) * cin + c8 * 8) : zero8();
#pragma unroll
    for (int j = 0; j < NT; ++j) bb[j] = sv ? load8(wrow[j] + s * 8) : zero8();
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bb[j], acc[i][j], 0, 0, 0);
    s += 4;
    c8 += 4;
    while (c8 >= C8) {
      c8 -= C8;
      ++tap;
    }
  }

  // Lane (m, g) holds D[4g + r][m]: tile row 4q + i, column 4g + r, channel n0 + 16j + m.
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int co = n0 + 16 * j + m;
    if (co >= nstore) continue;
    const float bv = bias[co];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int yy = ty0 + 4 * q + i;
      if (yy >= h) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int xx = tx0 + 4 * g + r;
        if (xx >= w) continue;
        const size_t o = ((size_t)(img * h + yy) * w + xx) * nstore + co;
        float v = acc[i][j][r] + bv;
        if (relu) v = fmaxf(v, 0.f);
        if (y) y[o] = (uint16_t)bf16_bits(v);
        if (yf) yf[o] = v;
      }
    }
  }
}

// LDS a tiled convolution may take: two blocks per CU (160 KB of LDS).
constexpr size_t kTileLdsMax = 80 * 1024;

template <int NT, bool UPCAT>
bool allow_tile_lds() {  // above 64 KB a kernel must ask for its dynamic LDS
  static const bool ok = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3_tile<NT, UPCAT>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTileLdsMax) ==
                         hipSuccess;
  return ok;
}

// The tiled forward 3x3 convolution specialised on the input channel count
// (C8 = cin / 8), for the layer shapes of the network.  Same work split and
// K order as k_conv3_tile (so the same bits), with:
//  - the K loop unrolled at compile time: segment s = 4 * step + g of lane
//    group g is (tap s / C8, channel segment s % C8); the weights' address is
//    a per-lane base plus an immediate, the tile's a per-lane base plus a
//    per-step offset computed by constant divisions;
//  - the tile's pixel stride padded by 16 channels where cin is a multiple
//    of 32 (cin 32 and 64): measured, pixel strides of 8 mod 16 words read
//    without LDS bank conflicts and strides of 0 or 4 mod 16 words with;
//  - the MFMA operands swapped (weights first), so a lane's accumulator
//    holds 4 consecutive output channels of one pixel: bias, ReLU and the
//    bf16 rounding then end in one 8-byte store per 4 channels.
template <int C8>
struct TileGeom {
  static constexpr int kCin = 8 * C8;
  static constexpr int kStride = (C8 % 4 == 0) ? kCin + 16 : kCin;  // halves per tile pixel
  static constexpr int kSegs = 9 * C8;
  static constexpr int kSteps = (kSegs + 3) / 4;
  static constexpr size_t kLds = (size_t)kHalo * kHalo * kStride * sizeof(uint16_t);
};

//  - with POOL, the 2 x 2 average of the bf16 output (the next encoder
//    level's input, as k_pool_forward computes it) written beside it: a
//    pooling window is two rows of one wave and two neighbouring lanes.
template <int C8, int NT, bool UPCAT, bool POOL = false>
__global__ __launch_bounds__(kThreads) void k_conv3_fast(const uint16_t* __restrict__ in,
                                                         const uint16_t* __restrict__ wgt,
                                                         const float* __restrict__ bias, uint16_t* __restrict__ y,
                                                         float* __restrict__ yf, int n_img, int h, int w, int nstore,
                                                         int relu, UpCat uc, uint16_t* __restrict__ ypool) {
  using G = TileGeom<C8>;
  extern __shared__ __align__(16) uint16_t tile[];  // kHalo * kHalo pixels x kStride
  const int tiles_x = (w + kTile - 1) / kTile, tiles_y = (h + kTile - 1) / kTile;
  const unsigned b = xcd_block();
  if (b >= (unsigned)(n_img * tiles_x * tiles_y)) return;  // the whole block, before any barrier
  const int img = (int)b / (tiles_x * tiles_y);
  const int rem = (int)b - img * tiles_x * tiles_y;
  const int ty0 = (rem / tiles_x) * kTile, tx0 = (rem - (rem / tiles_x) * tiles_x) * kTile;
  const int c8low = UPCAT ? uc.c_low >> 3 : 0;
  auto fill_src = [&](int pix, int c8, bool& inside) -> const uint16_t* {
    const int yy = ty0 - 1 + pix / kHalo, xx = tx0 - 1 + (pix - (pix / kHalo) * kHalo);
    inside = (unsigned)yy < (unsigned)h && (unsigned)xx < (unsigned)w;
    if (!UPCAT) return in + ((size_t)(img * h + yy) * w + xx) * G::kCin + c8 * 8;
    if (c8 < c8low) return uc.low + ((size_t)(img * (h >> 1) + (yy >> 1)) * (w >> 1) + (xx >> 1)) * uc.c_low + c8 * 8;
    return in + ((size_t)(img * h + yy) * w + xx) * (G::kCin - uc.c_low) + (c8 - c8low) * 8;
  };
  constexpr int kItems = kHalo * kHalo * C8, kPer = (kItems + kThreads - 1) / kThreads;
  if constexpr (C8 >= 10) {
    // Wide tiles: every thread issues all of its loads, then stores them,
    // so the block waits for one round trip to memory instead of one per
    // item (13-18 items per thread).  Narrow tiles stage three items at a
    // time: all of theirs cost occupancy (measured slower for cin 32-48).
    u32x4 fill[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int idx = threadIdx.x + k * kThreads;
      const int pix = idx / C8, c8 = idx - pix * C8;
      bool inside;
      const uint16_t* s = fill_src(pix, c8, inside);
      fill[k] = (idx < kItems && inside) ? *reinterpret_cast<const u32x4*>(s) : u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int idx = threadIdx.x + k * kThreads;
      if (idx < kItems) {
        const int pix = idx / C8, c8 = idx - pix * C8;
        *reinterpret_cast<u32x4*>(tile + pix * G::kStride + c8 * 8) = fill[k];
      }
    }
  } else {
    // Three items in flight per round.
#pragma unroll
    for (int k0 = 0; k0 < kPer; k0 += 3) {
      u32x4 fill[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int idx = threadIdx.x + (k0 + k) * kThreads;
        const int pix = idx / C8, c8 = idx - pix * C8;
        bool inside;
        const uint16_t* s = fill_src(pix, c8, inside);
        fill[k] = (k0 + k < kPer && idx < kItems && inside) ? *reinterpret_cast<const u32x4*>(s)
                                                            : u32x4{0u, 0u, 0u, 0u};
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int idx = threadIdx.x + (k0 + k) * kThreads;
        if (k0 + k < kPer && idx < kItems) {
          const int pix = idx / C8, c8 = idx - pix * C8;
          *reinterpret_cast<u32x4*>(tile + pix * G::kStride + c8 * 8) = fill[k];
        }
      }
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4, q = threadIdx.x >> 6;
  const int n0 = blockIdx.y * (16 * NT);
  constexpr int K = 9 * G::kCin;
  const uint16_t* wbase[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) wbase[j] = wgt + (size_t)(n0 + 16 * j + m) * K + g * 8;
  const uint16_t* abase = tile + (4 * q * kHalo + m) * G::kStride;
  f32x4 acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Operands of step k: the weights (global, L2 latency) are loaded two
  // steps ahead and the tile fragments (LDS) one step ahead, so a step's
  // MFMAs do not wait on the loads issued just before them.
  auto seg_ok = [&](int step) { return (4 * step + 3 < G::kSegs) || 4 * step + g < G::kSegs; };
  auto load_b = [&](int step, bf16x8* bb) {
    const bool sv = seg_ok(step);
#pragma unroll
    for (int j = 0; j < NT; ++j) bb[j] = sv ? load8(wbase[j] + 32 * step) : zero8();
  };
  auto load_a = [&](int step, bf16x8* a) {
    const int s = 4 * step + g;
    const bool sv = seg_ok(step);
    const int tap = s / C8, c8 = s - (s / C8) * C8;
    const int ty = tap / 3, tx = tap - (tap / 3) * 3;
    const uint16_t* ap = abase + (ty * kHalo + tx) * G::kStride + c8 * 8;
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = sv ? load8(ap + i * kHalo * G::kStride) : zero8();
  };
  bf16x8 bq[3][NT], aq[2][4];
  load_b(0, bq[0]);
  if (G::kSteps > 1) load_b(1, bq[1]);
  load_a(0, aq[0]);
#pragma unroll
  for (int step = 0; step < G::kSteps; ++step) {
    if (step + 2 < G::kSteps) load_b(step + 2, bq[(step + 2) % 3]);
    if (step + 1 < G::kSteps) load_a(step + 1, aq[(step + 1) % 2]);
    // Keep the scheduler from sinking those loads to their uses.
    __builtin_amdgcn_sched_barrier(0);
    const bf16x8* bb = bq[step % 3];
    const bf16x8* a = aq[step % 2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb[j], a[i], acc[i][j], 0, 0, 0);
  }

  // Lane (m, g) holds D[4g + r][m]: channel n0 + 16j + 4g + r of tile row
  // 4q + i, column m.  Every lane runs the loops (the pooling shuffles need
  // them all); stores are guarded.
  const int xx = tx0 + m;
  const bool xok = xx < w;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int co = n0 + 16 * j + 4 * g;
    const bool cok = co < nstore;  // nstore is a multiple of 4 (checked at launch)
    const f32x4 bv = cok ? *reinterpret_cast<const f32x4*>(bias + co) : f32x4{0.f, 0.f, 0.f, 0.f};
    float top[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int yy = ty0 + 4 * q + i;
      const bool ok = xok && cok && yy < h;
      const size_t o = ((size_t)(img * h + yy) * w + xx) * nstore + co;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        v[r] = acc[i][j][r] + bv[r];
        if (relu) v[r] = fmaxf(v[r], 0.f);
      }
      if (ok && y) *reinterpret_cast<uint2*>(y + o) = make_uint2(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]));
      if (ok && yf) *reinterpret_cast<f32x4*>(yf + o) = f32x4{v[0], v[1], v[2], v[3]};
      if constexpr (POOL) {
        float cur[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) cur[r] = bf16_value(bf16_bits(v[r]));  // the stored value
        if ((i & 1) == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r) top[r] = cur[r];
        } else {
          // Window rows yy - 1, yy; columns xx, xx + 1 (lanes m, m + 1).
          float pv[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float top_r = __shfl_xor(top[r], 1), bot_r = __shfl_xor(cur[r], 1);
            pv[r] = 0.25f * ((top[r] + top_r) + (cur[r] + bot_r));
          }
          if (ok && (m & 1) == 0) {
            const size_t po = ((size_t)(img * (h >> 1) + (yy >> 1)) * (w >> 1) + (xx >> 1)) * nstore + co;
            *reinterpret_cast<uint2*>(ypool + po) = make_uint2(pack_bf16(pv[0], pv[1]), pack_bf16(pv[2], pv[3]));
          }
        }
      }
    }
  }
}

template <int C8, int NT, bool UPCAT, bool POOL>
bool allow_fast_lds() {  // above 64 KB a kernel must ask for its dynamic LDS
  static const bool ok = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3_fast<C8, NT, UPCAT, POOL>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTileLdsMax) ==
                         hipSuccess;
  return ok;
}

template <int C8, bool UPCAT, bool POOL>
int go_conv3_fast(hipStream_t st, dim3 grid, int nt, const uint16_t* in, const uint16_t* w, const float* b, uint16_t* y,
                  float* yf, int n_img, int h, int wd, int nstore, int relu, UpCat uc, uint16_t* ypool) {
  constexpr size_t lds = TileGeom<C8>::kLds;
  static_assert(lds <= kTileLdsMax, "tile exceeds the LDS budget");
  bool ok = true;
  if (lds > 65536)
    ok = nt == 1 ? allow_fast_lds<C8, 1, UPCAT, POOL>()
         : nt == 2 ? allow_fast_lds<C8, 2, UPCAT, POOL>()
         : nt == 3 ? allow_fast_lds<C8, 3, UPCAT, POOL>()
                   : allow_fast_lds<C8, 4, UPCAT, POOL>();
  if (!ok) return -1;
#define HPT_F(NT_)                                                                                                   \
  hipLaunchKernelGGL((k_conv3_fast<C8, NT_, UPCAT, POOL>), grid, dim3(kThreads), lds, st, in, w, b, y, yf, n_img, h, \
                     wd, nstore, relu, uc, ypool)
  switch (nt) {
    case 1: HPT_F(1); break;
    case 2: HPT_F(2); break;
    case 3: HPT_F(3); break;
    default: HPT_F(4); break;
  }
#undef HPT_F
  return (int)hipGetLastError();
}

// The specialised kernel for the channel counts it is built for (16, 32,
// 48, 64, 80, 112); -1 for others.  POOL (not with UPCAT): also the 2 x 2
// average pooling of the output into ypool.
template <bool UPCAT, bool POOL = false>
int launch_conv3_fast(hipStream_t st, dim3 grid, int nt, const uint16_t* in, const uint16_t* w, const float* b,
                      uint16_t* y, float* yf, int n_img, int h, int wd, int cin, int nstore, int relu, UpCat uc,
                      uint16_t* ypool = nullptr) {
  if (nstore % 4) return -1;
  switch (cin) {
    case 16: return go_conv3_fast<2, UPCAT, POOL>(st, grid, nt, in, w, b, y, yf, n_img, h, wd, nstore, relu, uc, ypool);
    case 32: return go_conv3_fast<4, UPCAT, POOL>(st, grid, nt, in, w, b, y, yf, n_img, h, wd, nstore, relu, uc, ypool);
    case 48: return go_conv3_fast<6, UPCAT, POOL>(st, grid, nt, in, w, b, y, yf, n_img, h, wd, nstore, relu, uc, ypool);
    case 64: return go_conv3_fast<8, UPCAT, POOL>(st, grid, nt, in, w, b, y, yf, n_img, h, wd, nstore, relu, uc, ypool);
    case 80: return go_conv3_fast<10, UPCAT, POOL>(st, grid, nt, in, w, b, y, yf, n_img, h, wd, nstore, relu, uc, ypool);
    case 112: return go_conv3_fast<14, UPCAT, POOL>(st, grid, nt, in, w, b, y, yf, n_img, h, wd, nstore, relu, uc, ypool);
    default: return -1;
  }
}

// The tiled forward convolution when the 18 x 18 tile of all input channels
// fits in kTileLdsMax, the layer has at least 256 tiles (smaller images keep
// the CUs busier with the pixel-linear kernel and its channel split), and
// the output channels split into tiles of at most 4 x 16; returns -1 when
// it does not apply.
template <bool UPCAT, bool POOL = false>
int launch_conv3_tile(hipStream_t st, const uint16_t* in, const uint16_t* w, const float* b, uint16_t* y, float* yf,
                      int n_img, int h, int wd, int cin, int nout, int nstore, int relu, UpCat uc,
                      uint16_t* ypool = nullptr) {
  const size_t lds = (size_t)kHalo * kHalo * cin * sizeof(uint16_t);
  const unsigned blocks = (unsigned)(n_img * ((h + kTile - 1) / kTile) * ((wd + kTile - 1) / kTile));
  if (lds > kTileLdsMax || blocks < 256) return -1;
  const int tiles = nout / 16;
  int groups = (tiles + 3) / 4;
  while (tiles % groups) ++groups;
  while (groups < tiles && (long long)xcd_grid(blocks) * groups < 512) {
    do ++groups;
    while (tiles % groups);
  }
  const dim3 grid(xcd_grid(blocks), groups);
  const int nt = tiles / groups;
  if (kConvFast) {
    const int e = launch_conv3_fast<UPCAT, POOL>(st, grid, nt, in, w, b, y, yf, n_img, h, wd, cin, nstore, relu, uc,
                                                 ypool);
    if (e >= 0) return e;
  }
  if (POOL) return -1;  // only the specialised kernel pools
  if (lds > 65536) {
    const bool ok = nt == 1   ? allow_tile_lds<1, UPCAT>()
                    : nt == 2 ? allow_tile_lds<2, UPCAT>()
                    : nt == 3 ? allow_tile_lds<3, UPCAT>()
                              : allow_tile_lds<4, UPCAT>();
    if (!ok) return -1;
  }
  switch (nt) {
    case 1: hipLaunchKernelGGL((k_conv3_tile<1, UPCAT>), grid, dim3(kThreads), lds, st, in, w, b, y, yf, n_img, h, wd, cin, nstore, relu, uc); break;
    case 2: hipLaunchKernelGGL((k_conv3_tile<2, UPCAT>), grid, dim3(kThreads), lds, st, in, w, b, y, yf, n_img, h, wd, cin, nstore, relu, uc); break;
    case 3: hipLaunchKernelGGL((k_conv3_tile<3, UPCAT>), grid, dim3(kThreads), lds, st, in, w, b, y, yf, n_img, h, wd, cin, nstore, relu, uc); break;
    default: hipLaunchKernelGGL((k_conv3_tile<4, UPCAT>), grid, dim3(kThreads), lds, st, in, w, b, y, yf, n_img, h, wd, cin, nstore, relu, uc); break;
  }
  return (int)hipGetLastError();
}

// The pixel-linear forward 3x3 convolution specialised on the input channel
// count, for the low-resolution layers (too few 16 x 16 tiles for the tiled
// kernel to fill the GPU).  There a wave's K loop (up to 63 steps of loads
// from L2) is a chain of latencies and the waves are few, so the 4 waves of
// a block share one output tile of 64 pixels x 16*NT channels and split its
// K loop in four ranges; each range is unrolled at compile time with the
// operands of step k + 2 loaded while step k multiplies.  The partial sums
// meet in LDS and are added in wave order (deterministic; a different
// summation order than k_conv, within the same float32 bound).  MFMA
// operands are swapped as in k_conv3_fast for 8-byte stores.
template <int C8, int NT, bool UPCAT>
__global__ __launch_bounds__(kThreads) void k_conv3_lin(const uint16_t* __restrict__ in,
                                                        const uint16_t* __restrict__ wgt,
                                                        const float* __restrict__ bias, uint16_t* __restrict__ y,
                                                        float* __restrict__ yf, int P, int h, int w, int nstore,
                                                        int relu, UpCat uc) {
  constexpr int kCin = 8 * C8, kSegs = 9 * C8, kSteps = (kSegs + 3) / 4, K = 9 * kCin;
  __shared__ f32x4 part[kWavesPerBlock][4][NT][64];
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4, q = threadIdx.x >> 6;
  const long long p0l = (long long)xcd_block() * 64;
  if (p0l >= P) return;  // the whole block, before the barrier
  const int p0 = (int)p0l;
  const int n0 = blockIdx.y * (16 * NT);
  int px[4], py[4], pimg[4];
  bool pv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int p = p0 + 16 * i + m;
    pv[i] = p < P;
    const int pc = pv[i] ? p : 0;
    const int row = pc / w;
    px[i] = pc - row * w;
    py[i] = row % h;
    pimg[i] = row / h;
  }
  const int c8low = UPCAT ? uc.c_low >> 3 : 0, cskip = UPCAT ? kCin - uc.c_low : 0;
  const uint16_t* wbase[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) wbase[j] = wgt + (size_t)(n0 + 16 * j + m) * K + g * 8;

  auto seg_ok = [&](int step) { return (4 * step + 3 < kSegs) || 4 * step + g < kSegs; };
  auto load_b = [&](int step, bf16x8* bb) {
    const bool sv = seg_ok(step);
#pragma unroll
    for (int j = 0; j < NT; ++j) bb[j] = sv ? load8(wbase[j] + 32 * step) : zero8();
  };
  auto load_a = [&](int step, bf16x8* a) {
    const int s = 4 * step + g;
    const bool sv = seg_ok(step);
    const int tap = s / C8, c8 = s - (s / C8) * C8;
    const int oy = tap / 3 - 1, ox = tap - (tap / 3) * 3 - 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool ok = sv && pv[i] && (unsigned)(py[i] + oy) < (unsigned)h && (unsigned)(px[i] + ox) < (unsigned)w;
      if (!ok)
        a[i] = zero8();
      else if (!UPCAT)
        a[i] = load8(in + (size_t)(p0 + 16 * i + m + oy * w + ox) * kCin + c8 * 8);
      else if (c8 < c8low)
        a[i] = load8(uc.low + ((size_t)(pimg[i] * (h >> 1) + ((py[i] + oy) >> 1)) * (w >> 1) + ((px[i] + ox) >> 1)) *
                                  uc.c_low + c8 * 8);
      else
        a[i] = load8(in + (size_t)(p0 + 16 * i + m + oy * w + ox) * cskip + (c8 - c8low) * 8);
    }
  };

  f32x4 acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // Wave q takes steps [first, last) of the K loop.
  auto run = [&](auto first_c, auto last_c) {
    constexpr int first = decltype(first_c)::value, last = decltype(last_c)::value;
    bf16x8 bq[3][NT], aq[3][4];
    if constexpr (first < last) {
      load_b(first, bq[0]);
      load_a(first, aq[0]);
    }
    if constexpr (first + 1 < last) {
      load_b(first + 1, bq[1]);
      load_a(first + 1, aq[1]);
    }
#pragma unroll
    for (int step = first; step < last; ++step) {
      const int k = step - first;
      if (step + 2 < last) {
        load_b(step + 2, bq[(k + 2) % 3]);
        load_a(step + 2, aq[(k + 2) % 3]);
      }
      __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of this step's MFMAs
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[k % 3][j], aq[k % 3][i], acc[i][j], 0, 0, 0);
    }
  };
  constexpr int q1 = kSteps / 4, q2 = kSteps / 2, q3 = (3 * kSteps) / 4;
  switch (q) {  // wave-uniform
    case 0: run(std::integral_constant<int, 0>{}, std::integral_constant<int, q1>{}); break;
    case 1: run(std::integral_constant<int, q1>{}, std::integral_constant<int, q2>{}); break;
    case 2: run(std::integral_constant<int, q2>{}, std::integral_constant<int, q3>{}); break;
    default: run(std::integral_constant<int, q3>{}, std::integral_constant<int, kSteps>{}); break;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) part[q][i][j][lane] = acc[i][j];
  __syncthreads();

  // Wave q finishes pixel row i = q of the tile: the four partial sums in
  // wave order, then bias, ReLU, bf16.  Lane (m, g) holds D[4g + r][m]:
  // channel n0 + 16j + 4g + r of pixel p0 + 16q + m.
  const int i = q;
  if (!pv[i]) return;
  const size_t pix = (size_t)(p0 + 16 * i + m);
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int co = n0 + 16 * j + 4 * g;
    if (co >= nstore) continue;  // nstore is a multiple of 4 (checked at launch)
    f32x4 s = part[0][i][j][lane];
#pragma unroll
    for (int r = 1; r < kWavesPerBlock; ++r) s += part[r][i][j][lane];
    const f32x4 bv = *reinterpret_cast<const f32x4*>(bias + co);
    float v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[r] = s[r] + bv[r];
      if (relu) v[r] = fmaxf(v[r], 0.f);
    }
    const size_t o = pix * nstore + co;
    if (y) *reinterpret_cast<uint2*>(y + o) = make_uint2(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]));
    if (yf) *reinterpret_cast<f32x4*>(yf + o) = f32x4{v[0], v[1], v[2], v[3]};
  }
}

template <int C8, bool UPCAT>
int go_conv3_lin(hipStream_t st, dim3 grid, int nt, const uint16_t* in, const uint16_t* w, const float* b, uint16_t* y,
                 float* yf, int P, int h, int wd, int nstore, int relu, UpCat uc) {
#define HPT_L(NT_)                                                                                                  \
  hipLaunchKernelGGL((k_conv3_lin<C8, NT_, UPCAT>), grid, dim3(kThreads), 0, st, in, w, b, y, yf, P, h, wd, nstore, \
                     relu, uc)
  switch (nt) {
    case 1: HPT_L(1); break;
    case 2: HPT_L(2); break;
    case 3: HPT_L(3); break;
    default: HPT_L(4); break;
  }
#undef HPT_L
  return (int)hipGetLastError();
}

// The specialised pixel-linear kernel for the channel counts it is built for
// (64, 96, 128, 160, 224): one block per 64-pixel tile and grid row of
// output channels, the channels split into as few rows (at most 4 tiles of
// 16 each) as still give 4 blocks per CU; -1 for other channel counts.
template <bool UPCAT>
int launch_conv3_lin(hipStream_t st, const uint16_t* in, const uint16_t* w, const float* b, uint16_t* y, float* yf,
                     int P, int h, int wd, int cin, int nout, int nstore, int relu, UpCat uc) {
  if (nstore % 4) return -1;
  if (cin != 64 && cin != 96 && cin != 128 && cin != 160 && cin != 224) return -1;
  const int tiles = nout / 16;
  const unsigned pixel_blocks = xcd_grid((unsigned)((P + 63) / 64));
  int groups = (tiles + 3) / 4;
  while (tiles % groups) ++groups;
  while (groups < tiles && (long long)pixel_blocks * groups < 1024) {
    do ++groups;
    while (tiles % groups);
  }
  const int nt = tiles / groups;
  const dim3 grid(pixel_blocks, groups);
  switch (cin) {
    case 64: return go_conv3_lin<8, UPCAT>(st, grid, nt, in, w, b, y, yf, P, h, wd, nstore, relu, uc);
    case 96: return go_conv3_lin<12, UPCAT>(st, grid, nt, in, w, b, y, yf, P, h, wd, nstore, relu, uc);
    case 128: return go_conv3_lin<16, UPCAT>(st, grid, nt, in, w, b, y, yf, P, h, wd, nstore, relu, uc);
    case 160: return go_conv3_lin<20, UPCAT>(st, grid, nt, in, w, b, y, yf, P, h, wd, nstore, relu, uc);
    default: return go_conv3_lin<28, UPCAT>(st, grid, nt, in, w, b, y, yf, P, h, wd, nstore, relu, uc);
  }
}

template <typename InT, int KS, int MT, int NT, bool FORWARD>
int go_conv(hipStream_t st, const InT* in, const uint16_t* w, const float* b, uint16_t* y, float* yf, int P, int h,
            int wd, int cin, int nstore, int relu, int groups, UpCat uc) {
  const int tiles = (P + 16 * MT - 1) / (16 * MT);
  const dim3 grid(xcd_grid((tiles + kWavesPerBlock - 1) / kWavesPerBlock), groups);
  if constexpr (KS == 3 && FORWARD && sizeof(InT) == 2) {
    if (uc.low) {
      hipLaunchKernelGGL((k_conv<InT, KS, MT, NT, FORWARD, true>), grid, dim3(kThreads), 0, st, in, w, b, y, yf, P, h,
                         wd, cin, nstore, relu, uc);
      return (int)hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_conv<InT, KS, MT, NT, FORWARD>), grid, dim3(kThreads), 0, st, in, w, b, y, yf, P, h, wd, cin,
                     nstore, relu, uc);
  return (int)hipGetLastError();
}

// N = nout channels (a multiple of 16) split into `groups` grid rows of NT <= 8 tiles each.
template <typename InT, int KS, bool FORWARD>
int launch_conv_ks(hipStream_t st, const InT* in, const uint16_t* w, const float* b, uint16_t* y, float* yf, int P,
                   int h, int wd, int cin, int nout, int nstore, int relu, UpCat uc) {
  if constexpr (KS == 3 && FORWARD && sizeof(InT) == 2) {
    if (kConvFast) {
      const int e = uc.low ? launch_conv3_lin<true>(st, in, w, b, y, yf, P, h, wd, cin, nout, nstore, relu, uc)
                           : launch_conv3_lin<false>(st, in, w, b, y, yf, P, h, wd, cin, nout, nstore, relu, uc);
      if (e >= 0) return e;
    }
  }
  const int tiles = nout / 16;
  int groups = (tiles + 7) / 8;
  while (tiles % groups) ++groups;
  // Low-resolution layers: split the output channels into more grid rows
  // until the grid gives every CU two workgroups (256 CUs).  Each output
  // still sums K in the same order, so the result does not change.
  auto blocks = [&](int gr) {
    const int mt = tiles / gr <= 4 ? 4 : 2;
    const long long waves = (P + 16LL * mt - 1) / (16LL * mt);
    return (long long)xcd_grid((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock)) * gr;
  };
  while (groups < tiles && blocks(groups) < 512) {
    do ++groups;
    while (tiles % groups);
  }
  switch (tiles / groups) {
    case 1: return go_conv<InT, KS, 4, 1, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
    case 2: return go_conv<InT, KS, 4, 2, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
    case 3: return go_conv<InT, KS, 4, 3, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
    case 4: return go_conv<InT, KS, 4, 4, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
    case 5: return go_conv<InT, KS, 2, 5, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
    case 6: return go_conv<InT, KS, 2, 6, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
    case 7: return go_conv<InT, KS, 2, 7, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
    default: return go_conv<InT, KS, 2, 8, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nstore, relu, groups, uc);
  }
}

// A 1x1 forward convolution to at most 4 channels (the network's output
// layer, 32 -> 3): one thread per pixel reads the pixel's channels once
// (16-byte loads) and sums each output over them in channel order; the
// weights are the same for every thread (scalar loads).  As an MFMA GEMM a
// wave would keep 13 of its 16 output columns idle and store 4 bytes at a
// time; this way the layer streams its input at close to the HBM rate.
template <int C8, int NO>
__global__ __launch_bounds__(kThreads) void k_conv1_narrow(const uint16_t* __restrict__ in,
                                                           const uint16_t* __restrict__ wgt,
                                                           const float* __restrict__ bias, uint16_t* __restrict__ y,
                                                           float* __restrict__ yf, int P, int relu) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= P) return;
  constexpr int kCin = 8 * C8;
  float acc[NO];
#pragma unroll
  for (int o = 0; o < NO; ++o) acc[o] = 0.f;
#pragma unroll
  for (int s = 0; s < C8; ++s) {
    const u32x4 v = *reinterpret_cast<const u32x4*>(in + (size_t)p * kCin + s * 8);
    const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float x = bf16_value((wv[e >> 1] >> (16 * (e & 1))) & 0xFFFFu);
#pragma unroll
      for (int o = 0; o < NO; ++o) acc[o] = fmaf(x, bf16_value(wgt[o * kCin + s * 8 + e]), acc[o]);
    }
  }
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    float v = acc[o] + bias[o];
    if (relu) v = fmaxf(v, 0.f);
    if (y) y[(size_t)p * NO + o] = (uint16_t)bf16_bits(v);
    if (yf) yf[(size_t)p * NO + o] = v;
  }
}

template <typename InT, bool FORWARD>
int launch_conv(hipStream_t st, const InT* in, const uint16_t* w, const float* b, uint16_t* y, float* yf, int P,
                int h, int wd, int cin, int nout, int nstore, int relu, int k, UpCat uc = UpCat{nullptr, 0}) {
  if constexpr (FORWARD && sizeof(InT) == 2) {
    if (k == 1 && kConvFast && cin == 32 && nstore == 3) {
      hipLaunchKernelGGL((k_conv1_narrow<4, 3>), dim3((unsigned)((P + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, in, w, b, y, yf,
                         P, relu);
      return (int)hipGetLastError();
    }
  }
  if (k == 1) return launch_conv_ks<InT, 1, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nout, nstore, relu, uc);
  return launch_conv_ks<InT, 3, FORWARD>(st, in, w, b, y, yf, P, h, wd, cin, nout, nstore, relu, uc);
}

// ------------------------------------------------------------------ weight gradient
// partial[chunk][co][t][ci] = sum over the chunk's pixels p of dy[p][co] * x[p + off(t)][ci].
// K = pixels: lane group g takes the 8 consecutive pixels base + 8g .. base + 8g + 7 of each
// 32-pixel step, for A (dy, rows = output channels) and B (x, columns = input channels) alike.
template <int KS, int MT, int NT>
__global__ __launch_bounds__(kThreads) void k_wgrad(const float* __restrict__ dy, const uint16_t* __restrict__ x,
                                                    float* __restrict__ partial, int P, int h, int w, int cin,
                                                    int cout_pad, int chunks, int co_groups, int ci_groups) {
  constexpr int taps = KS * KS;
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  int item = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (item >= chunks * taps * co_groups * ci_groups) return;  // uniform over the wave
  const int cig = item % ci_groups;
  item /= ci_groups;
  const int cog = item % co_groups;
  item /= co_groups;
  const int tap = item % taps, chunk = item / taps;
  const int co0 = cog * 16 * MT, ci0 = cig * 16 * NT;
  const int oy = tap / KS - KS / 2, ox = tap % KS - KS / 2;
  const int pbeg = chunk * kWgradChunk;
  const int pend = min(P, pbeg + kWgradChunk);

  int p = pbeg + 8 * g;  // this lane group's first pixel in the step
  int row = p / w;
  int px = p - row * w, py = row % h;

  f32x4 acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int base = pbeg; base < pend; base += 32) {
    uint32_t ab[MT][4], bb[NT][4];
    int qx = px, qy = py;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      float av[MT][2], bv[NT][2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int pi = p + e + u;
        const bool live = pi < pend;
        const bool ok = live && (unsigned)(qy + oy) < (unsigned)h && (unsigned)(qx + ox) < (unsigned)w;
#pragma unroll
        for (int i = 0; i < MT; ++i) av[i][u] = live ? dy[(size_t)pi * cout_pad + co0 + 16 * i + m] : 0.f;
#pragma unroll
        for (int j = 0; j < NT; ++j)
          bv[j][u] = ok ? bf16_value(x[(size_t)(pi + oy * w + ox) * cin + ci0 + 16 * j + m]) : 0.f;
        if (++qx == w) {
          qx = 0;
          if (++qy == h) qy = 0;
        }
      }
#pragma unroll
      for (int i = 0; i < MT; ++i) ab[i][e >> 1] = pack_bf16(av[i][0], av[i][1]);
#pragma unroll
      for (int j = 0; j < NT; ++j) bb[j][e >> 1] = pack_bf16(bv[j][0], bv[j][1]);
    }
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const bf16x8 a = __builtin_bit_cast(bf16x8, u32x4{ab[i][0], ab[i][1], ab[i][2], ab[i][3]});
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const bf16x8 b = __builtin_bit_cast(bf16x8, u32x4{bb[j][0], bb[j][1], bb[j][2], bb[j][3]});
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
      }
    }
    p += 32;
    px += 32;
    while (px >= w) {
      px -= w;
      if (++py == h) py = 0;
    }
  }

  // Lane (m, g) holds D[4g + r][m]: output channel co0 + 16i + 4g + r, input channel ci0 + 16j + m.
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = co0 + 16 * i + 4 * g + r, ci = ci0 + 16 * j + m;
        partial[((size_t)(chunk * cout_pad + co) * taps + tap) * cin + ci] = acc[i][j][r];
      }
}

// partial_b[chunk][co] = sum over the chunk's pixels of dy[p][co]: the block's
// threads form kThreads / cout_pad row groups (group g sums pixels g, g + groups,
// ...; consecutive threads read consecutive channels of a pixel), then thread co
// adds the groups in order (deterministic).
__global__ __launch_bounds__(kThreads) void k_bias_partial(const float* __restrict__ dy, float* __restrict__ partial_b,
                                                           int P, int cout_pad) {
  __shared__ float part[kThreads];
  const int groups = kThreads / cout_pad;
  const int g = threadIdx.x / cout_pad, co = threadIdx.x - g * cout_pad;
  const int pbeg = blockIdx.x * kWgradChunk, pend = min(P, pbeg + kWgradChunk);
  float s = 0.f;
  if (g < groups)
    for (int p = pbeg + g; p < pend; p += groups) s += dy[(size_t)p * cout_pad + co];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < cout_pad) {
    float t = 0.f;
    for (int k = 0; k < groups; ++k) t += part[k * cout_pad + threadIdx.x];
    partial_b[(size_t)blockIdx.x * cout_pad + threadIdx.x] = t;
  }
}

// dw[co][t][ci] (co < cout) = sum of the chunks in order; db likewise.
__global__ __launch_bounds__(kThreads) void k_wgrad_reduce(const float* __restrict__ partial,
                                                           const float* __restrict__ partial_b, float* __restrict__ dw,
                                                           float* __restrict__ db, int chunks, int cout, int cout_pad,
                                                           int row) {
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const size_t n = (size_t)cout * row;
  if (idx < n) {
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += partial[(size_t)c * cout_pad * row + idx];
    dw[idx] = s;
  } else if (idx < n + (size_t)cout) {
    const int co = (int)(idx - n);
    float s = 0.f;
    for (int c = 0; c < chunks; ++c) s += partial_b[(size_t)c * cout_pad + co];
    db[co] = s;
  }
}

template <int KS, int MT, int NT>
int go_wgrad(hipStream_t st, const float* dy, const uint16_t* x, float* partial, int P, int h, int wd, int cin,
             int cout_pad, int chunks) {
  const int co_groups = cout_pad / (16 * MT), ci_groups = cin / (16 * NT);
  const long items = (long)chunks * KS * KS * co_groups * ci_groups;
  hipLaunchKernelGGL((k_wgrad<KS, MT, NT>), dim3((unsigned)((items + kWavesPerBlock - 1) / kWavesPerBlock)),
                     dim3(kThreads), 0, st, dy, x, partial, P, h, wd, cin, cout_pad, chunks, co_groups, ci_groups);
  return (int)hipGetLastError();
}

template <int KS>
int launch_wgrad_ks(hipStream_t st, const float* dy, const uint16_t* x, float* partial, int P, int h, int wd, int cin,
                    int cout_pad, int chunks) {
  const bool co2 = (cout_pad / 16) % 2 == 0, ci2 = (cin / 16) % 2 == 0;
  if (co2 && ci2) return go_wgrad<KS, 2, 2>(st, dy, x, partial, P, h, wd, cin, cout_pad, chunks);
  if (co2) return go_wgrad<KS, 2, 1>(st, dy, x, partial, P, h, wd, cin, cout_pad, chunks);
  if (ci2) return go_wgrad<KS, 1, 2>(st, dy, x, partial, P, h, wd, cin, cout_pad, chunks);
  return go_wgrad<KS, 1, 1>(st, dy, x, partial, P, h, wd, cin, cout_pad, chunks);
}

// ------------------------------------------------------------------ elementwise steps
__global__ void k_prepare_weights(const float* __restrict__ master, uint16_t* __restrict__ w, uint16_t* __restrict__ wt,
                                  int cin, int cout, int cout_pad, int taps) {
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const size_t row = (size_t)taps * cin;
  if (idx >= (size_t)cout_pad * row) return;
  const int co = (int)(idx / row);
  const int rem = (int)(idx - (size_t)co * row);
  const int t = rem / cin, ci = rem - t * cin;
  const uint16_t b = (uint16_t)(co < cout ? bf16_bits(master[(size_t)co * row + rem]) : 0u);
  if (w) w[idx] = b;
  if (wt) wt[((size_t)ci * taps + (taps - 1 - t)) * cout_pad + co] = b;
}

__global__ void k_relu_backward(float* __restrict__ dy, const uint16_t* __restrict__ y, size_t count) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < count && (int16_t)y[i] <= 0) dy[i] = 0.f;
}

__global__ void k_pool_forward(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int n, int h, int w, int c) {
  const int ho = h / 2, wo = w / 2;
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= (size_t)n * ho * wo * c) return;
  const int ch = (int)(idx % c);
  const size_t pix = idx / c;
  const int xo = (int)(pix % wo);
  const size_t r = pix / wo;  // img * ho + yo
  const int yo = (int)(r % ho);
  const size_t img = r / ho;
  const size_t i00 = ((img * h + 2 * yo) * w + 2 * xo) * c + ch;
  const size_t dr = (size_t)w * c;
  const float s = (bf16_value(x[i00]) + bf16_value(x[i00 + c])) + (bf16_value(x[i00 + dr]) + bf16_value(x[i00 + dr + c]));
  y[idx] = (uint16_t)bf16_bits(0.25f * s);
}

__global__ void k_pool_backward(const float* __restrict__ dy, float* __restrict__ dx, int n, int h, int w, int c) {
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= (size_t)n * h * w * c) return;
  const int ch = (int)(idx % c);
  const size_t pix = idx / c;
  const int xx = (int)(pix % w);
  const size_t r = pix / w;
  const int yy = (int)(r % h);
  const size_t img = r / h;
  dx[idx] = 0.25f * dy[((img * (h / 2) + yy / 2) * (w / 2) + xx / 2) * c + ch];
}

__global__ void k_upsample_forward(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int n, int h, int w,
                                   int c) {
  const int ho = 2 * h, wo = 2 * w;
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= (size_t)n * ho * wo * c) return;
  const int ch = (int)(idx % c);
  const size_t pix = idx / c;
  const int xo = (int)(pix % wo);
  const size_t r = pix / wo;
  const int yo = (int)(r % ho);
  const size_t img = r / ho;
  y[idx] = x[((img * h + yo / 2) * w + xo / 2) * c + ch];
}

__global__ void k_upsample_backward(const float* __restrict__ dy, float* __restrict__ dx, int n, int h, int w, int c) {
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= (size_t)n * h * w * c) return;
  const int ch = (int)(idx % c);
  const size_t pix = idx / c;
  const int xx = (int)(pix % w);
  const size_t r = pix / w;
  const int yy = (int)(r % h);
  const size_t img = r / h;
  const size_t i00 = ((img * 2 * h + 2 * yy) * 2 * w + 2 * xx) * c + ch;
  const size_t dr = (size_t)2 * w * c;
  dx[idx] = (dy[i00] + dy[i00 + c]) + (dy[i00 + dr] + dy[i00 + dr + c]);
}

__global__ void k_concat_forward(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                                 uint16_t* __restrict__ y, size_t pixels, int ca, int cb) {
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const int c = ca + cb;
  if (idx >= pixels * c) return;
  const size_t p = idx / c;
  const int ch = (int)(idx - p * c);
  y[idx] = ch < ca ? a[p * ca + ch] : b[p * cb + ch - ca];
}

__global__ void k_concat_backward(const float* __restrict__ dy, float* __restrict__ da, float* __restrict__ db,
                                  size_t pixels, int ca, int cb) {
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const int c = ca + cb;
  if (idx >= pixels * c) return;
  const size_t p = idx / c;
  const int ch = (int)(idx - p * c);
  if (ch < ca) {
    if (da) da[p * ca + ch] = dy[idx];
  } else if (db) {
    db[p * cb + ch - ca] = dy[idx];
  }
}

__global__ void k_accumulate(float* __restrict__ dx, const float* __restrict__ dy, size_t count) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < count) dx[i] += dy[i];
}

// Source pixel of window pixel (xx, yy), mirrored as `mirror` asks.
__device__ __forceinline__ void window_source(int xx, int yy, int w, int h, int x0, int y0, int mirror, int& sx,
                                              int& sy) {
  sx = x0 + ((mirror & 1) ? w - 1 - xx : xx);
  sy = y0 + ((mirror & 2) ? h - 1 - yy : yy);
}

__global__ void k_make_input(const float4* __restrict__ color, const float4* __restrict__ albedo,
                             const float4* __restrict__ normal, int src_w, int src_h, int x0, int y0,
                             uint32_t* __restrict__ x, int h, int w, int mirror, float aux) {
  const int idx = blockIdx.x * kThreads + threadIdx.x;
  if (idx >= h * w) return;
  const int yy = idx / w, xx = idx - yy * w;
  int sx, sy;
  window_source(xx, yy, w, h, x0, y0, mirror, sx, sy);
  float v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (sx >= 0 && sy >= 0 && sx < src_w && sy < src_h) {
    const size_t s = (size_t)sy * src_w + sx;
    const float4 c = color[s], a = albedo[s], nn = normal[s];
    v[0] = log1pf(fmaxf(c.x, 0.f));
    v[1] = log1pf(fmaxf(c.y, 0.f));
    v[2] = log1pf(fmaxf(c.z, 0.f));
    v[3] = a.x;
    v[4] = a.y;
    v[5] = a.z;
    v[6] = nn.x * 0.5f + 0.5f;
    v[7] = nn.y * 0.5f + 0.5f;
    v[8] = nn.z * 0.5f + 0.5f;
  }
  u32x4* o = reinterpret_cast<u32x4*>(x + (size_t)idx * 8);
  o[0] = u32x4{pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]), pack_bf16(v[4], v[5]), pack_bf16(v[6], v[7])};
  const float a = (sx >= 0 && sy >= 0 && sx < src_w && sy < src_h) ? aux : 0.f;
  o[1] = u32x4{pack_bf16(v[8], a), 0u, 0u, 0u};
}

__global__ void k_make_target(const float4* __restrict__ color, int src_w, int src_h, int x0, int y0,
                              float* __restrict__ out, int h, int w, int mirror) {
  const int idx = blockIdx.x * kThreads + threadIdx.x;
  if (idx >= h * w) return;
  const int yy = idx / w, xx = idx - yy * w;
  int sx, sy;
  window_source(xx, yy, w, h, x0, y0, mirror, sx, sy);
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
  if (sx >= 0 && sy >= 0 && sx < src_w && sy < src_h) c = color[(size_t)sy * src_w + sx];
  out[(size_t)idx * 3 + 0] = log1pf(fmaxf(c.x, 0.f));
  out[(size_t)idx * 3 + 1] = log1pf(fmaxf(c.y, 0.f));
  out[(size_t)idx * 3 + 2] = log1pf(fmaxf(c.z, 0.f));
}

__global__ void k_add_residual(float* __restrict__ out, const uint16_t* __restrict__ x, size_t pixels, int xc) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= pixels * 3) return;
  const size_t p = i / 3;
  out[i] += bf16_value(x[p * xc + (i - p * 3)]);
}

__global__ __launch_bounds__(kThreads) void k_l1_loss(const float* __restrict__ out, const float* __restrict__ target,
                                                      float* __restrict__ dy, float* __restrict__ loss_sum,
                                                      size_t pixels, int cout_pad) {
  __shared__ float part[kWavesPerBlock];
  const size_t p = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const float scale = 1.0f / (3.0f * (float)pixels);
  float l = 0.f;
  if (p < pixels) {
    for (int c = 0; c < 3; ++c) {
      const float d = out[p * 3 + c] - target[p * 3 + c];
      l += fabsf(d);
      dy[p * cout_pad + c] = d > 0.f ? scale : (d < 0.f ? -scale : 0.f);
    }
    for (int c = 3; c < cout_pad; ++c) dy[p * cout_pad + c] = 0.f;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < kWavesPerBlock; ++i) s += part[i];
    atomicAdd(loss_sum, s * scale);
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       size_t count, float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_write_output(const float* __restrict__ out, float4* __restrict__ dst, int dst_w, int dst_h, int x0,
                               int y0, int h, int w) {
  const int idx = blockIdx.x * kThreads + threadIdx.x;
  if (idx >= h * w) return;
  const int yy = idx / w, xx = idx - yy * w;
  const int dx = x0 + xx, dyy = y0 + yy;
  if (dx < 0 || dyy < 0 || dx >= dst_w || dyy >= dst_h) return;
  const float* o = out + (size_t)idx * 3;
  dst[(size_t)dyy * dst_w + dx] = make_float4(fmaxf(expm1f(o[0]), 0.f), fmaxf(expm1f(o[1]), 0.f),
                                              fmaxf(expm1f(o[2]), 0.f), 1.f);
}

__global__ void k_fill(float* __restrict__ p, float v, size_t count) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < count) p[i] = v;
}

unsigned blocks_for(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int bad() { return (int)hipErrorInvalidValue; }

int last() { return (int)hipGetLastError(); }

bool conv_shape_ok(int n, int h, int wd, int cin, int cout_pad, int k) {
  if (n <= 0 || h <= 0 || wd <= 0 || cin <= 0 || cout_pad <= 0) return false;
  if (cin % 8 || cout_pad % 16 || (k != 1 && k != 3)) return false;
  return (long long)n * h * wd < (1ll << 31) && (long long)n * h * wd * (cin > cout_pad ? cin : cout_pad) < (1ll << 40);
}

int wgrad_chunks(long long P) { return (int)((P + kWgradChunk - 1) / kWgradChunk); }

}  // namespace

extern "C" {

int hpt_nn_conv_forward(void* stream, const uint16_t* x, const uint16_t* w, const float* b, uint16_t* y, float* y_f32,
                        int n, int h, int wd, int cin, int cout, int cout_pad, int k, int relu) {
  if (!conv_shape_ok(n, h, wd, cin, cout_pad, k) || cout <= 0 || cout > cout_pad || !x || !w || !b || (!y && !y_f32))
    return bad();
  if (k == 3 && kConvTile) {
    const int e = launch_conv3_tile<false>((hipStream_t)stream, x, w, b, y, y_f32, n, h, wd, cin, cout_pad, cout, relu,
                                           UpCat{nullptr, 0});
    if (e >= 0) return e;
  }
  return launch_conv<uint16_t, true>((hipStream_t)stream, x, w, b, y, y_f32, n * h * wd, h, wd, cin, cout_pad, cout,
                                     relu, k);
}

int hpt_nn_conv_forward_pool(void* stream, const uint16_t* x, const uint16_t* w, const float* b, uint16_t* y,
                             uint16_t* y_pool, int n, int h, int wd, int cin, int cout, int cout_pad, int relu) {
  if (!conv_shape_ok(n, h, wd, cin, cout_pad, 3) || cout <= 0 || cout > cout_pad || (h | wd) & 1 || !x || !w || !b ||
      !y || !y_pool)
    return bad();
  if (kConvTile) {
    const int e = launch_conv3_tile<false, true>((hipStream_t)stream, x, w, b, y, nullptr, n, h, wd, cin, cout_pad,
                                                 cout, relu, UpCat{nullptr, 0}, y_pool);
    if (e >= 0) return e;
  }
  const int e = hpt_nn_conv_forward(stream, x, w, b, y, nullptr, n, h, wd, cin, cout, cout_pad, 3, relu);
  if (e) return e;
  return hpt_nn_pool_forward(stream, y, y_pool, n, h, wd, cout);
}

int hpt_nn_conv_forward_upcat(void* stream, const uint16_t* low, const uint16_t* skip, const uint16_t* w,
                              const float* b, uint16_t* y, float* y_f32, int n, int h, int wd, int c_low, int c_skip,
                              int cout, int cout_pad, int relu) {
  const int cin = c_low + c_skip;
  if (!conv_shape_ok(n, h, wd, cin, cout_pad, 3) || h % 2 || wd % 2 || c_low <= 0 || c_low % 8 || c_skip <= 0 ||
      c_skip % 8 || cout <= 0 || cout > cout_pad || !low || !skip || !w || !b || (!y && !y_f32))
    return bad();
  const UpCat uc{low, c_low};
  if (kConvTile) {
    const int e =
        launch_conv3_tile<true>((hipStream_t)stream, skip, w, b, y, y_f32, n, h, wd, cin, cout_pad, cout, relu, uc);
    if (e >= 0) return e;
  }
  return launch_conv<uint16_t, true>((hipStream_t)stream, skip, w, b, y, y_f32, n * h * wd, h, wd, cin, cout_pad,
                                     cout, relu, 3, uc);
}

int hpt_nn_conv_backward_data(void* stream, const float* dy, const uint16_t* wt, float* dx, int n, int h, int wd,
                              int cin, int cout, int cout_pad, int k) {
  if (!conv_shape_ok(n, h, wd, cin, cout_pad, k) || cin % 16 || cout <= 0 || cout > cout_pad || !dy || !wt || !dx)
    return bad();
  return launch_conv<float, false>((hipStream_t)stream, dy, wt, nullptr, nullptr, dx, n * h * wd, h, wd, cout_pad,
                                   cin, cin, 0, k);
}

size_t hpt_nn_wgrad_scratch_floats(int n, int h, int wd, int cin, int cout_pad, int k) {
  const long long P = (long long)n * h * wd;
  const size_t chunks = (size_t)wgrad_chunks(P);
  return chunks * (size_t)cout_pad * (size_t)(k * k) * (size_t)cin + chunks * (size_t)cout_pad;
}

int hpt_nn_conv_backward_weights(void* stream, const float* dy, const uint16_t* x, float* dw, float* db,
                                 float* partial, int n, int h, int wd, int cin, int cout, int cout_pad, int k) {
  if (!conv_shape_ok(n, h, wd, cin, cout_pad, k) || cin % 16 || cout_pad > kThreads || cout <= 0 ||
      cout > cout_pad || !dy || !x || !dw || !db || !partial)
    return bad();
  hipStream_t st = (hipStream_t)stream;
  const int P = n * h * wd, chunks = wgrad_chunks(P);
  const int e = k == 1 ? launch_wgrad_ks<1>(st, dy, x, partial, P, h, wd, cin, cout_pad, chunks)
                       : launch_wgrad_ks<3>(st, dy, x, partial, P, h, wd, cin, cout_pad, chunks);
  if (e) return e;
  float* partial_b = partial + (size_t)chunks * cout_pad * k * k * cin;
  hipLaunchKernelGGL(k_bias_partial, dim3(chunks), dim3(kThreads), 0, st, dy, partial_b, P, cout_pad);
  if (const int e2 = last()) return e2;
  const int row = k * k * cin;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for((size_t)cout * row + cout)), dim3(kThreads), 0, st, partial,
                     partial_b, dw, db, chunks, cout, cout_pad, row);
  return last();
}

int hpt_nn_prepare_weights(void* stream, const float* master, uint16_t* w, uint16_t* wt, int cin, int cout,
                           int cout_pad, int k) {
  if (!master || (!w && !wt) || cin <= 0 || cout <= 0 || cout > cout_pad || (k != 1 && k != 3)) return bad();
  hipLaunchKernelGGL(k_prepare_weights, dim3(blocks_for((size_t)cout_pad * k * k * cin)), dim3(kThreads), 0,
                     (hipStream_t)stream, master, w, wt, cin, cout, cout_pad, k * k);
  return last();
}

int hpt_nn_relu_backward(void* stream, float* dy, const uint16_t* y, size_t count) {
  if (!count) return 0;
  if (!dy || !y) return bad();
  hipLaunchKernelGGL(k_relu_backward, dim3(blocks_for(count)), dim3(kThreads), 0, (hipStream_t)stream, dy, y, count);
  return last();
}

int hpt_nn_pool_forward(void* stream, const uint16_t* x, uint16_t* y, int n, int h, int wd, int c) {
  if (!x || !y || n <= 0 || h <= 0 || wd <= 0 || c <= 0 || (h | wd) & 1) return bad();
  hipLaunchKernelGGL(k_pool_forward, dim3(blocks_for((size_t)n * (h / 2) * (wd / 2) * c)), dim3(kThreads), 0,
                     (hipStream_t)stream, x, y, n, h, wd, c);
  return last();
}

int hpt_nn_pool_backward(void* stream, const float* dy, float* dx, int n, int h, int wd, int c) {
  if (!dy || !dx || n <= 0 || h <= 0 || wd <= 0 || c <= 0 || (h | wd) & 1) return bad();
  hipLaunchKernelGGL(k_pool_backward, dim3(blocks_for((size_t)n * h * wd * c)), dim3(kThreads), 0,
                     (hipStream_t)stream, dy, dx, n, h, wd, c);
  return last();
}

int hpt_nn_upsample_forward(void* stream, const uint16_t* x, uint16_t* y, int n, int h, int wd, int c) {
  if (!x || !y || n <= 0 || h <= 0 || wd <= 0 || c <= 0) return bad();
  hipLaunchKernelGGL(k_upsample_forward, dim3(blocks_for((size_t)n * 4 * h * wd * c)), dim3(kThreads), 0,
                     (hipStream_t)stream, x, y, n, h, wd, c);
  return last();
}

int hpt_nn_upsample_backward(void* stream, const float* dy, float* dx, int n, int h, int wd, int c) {
  if (!dy || !dx || n <= 0 || h <= 0 || wd <= 0 || c <= 0) return bad();
  hipLaunchKernelGGL(k_upsample_backward, dim3(blocks_for((size_t)n * h * wd * c)), dim3(kThreads), 0,
                     (hipStream_t)stream, dy, dx, n, h, wd, c);
  return last();
}

int hpt_nn_concat_forward(void* stream, const uint16_t* a, const uint16_t* b, uint16_t* y, size_t pixels, int ca,
                          int cb) {
  if (!pixels) return 0;
  if (!a || !b || !y || ca <= 0 || cb <= 0) return bad();
  hipLaunchKernelGGL(k_concat_forward, dim3(blocks_for(pixels * (ca + cb))), dim3(kThreads), 0, (hipStream_t)stream,
                     a, b, y, pixels, ca, cb);
  return last();
}

int hpt_nn_concat_backward(void* stream, const float* dy, float* da, float* db, size_t pixels, int ca, int cb) {
  if (!pixels) return 0;
  if (!dy || ca <= 0 || cb <= 0) return bad();
  hipLaunchKernelGGL(k_concat_backward, dim3(blocks_for(pixels * (ca + cb))), dim3(kThreads), 0, (hipStream_t)stream,
                     dy, da, db, pixels, ca, cb);
  return last();
}

int hpt_nn_accumulate(void* stream, float* dx, const float* dy, size_t count) {
  if (!count) return 0;
  if (!dx || !dy) return bad();
  hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(count)), dim3(kThreads), 0, (hipStream_t)stream, dx, dy, count);
  return last();
}

int hpt_nn_make_input(void* stream, const float* color, const float* albedo, const float* normal, int src_w, int src_h,
                      int x0, int y0, uint16_t* x, int h, int wd, int mirror, float aux) {
  if (!color || !albedo || !normal || !x || src_w <= 0 || src_h <= 0 || h <= 0 || wd <= 0 ||
      (long long)h * wd >= (1ll << 31))
    return bad();
  hipLaunchKernelGGL(k_make_input, dim3(blocks_for((size_t)h * wd)), dim3(kThreads), 0, (hipStream_t)stream,
                     reinterpret_cast<const float4*>(color), reinterpret_cast<const float4*>(albedo),
                     reinterpret_cast<const float4*>(normal), src_w, src_h, x0, y0, reinterpret_cast<uint32_t*>(x), h,
                     wd, mirror, aux);
  return last();
}

int hpt_nn_make_target(void* stream, const float* color, int src_w, int src_h, int x0, int y0, float* out, int h,
                       int wd, int mirror) {
  if (!color || !out || src_w <= 0 || src_h <= 0 || h <= 0 || wd <= 0 || (long long)h * wd >= (1ll << 31))
    return bad();
  hipLaunchKernelGGL(k_make_target, dim3(blocks_for((size_t)h * wd)), dim3(kThreads), 0, (hipStream_t)stream,
                     reinterpret_cast<const float4*>(color), src_w, src_h, x0, y0, out, h, wd, mirror);
  return last();
}

int hpt_nn_add_residual(void* stream, float* out, const uint16_t* x, size_t pixels, int x_channels) {
  if (!pixels) return 0;
  if (!out || !x || x_channels < 3) return bad();
  hipLaunchKernelGGL(k_add_residual, dim3(blocks_for(pixels * 3)), dim3(kThreads), 0, (hipStream_t)stream, out, x,
                     pixels, x_channels);
  return last();
}

int hpt_nn_l1_loss(void* stream, const float* out, const float* target_log, float* dy, float* loss_sum, size_t pixels,
                   int cout_pad) {
  if (!pixels) return 0;
  if (!out || !target_log || !dy || !loss_sum || cout_pad < 3) return bad();
  hipLaunchKernelGGL(k_l1_loss, dim3(blocks_for(pixels)), dim3(kThreads), 0, (hipStream_t)stream, out, target_log, dy,
                     loss_sum, pixels, cout_pad);
  return last();
}

int hpt_nn_adam(void* stream, float* p, const float* g, float* m, float* v, size_t count, float lr, float beta1,
                float beta2, float eps, int t) {
  if (!count) return 0;
  if (!p || !g || !m || !v || t < 1) return bad();
  const float c1 = 1.f - (float)std::pow((double)beta1, (double)t);
  const float c2 = 1.f - (float)std::pow((double)beta2, (double)t);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(count)), dim3(kThreads), 0, (hipStream_t)stream, p, g, m, v, count, lr,
                     beta1, beta2, eps, c1, c2);
  return last();
}

int hpt_nn_write_output(void* stream, const float* out, float* dst, int dst_w, int dst_h, int x0, int y0, int h,
                        int wd) {
  if (!out || !dst || dst_w <= 0 || dst_h <= 0 || h <= 0 || wd <= 0 || (long long)h * wd >= (1ll << 31)) return bad();
  hipLaunchKernelGGL(k_write_output, dim3(blocks_for((size_t)h * wd)), dim3(kThreads), 0, (hipStream_t)stream, out,
                     reinterpret_cast<float4*>(dst), dst_w, dst_h, x0, y0, h, wd);
  return last();
}

// Zeroing runs on a non-blocking stream of the library per device, not on
// the legacy null stream: work there would depend on every blocking stream,
// and HIP refuses it while another thread captures a graph.
static hipError_t zeroing_stream(hipStream_t* out) {
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!streams[dev]) {
    e = hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
    if (e != hipSuccess) return e;
  }
  *out = streams[dev];
  return hipSuccess;
}

int hpt_nn_alloc(size_t bytes, void** ptr) {
  if (!ptr) return bad();
  *ptr = nullptr;
  if (!bytes) return 0;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return (int)e;
  hipStream_t s = nullptr;
  e = zeroing_stream(&s);
  if (e == hipSuccess) e = hipMemsetAsync(*ptr, 0, bytes, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(*ptr);
    *ptr = nullptr;
  }
  return (int)e;
}

int hpt_nn_stream_create(void** stream) {
  if (!stream) return bad();
  return (int)hipStreamCreateWithFlags(reinterpret_cast<hipStream_t*>(stream), hipStreamNonBlocking);
}

int hpt_nn_free(void* ptr) { return ptr ? (int)hipFree(ptr) : 0; }

int hpt_nn_copy(void* stream, void* dst, const void* src, size_t bytes) {
  if (!bytes) return 0;
  if (!dst || !src) return bad();
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream);
}

int hpt_nn_fill(void* stream, float* p, float value, size_t count) {
  if (!count) return 0;
  if (!p) return bad();
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(count)), dim3(kThreads), 0, (hipStream_t)stream, p, value, count);
  return last();
}

int hpt_nn_stream_synchronize(void* stream) { return (int)hipStreamSynchronize((hipStream_t)stream); }

int hpt_nn_event_create(void** event) {
  if (!event) return bad();
  return (int)hipEventCreate(reinterpret_cast<hipEvent_t*>(event));
}

int hpt_nn_event_destroy(void* event) { return event ? (int)hipEventDestroy((hipEvent_t)event) : 0; }

int hpt_nn_event_record(void* event, void* stream) {
  if (!event) return bad();
  return (int)hipEventRecord((hipEvent_t)event, (hipStream_t)stream);
}

int hpt_nn_event_elapsed(void* start, void* stop, float* ms) {
  if (!start || !stop || !ms) return bad();
  const hipError_t e = hipEventSynchronize((hipEvent_t)stop);
  if (e != hipSuccess) return (int)e;
  return (int)hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
}

}  // extern "C"
