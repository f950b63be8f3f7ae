// bvh_gpu.h — on-GPU build of the 8-wide compressed BVH the kernels walk.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.h"

namespace hpt {

struct GpuBvh {
  Node8* nodes = nullptr;    // device, node_count entries (caller frees with hipFree)
  LeafTri* tris = nullptr;   // device, tri_count entries (caller frees with hipFree)
  uint32_t node_count = 0;
  uint32_t depth = 0;        // levels of the 8-wide tree
};

// Builds from world-space triangles on the device (`d_pos`: 9 floats per
// triangle, triangle i has global id i) on `stream`, with the same binned SAH
// (32 bins, leaves of at most 3 triangles, stable splits), greedy 8-wide
// collapse, octant slot order and quantisation as build_bvh8 -- the output
// is the same bytes.  Returns false with a message on invalid input or when
// a split of more than 1024 triangles finds all their centroids equal (the
// caller then builds on the host).
bool build_bvh8_gpu(const float* d_pos, uint32_t tri_count, hipStream_t stream, GpuBvh& out, std::string& err);

}  // namespace hpt
