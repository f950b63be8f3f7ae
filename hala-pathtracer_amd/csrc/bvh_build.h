// bvh_build.h — host build of the 8-wide compressed BVH the kernels walk.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "common.h"

namespace hpt {

struct BvhBuildResult {
  std::vector<Node8> nodes;
  std::vector<LeafTri> tris;
  uint32_t depth = 0;  // levels of the 8-wide tree
};

// Builds from world-space triangles: `pos` holds 9 floats per triangle
// (v0, v1, v2), triangle i has global id i.  Binned SAH (32 bins) over
// centroids down to leaves of at most 3 triangles, then a greedy collapse of
// the binary tree to 8-wide nodes (open the largest-area inner child until 8
// slots are used), child-to-slot assignment by octant, 8-bit quantisation.
// Returns false with a message on invalid input.
bool build_bvh8(const float* pos, uint32_t tri_count, BvhBuildResult& out, std::string& err);

}  // namespace hpt
