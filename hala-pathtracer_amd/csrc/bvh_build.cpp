// bvh_build.cpp — host build of the 8-wide compressed BVH (Node8/LeafTri).
//
// Binary SAH tree over triangle centroids (32 bins per axis), leaves of at
// most 3 triangles (the split stable, so a leaf lists its triangles in id
// order: the GPU build in bvh_gpu.hip reproduces this tree byte for byte),
// collapsed greedily to 8-wide nodes and quantised to 8 bits
// per plane (the compressed-wide-BVH node of Ylitie, Karras and Laine, "Efficient
// Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017,
// with this project's own field layout; see common.h).
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>

namespace hpt {
namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<float>::infinity();
      hi[a] = -std::numeric_limits<float>::infinity();
    }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  double area() const {
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = std::max(0.0, (double)hi[a] - (double)lo[a]);
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }
};

struct BNode {
  Box box;
  uint32_t left = 0, right = 0;  // inner: children
  uint32_t first = 0, count = 0;  // leaf: range in order[]
  bool leaf = false;
};

constexpr int kBins = 32;
constexpr uint32_t kMaxLeaf = 3;

struct Builder {
  const float* pos;
  uint32_t n;
  std::vector<Box> tbox;
  std::vector<float> cen;  // 3 per triangle
  std::vector<uint32_t> order;
  std::vector<BNode> nodes;

  void build() {
    tbox.resize(n);
    cen.resize((size_t)n * 3);
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      Box b;
      b.reset();
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
          float x = pos[(size_t)i * 9 + k * 3 + a];
          b.lo[a] = std::min(b.lo[a], x);
          b.hi[a] = std::max(b.hi[a], x);
        }
      tbox[i] = b;
      for (int a = 0; a < 3; ++a) cen[(size_t)i * 3 + a] = 0.5f * (b.lo[a] + b.hi[a]);
      order[i] = i;
    }
    nodes.reserve(2 * (size_t)n / 2 + 1);
    nodes.push_back(BNode{});
    struct Item {
      uint32_t node, first, count;
    };
    std::vector<Item> stack;
    stack.push_back({0, 0, n});
    while (!stack.empty()) {
      Item it = stack.back();
      stack.pop_back();
      BNode nd;
      nd.box.reset();
      Box cb;
      cb.reset();
      for (uint32_t k = it.first; k < it.first + it.count; ++k) {
        uint32_t t = order[k];
        nd.box.grow(tbox[t]);
        for (int a = 0; a < 3; ++a) {
          cb.lo[a] = std::min(cb.lo[a], cen[(size_t)t * 3 + a]);
          cb.hi[a] = std::max(cb.hi[a], cen[(size_t)t * 3 + a]);
        }
      }
      uint32_t mid = 0;
      bool make_leaf = it.count <= 1;
      if (!make_leaf) {
        // Binned SAH.
        double best_cost = std::numeric_limits<double>::infinity();
        int best_axis = -1, best_split = -1;
        const double parent_area = std::max(nd.box.area(), 1e-30);
        for (int a = 0; a < 3; ++a) {
          const float ext = cb.hi[a] - cb.lo[a];
          if (!(ext > 0.0f)) continue;
          uint32_t cnt[kBins] = {};
          Box bb[kBins];
          for (int b = 0; b < kBins; ++b) bb[b].reset();
          const float scale = (float)kBins / ext;
          for (uint32_t k = it.first; k < it.first + it.count; ++k) {
            uint32_t t = order[k];
            int b = (int)((cen[(size_t)t * 3 + a] - cb.lo[a]) * scale);
            b = std::min(std::max(b, 0), kBins - 1);
            cnt[b]++;
            bb[b].grow(tbox[t]);
          }
          double rarea[kBins];
          uint32_t rcnt[kBins];
          Box acc;
          acc.reset();
          uint32_t c = 0;
          for (int b = kBins - 1; b > 0; --b) {
            acc.grow(bb[b]);
            c += cnt[b];
            rarea[b] = acc.area();
            rcnt[b] = c;
          }
          acc.reset();
          c = 0;
          for (int b = 0; b < kBins - 1; ++b) {
            acc.grow(bb[b]);
            c += cnt[b];
            if (c == 0 || rcnt[b + 1] == 0) continue;
            double cost = 1.0 + (acc.area() * c + rarea[b + 1] * rcnt[b + 1]) / parent_area;
            if (cost < best_cost) {
              best_cost = cost;
              best_axis = a;
              best_split = b + 1;
            }
          }
        }
        if (it.count <= kMaxLeaf && (double)it.count <= best_cost) make_leaf = true;
        if (!make_leaf) {
          if (best_axis >= 0) {
            const float ext = cb.hi[best_axis] - cb.lo[best_axis];
            const float scale = (float)kBins / ext;
            auto beg = order.begin() + it.first, end = beg + it.count;
            auto m = std::stable_partition(beg, end, [&](uint32_t t) {
              int b = (int)((cen[(size_t)t * 3 + best_axis] - cb.lo[best_axis]) * scale);
              b = std::min(std::max(b, 0), kBins - 1);
              return b < best_split;
            });
            mid = (uint32_t)(m - order.begin());
          }
          if (best_axis < 0 || mid == it.first || mid == it.first + it.count) {
            // All centroids coincide: split the range in half.
            mid = it.first + it.count / 2;
          }
        }
      }
      if (make_leaf) {
        nd.leaf = true;
        nd.first = it.first;
        nd.count = it.count;
        nodes[it.node] = nd;
        continue;
      }
      nd.left = (uint32_t)nodes.size();
      nd.right = nd.left + 1;
      nodes[it.node] = nd;
      nodes.push_back(BNode{});
      nodes.push_back(BNode{});
      stack.push_back({nd.left, it.first, mid - it.first});
      stack.push_back({nd.right, mid, it.first + it.count - mid});
    }
  }
};

// Smallest power-of-two scale with 255*scale >= ext, as a biased exponent.
int exponent_for(double ext) {
  if (!(ext > 0.0)) return 1;
  int x;
  double m = std::frexp(ext / 255.0, &x);  // ext/255 = m * 2^x, m in [0.5, 1)
  int e = (m == 0.5) ? x - 1 : x;
  return std::min(std::max(e + 127, 1), 254);
}

float scale_of(int eb) {
  uint32_t bits = (uint32_t)eb << 23;
  float s;
  std::memcpy(&s, &bits, 4);
  return s;
}

}  // namespace

bool build_bvh8(const float* pos, uint32_t tri_count, BvhBuildResult& out, std::string& err) {
  out.nodes.clear();
  out.tris.clear();
  out.depth = 0;
  if (tri_count == 0) return true;
  for (size_t i = 0; i < (size_t)tri_count * 9; ++i)
    if (!std::isfinite(pos[i])) {
      err = "triangle vertex is not finite";
      return false;
    }
  Builder b;
  b.pos = pos;
  b.n = tri_count;
  b.build();

  struct Work {
    uint32_t bnode, out, depth;
  };
  std::deque<Work> queue;
  out.nodes.push_back(Node8{});
  out.tris.reserve(tri_count);
  queue.push_back({0, 0, 1});
  while (!queue.empty()) {
    Work w = queue.front();
    queue.pop_front();
    out.depth = std::max(out.depth, w.depth);
    // Gather up to 8 children by opening the largest inner child.
    std::vector<uint32_t> ch;
    const BNode& root = b.nodes[w.bnode];
    if (root.leaf) {
      ch.push_back(w.bnode);
    } else {
      ch.push_back(root.left);
      ch.push_back(root.right);
      while (ch.size() < 8) {
        int pick = -1;
        double best = -1.0;
        for (size_t i = 0; i < ch.size(); ++i) {
          const BNode& c = b.nodes[ch[i]];
          if (c.leaf) continue;
          double a = c.box.area();
          if (a > best) {
            best = a;
            pick = (int)i;
          }
        }
        if (pick < 0) break;
        uint32_t open = ch[pick];
        ch[pick] = b.nodes[open].left;
        ch.push_back(b.nodes[open].right);
      }
    }
    // Node bounds and the octant of each child around the node centre.
    Box nb;
    nb.reset();
    for (uint32_t c : ch) nb.grow(b.nodes[c].box);
    double cx = 0.5 * ((double)nb.lo[0] + nb.hi[0]), cy = 0.5 * ((double)nb.lo[1] + nb.hi[1]),
           cz = 0.5 * ((double)nb.lo[2] + nb.hi[2]);
    int slot_of[8];
    bool used_slot[8] = {};
    bool done[8] = {};
    for (size_t r = 0; r < ch.size(); ++r) {
      double best = -std::numeric_limits<double>::infinity();
      int bi = -1, bs = -1;
      for (size_t i = 0; i < ch.size(); ++i) {
        if (done[i]) continue;
        const Box& cb = b.nodes[ch[i]].box;
        double dx = 0.5 * ((double)cb.lo[0] + cb.hi[0]) - cx;
        double dy = 0.5 * ((double)cb.lo[1] + cb.hi[1]) - cy;
        double dz = 0.5 * ((double)cb.lo[2] + cb.hi[2]) - cz;
        for (int s = 0; s < 8; ++s) {
          if (used_slot[s]) continue;
          double cost = ((s & 1) ? dx : -dx) + ((s & 2) ? dy : -dy) + ((s & 4) ? dz : -dz);
          if (cost > best) {
            best = cost;
            bi = (int)i;
            bs = s;
          }
        }
      }
      done[bi] = true;
      used_slot[bs] = true;
      slot_of[bi] = bs;
    }
    int child_at[8];
    for (int s = 0; s < 8; ++s) child_at[s] = -1;
    for (size_t i = 0; i < ch.size(); ++i) child_at[slot_of[i]] = (int)i;

    Node8 nd;
    std::memset(&nd, 0, sizeof nd);
    for (int a = 0; a < 3; ++a) nd.p[a] = nb.lo[a];
    // Inner children first: contiguous in slot order.
    uint32_t inner = 0;
    for (int s = 0; s < 8; ++s)
      if (child_at[s] >= 0 && !b.nodes[ch[child_at[s]]].leaf) inner++;
    nd.child_base = (uint32_t)out.nodes.size();
    nd.tri_base = (uint32_t)out.tris.size();
    out.nodes.resize(out.nodes.size() + inner);
    uint32_t rank = 0, toff = 0;
    for (int s = 0; s < 8; ++s) {
      if (child_at[s] < 0) continue;
      const BNode& c = b.nodes[ch[child_at[s]]];
      if (!c.leaf) {
        nd.imask |= (uint8_t)(1u << s);
        nd.meta[s] = 0x1F;
        queue.push_back({ch[child_at[s]], nd.child_base + rank, w.depth + 1});
        rank++;
      } else {
        const uint32_t cnt = c.count;
        nd.meta[s] = (uint8_t)((((1u << cnt) - 1u) << 5) | toff);
        for (uint32_t k = 0; k < cnt; ++k) {
          uint32_t t = b.order[c.first + k];
          LeafTri lt;
          std::memcpy(lt.v0, pos + (size_t)t * 9 + 0, 12);
          std::memcpy(lt.v1, pos + (size_t)t * 9 + 3, 12);
          std::memcpy(lt.v2, pos + (size_t)t * 9 + 6, 12);
          lt.id = t;
          lt.pad1 = lt.pad2 = 0;
          out.tris.push_back(lt);
        }
        toff += cnt;
      }
    }
    // Quantise: lo = fma(qlo, scale, p) <= child lo and fma(qhi, scale, p) >= child hi.
    for (int a = 0; a < 3; ++a) {
      int eb = exponent_for((double)nb.hi[a] - (double)nb.lo[a]);
      for (;;) {
        const float sc = scale_of(eb);
        bool ok = true;
        for (int s = 0; s < 8 && ok; ++s) {
          if (child_at[s] < 0) {
            nd.qlo[a][s] = 0;
            nd.qhi[a][s] = 0;
            continue;
          }
          const Box& cb = b.nodes[ch[child_at[s]]].box;
          double ql = std::floor(((double)cb.lo[a] - (double)nd.p[a]) / (double)sc);
          double qh = std::ceil(((double)cb.hi[a] - (double)nd.p[a]) / (double)sc);
          int il = (int)std::min(std::max(ql, 0.0), 255.0);
          int ih = (int)std::min(std::max(qh, 0.0), 255.0);
          while (il > 0 && std::fma((float)il, sc, nd.p[a]) > cb.lo[a]) il--;
          while (ih < 255 && std::fma((float)ih, sc, nd.p[a]) < cb.hi[a]) ih++;
          if (std::fma((float)il, sc, nd.p[a]) > cb.lo[a] || std::fma((float)ih, sc, nd.p[a]) < cb.hi[a]) {
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
          err = "scene extent too large to quantise";
          return false;
        }
        eb++;
      }
    }
    out.nodes[w.out] = nd;
  }
  return true;
}

}  // namespace hpt
