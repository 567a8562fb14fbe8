// rt_search_bvh.cpp -- the search BVH of the per-ray path.
//
// Incoherent ray groups (random shadow rays, diffuse bounces) have direction boxes that cover
// most of the sphere, so the reference's cull chain keeps nearly every (leaf, triangle) pair
// of such a group and enumerating the candidate set costs a leaf test per pair.  The per-ray
// path finds the same answer from the other side: every candidate that can change a ray
// passes the strict leaf test, so each ray walks a BVH over the triangles, runs the strict
// leaf test on the triangles of the leaves it reaches, and evaluates the reference's exact
// cull chain only for triangles that pass it.  The BVH only has to be conservative: a box
// test may pass for a box the ray misses, never the other way round.
//
// The tree is built over the triangles in ascending 30-bit Morton order of their centroids
// (the scene's sorted traversal copy), splitting each range at its highest differing code
// bit, or at the middle when that split would make the tree deeper than kSearchDepth
// interior levels or the codes are equal.  Leaves hold at most kSearchLeaf consecutive
// triangles.  Nodes are stored depth first, each holding the boxes and references of its two
// children.  Every box is the FP32 min/max of its triangles' vertices grown outward by
// pad = 2^-11 * max(1, largest vertex coordinate magnitude), which covers the rounding of
// the strict leaf test's hit point and of the slab test (the argument and its limit are in
// the design notes of the per-ray path).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rt_internal.h"

namespace rti {
namespace {

struct SBox {
  float lo[3], hi[3];
};

// Interior levels a balanced (median split) tree needs over n triangles.
int levels_needed(uint32_t n) {
  int l = 0;
  while (n > (uint32_t)kSearchLeaf) {
    n = (n + 1) / 2;
    ++l;
  }
  return l;
}

class SearchBuilder {
 public:
  SearchBuilder(const float* v9, const uint32_t* code, std::vector<float4>& out, float pad)
      : v9_(v9), code_(code), out_(out), pad_(pad) {}

  // Reference of the subtree over [a, b) whose parent sits at interior depth `depth`
  // (the root node is depth 1); its box (unpadded) goes to `box`.
  uint32_t build(uint32_t a, uint32_t b, int depth, SBox& box) {
    const uint32_t n = b - a;
    if (n <= (uint32_t)kSearchLeaf) {
      tri_box(a, b, box);
      return 0x80000000u | ((n - 1) << 27) | a;
    }
    const uint32_t node = reserve();
    const uint32_t m = split(a, b, depth);
    SBox b0, b1;
    const uint32_t r0 = build(a, m, depth + 1, b0);
    const uint32_t r1 = build(m, b, depth + 1, b1);
    write(node, b0, r0, b1, r1);
    for (int k = 0; k < 3; ++k) {
      box.lo[k] = std::min(b0.lo[k], b1.lo[k]);
      box.hi[k] = std::max(b0.hi[k], b1.hi[k]);
    }
    return node;
  }

  // Root node over every triangle; a scene of at most kSearchLeaf triangles gets a node
  // with two leaves (one triangle repeated when there is only one).
  void build_root(uint32_t ntri) {
    if (ntri <= (uint32_t)kSearchLeaf) {
      const uint32_t node = reserve();
      const uint32_t m = ntri == 1 ? 1 : ntri / 2;
      SBox b0, b1;
      tri_box(0, m, b0);
      const uint32_t r0 = 0x80000000u | ((m - 1) << 27);
      uint32_t r1 = r0;
      b1 = b0;
      if (ntri > 1) {
        tri_box(m, ntri, b1);
        r1 = 0x80000000u | ((ntri - m - 1) << 27) | m;
      }
      write(node, b0, r0, b1, r1);
      return;
    }
    SBox box;
    build(0, ntri, 1, box);
  }

 private:
  uint32_t reserve() {
    const uint32_t node = (uint32_t)(out_.size() / 4);
    out_.resize(out_.size() + 4);
    return node;
  }

  void tri_box(uint32_t a, uint32_t b, SBox& box) const {
    for (int k = 0; k < 3; ++k) {
      box.lo[k] = INFINITY;
      box.hi[k] = -INFINITY;
    }
    for (uint32_t i = a; i < b; ++i)
      for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) {
          const float x = v9_[9 * (size_t)i + 3 * v + k];
          box.lo[k] = std::min(box.lo[k], x);
          box.hi[k] = std::max(box.hi[k], x);
        }
  }

  uint32_t split(uint32_t a, uint32_t b, int depth) const {
    const uint32_t mid = a + (b - a) / 2;
    const uint32_t ca = code_[a], cb = code_[b - 1];
    if (ca == cb) return mid;
    const int bit = 31 - __builtin_clz(ca ^ cb);
    // codes are sorted and agree above `bit`: the first code with the bit set starts the
    // upper half
    uint32_t lo = a, hi = b - 1;
    while (lo < hi) {
      const uint32_t md = lo + (hi - lo) / 2;
      if (code_[md] & (1u << bit)) hi = md;
      else lo = md + 1;
    }
    const uint32_t m = lo;
    // children sit at interior depth depth + 1 (if interior); keep the whole tree within
    // kSearchDepth interior levels
    if (depth + levels_needed(m - a) > kSearchDepth || depth + levels_needed(b - m) > kSearchDepth)
      return mid;
    return m;
  }

  void write(uint32_t node, const SBox& b0, uint32_t r0, const SBox& b1, uint32_t r1) {
    float lo0[3], hi0[3], lo1[3], hi1[3];
    for (int k = 0; k < 3; ++k) {
      lo0[k] = b0.lo[k] - pad_;
      hi0[k] = b0.hi[k] + pad_;
      lo1[k] = b1.lo[k] - pad_;
      hi1[k] = b1.hi[k] + pad_;
    }
    float f0, f1;
    std::memcpy(&f0, &r0, 4);
    std::memcpy(&f1, &r1, 4);
    float4* p = &out_[4 * (size_t)node];
    p[0] = make_float4(lo0[0], lo0[1], lo0[2], hi0[0]);
    p[1] = make_float4(hi0[1], hi0[2], lo1[0], lo1[1]);
    p[2] = make_float4(lo1[2], hi1[0], hi1[1], hi1[2]);
    p[3] = make_float4(f0, f1, 0.f, 0.f);
  }

  const float* v9_;
  const uint32_t* code_;
  std::vector<float4>& out_;
  float pad_;
};

}  // namespace

void build_search_bvh(const float* v9, const uint32_t* code, uint32_t ntri,
                      std::vector<float4>& nodes) {
  nodes.clear();
  if (ntri == 0) return;
  float big = 1.f;
  for (size_t i = 0; i < 9 * (size_t)ntri; ++i) big = std::max(big, std::fabs(v9[i]));
  const float pad = std::ldexp(big, -11);
  nodes.reserve(4 * (size_t)std::max<uint32_t>(1, ntri / 2));
  SearchBuilder b(v9, code, nodes, pad);
  b.build_root(ntri);
}

}  // namespace rti
